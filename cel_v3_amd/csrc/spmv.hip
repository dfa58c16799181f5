// spmv.hip -- SpMV, face packing, generators and set-up kernels for gfx950.
#include "cel_internal.hpp"
#include "spmv_device.hpp"

namespace cel {

// ------------------------------------------------------------------ CSR / COO / JAD SpMV
// CSR (cel_fop_mv_module.f90:215-229): y(i) = y(i) + tmp, tmp the row sum
// from 0 (y zeroed first unless accumulating).  COO
// (cel_fop_mv_module.f90:380-391): y(row) = y(row) + a x for each entry in
// order; the block's entries are row-sorted (the order its CSR offsets
// assume, cel_sp_mat_module.f90:430-456), so a row's run of entries is
// added to y one by one, in entry order, with no atomics: the row sum
// starts from y(row) itself (INIT) or from 0 when y is overwritten.
template <typename VT, bool INIT>
__global__ __launch_bounds__(kBlock) void k_csr_spmv(int n, const int* __restrict__ xadj,
                                                     const int* __restrict__ cols,
                                                     const VT* __restrict__ vals,
                                                     const VT* __restrict__ x, VT* __restrict__ y,
                                                     int accumulate, int coo, int ntiles,
                                                     const int* __restrict__ skip)
{
    if (skip && *skip) return;  // uniform across the grid
    __shared__ CsrStage<VT> sm;
    CsrTileStream<VT> ts;
    const int tfirst = (int)group_tile(0, ntiles);
    if (tfirst >= 0) {  // uniform across the block
        ts.start(n, tfirst, (int)group_tile(1, ntiles), xadj, cols, vals, sm);
        ts.stage(sm);
        __syncthreads();
    }
    int step = 0;
    for (int t = tfirst; t >= 0; t = (int)group_tile(++step, ntiles)) {
        const int row = t * kTileRows + threadIdx.x;
        const int nxt = (int)group_tile(step + 1, ntiles);
        const int aft = (int)group_tile(step + 2, ntiles);
        VT init = VT(0);
        if constexpr (INIT) init = y[row < n ? row : 0];
        VT s = ts.template rows<VT>(n, t, nxt, aft, xadj, cols, vals, x, sm, row, nullptr, init);
        __syncthreads();
        if (nxt >= 0) {
            ts.stage(sm);
            __syncthreads();
        }
        if (row < n) y[row] = coo ? s : accumulate ? y[row] + s : VT(0) + s;
    }
}

// FP64 product fused with the next dot product (kFuseDot / kFuseResid,
// cel_internal.hpp): block b walks the 256-row tiles of its row chunk
// [b*chunk, b*chunk + chunk) with the same tile stream as k_csr_spmv (same
// row sums), each thread its rows in ascending order.
template <int FUSE>
__global__ __launch_bounds__(kBlock) void k_csr_spmv_fused(int n, long long chunk,
                                                           const int* __restrict__ xadj,
                                                           const int* __restrict__ cols,
                                                           const double* __restrict__ vals,
                                                           const double* __restrict__ x,
                                                           double* __restrict__ y,
                                                           const double* __restrict__ aux,
                                                           double* __restrict__ part,
                                                           const int* __restrict__ skip)
{
    if (skip && *skip) return;  // uniform across the grid
    __shared__ CsrStage<double> sm;
    __shared__ double red[kWaves];
    const long long c0 = (long long)blockIdx.x * chunk;
    const long long c1 = c0 + chunk < n ? c0 + chunk : n;
    const int t0 = (int)(c0 / kTileRows);  // chunk is a multiple of the tile
    const int t1 = c0 < c1 ? (int)((c1 + kTileRows - 1) / kTileRows) : t0;
    CsrTileStream<double> ts;
    if (t0 < t1) {  // uniform across the block
        ts.start(n, t0, t0 + 1 < t1 ? t0 + 1 : -1, xadj, cols, vals, sm);
        ts.stage(sm);
        __syncthreads();
    }
    double acc = 0.0;
    for (int t = t0; t < t1; ++t) {
        const int row = t * kTileRows + threadIdx.x;
        const int nxt = t + 1 < t1 ? t + 1 : -1;
        const int aft = t + 2 < t1 ? t + 2 : -1;
        double s = ts.template rows<double>(n, t, nxt, aft, xadj, cols, vals, x, sm, row, nullptr);
        __syncthreads();
        if (nxt >= 0) {
            ts.stage(sm);
            __syncthreads();
        }
        if (row < n) acc = acc + fuse_row<FUSE>(0.0 + s, row, aux, y);
    }
    const double bs = block_sum(acc, red);
    if (threadIdx.x == 0) part[blockIdx.x] = bs;
}

template <typename VT>
__global__ __launch_bounds__(kBlock) void k_jad_spmv(int n, int mmax, const int* __restrict__ begin,
                                                     const int* __restrict__ jcols,
                                                     const VT* __restrict__ jvals,
                                                     const int* __restrict__ perm,
                                                     const VT* __restrict__ x, VT* __restrict__ y,
                                                     int accumulate, int ntiles,
                                                     const int* __restrict__ skip)
{
    if (skip && *skip) return;
    // single tiles dealt round-robin (tile t to block t mod G): for JAD,
    // whose tiles need no staging, measured faster than tile pairs (FP64
    // 346 against 360-366 us on the 256^3 block)
    for (int t = blockIdx.x; t < ntiles; t += gridDim.x) {
        const int k = t * kTileRows + threadIdx.x;
        if (k < n) {
            VT s = jad_row<VT, VT>(k, mmax, begin, jcols, jvals, x);
            const int row = perm[k];
            // the reference overwrites y(perm(k)) = tmp(k) (cel_fop_mv_module.f90:323-326)
            y[row] = accumulate ? y[row] + s : s;
        }
    }
}

// A block with no columns has no entries: y = 0 (or unchanged).  The tile
// kernels are not launched for it -- their branch-free gathers read x[0].
template <typename T>
static int spmv_no_columns(const Mat& A, T* y, int accumulate)
{
    if (!accumulate) CEL_HIP(hipMemsetAsync(y, 0, sizeof(T) * (size_t)A.n, rt().stream));
    return CEL_OK;
}

template <typename VT>
static int launch_csr_tiles(const Mat& A, const VT* vals, const VT* x, VT* y, int accumulate,
                            int coo, const int* skip)
{
    const int ntiles = (A.n + kTileRows - 1) / kTileRows;
    if (coo && accumulate) {
        const int G = resident_grid((const void*)k_csr_spmv<VT, true>, ntiles);
        hipLaunchKernelGGL((k_csr_spmv<VT, true>), dim3(G), dim3(kBlock), 0, rt().stream, A.n,
                           A.xadj.p, A.cols.p, vals, x, y, accumulate, coo, ntiles, skip);
    } else {
        const int G = resident_grid((const void*)k_csr_spmv<VT, false>, ntiles);
        hipLaunchKernelGGL((k_csr_spmv<VT, false>), dim3(G), dim3(kBlock), 0, rt().stream, A.n,
                           A.xadj.p, A.cols.p, vals, x, y, accumulate, coo, ntiles, skip);
    }
    CEL_HIP(hipGetLastError());
    return CEL_OK;
}

int launch_csr_spmv(const Mat& A, const double* x, double* y, int accumulate, const int* skip,
                    int coo)
{
    if (A.n == 0) return CEL_OK;
    if (A.ncols == 0) return spmv_no_columns(A, y, accumulate);
    return launch_csr_tiles<double>(A, A.vals.p, x, y, accumulate, coo, skip);
}

int launch_csr_spmv_fused(const Mat& A, int fuse, long long chunk, int G, const double* x,
                          double* y, const double* aux, double* part, const int* skip)
{
    if (A.n == 0 || A.ncols == 0 || A.n != A.ncols)
        return fail(CEL_ERR_ARG, "fused product: a square block with rows");
    if (chunk <= 0 || chunk % kTileRows != 0 || (long long)G * chunk < A.n)
        return fail(CEL_ERR_ARG, "fused product: chunks must be whole tiles covering every row");
    if (fuse == kFuseDot)
        hipLaunchKernelGGL(k_csr_spmv_fused<kFuseDot>, dim3(G), dim3(kBlock), 0, rt().stream, A.n,
                           chunk, A.xadj.p, A.cols.p, A.vals.p, x, y, aux, part, skip);
    else if (fuse == kFuseResid)
        hipLaunchKernelGGL(k_csr_spmv_fused<kFuseResid>, dim3(G), dim3(kBlock), 0, rt().stream, A.n,
                           chunk, A.xadj.p, A.cols.p, A.vals.p, x, y, aux, part, skip);
    else
        return fail(CEL_ERR_ARG, "fused product: unknown epilogue");
    CEL_HIP(hipGetLastError());
    return CEL_OK;
}

int launch_csr_spmv_f32(const Mat& A, const float* x, float* y, int accumulate, int coo)
{
    if (A.n == 0) return CEL_OK;
    if (A.ncols == 0) return spmv_no_columns(A, y, accumulate);
    if (!A.vals32.p) return fail(CEL_ERR_STATE, "FP32 values not prepared");
    return launch_csr_tiles<float>(A, A.vals32.p, x, y, accumulate, coo, nullptr);
}

int launch_jad_spmv(const Mat& A, const double* x, double* y, int accumulate, const int* skip)
{
    if (A.n == 0) return CEL_OK;
    int ntiles = (A.n + kTileRows - 1) / kTileRows;
    int G = resident_grid((const void*)k_jad_spmv<double>, ntiles);
    hipLaunchKernelGGL(k_jad_spmv<double>, dim3(G), dim3(kBlock), 0, rt().stream, A.n, A.jad_mmax,
                       A.jad_begin.p, A.jad_cols.p, A.jad_vals.p, A.jad_perm.p, x, y, accumulate,
                       ntiles, skip);
    CEL_HIP(hipGetLastError());
    return CEL_OK;
}

int launch_jad_spmv_f32(const Mat& A, const float* x, float* y, int accumulate)
{
    if (A.n == 0) return CEL_OK;
    if (!A.jad_vals32.p) return fail(CEL_ERR_STATE, "FP32 JAD values not prepared");
    int ntiles = (A.n + kTileRows - 1) / kTileRows;
    int G = resident_grid((const void*)k_jad_spmv<float>, ntiles);
    hipLaunchKernelGGL(k_jad_spmv<float>, dim3(G), dim3(kBlock), 0, rt().stream, A.n, A.jad_mmax,
                       A.jad_begin.p, A.jad_cols.p, A.jad_vals32.p, A.jad_perm.p, x, y, accumulate,
                       ntiles, nullptr);
    CEL_HIP(hipGetLastError());
    return CEL_OK;
}

// ------------------------------------------------------------------ halo
// Adds the halo entries (compact rows, block order) to y.  One thread per
// compact halo row; the local sum is already in y.
__global__ void k_halo_spmv(int nrows, const int* __restrict__ hrows, const int* __restrict__ hptr,
                            const int* __restrict__ hcols, const double* __restrict__ hvals,
                            const double* __restrict__ ghost, double* __restrict__ y)
{
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= nrows) return;
    const int row = hrows[q];
    double s = y[row];
    for (int j = hptr[q]; j < hptr[q + 1]; ++j) {
        double prod = hvals[j] * ghost[hcols[j]];
        s = s + prod;
    }
    y[row] = s;
}

int launch_halo_spmv(const Mat& A, const double* ghost, double* y)
{
    if (A.halo_nrows == 0) return CEL_OK;
    int G = (A.halo_nrows + kBlock - 1) / kBlock;
    hipLaunchKernelGGL(k_halo_spmv, dim3(G), dim3(kBlock), 0, rt().stream, A.halo_nrows,
                       A.halo_rows.p, A.halo_ptr.p, A.halo_cols.p, A.halo_vals.p, ghost, y);
    CEL_HIP(hipGetLastError());
    return CEL_OK;
}

__global__ void k_gather(int len, const int* __restrict__ perm, const double* __restrict__ x,
                         double* __restrict__ buf)
{
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j < len) buf[j] = x[perm[j]];
}

int launch_face_pack(const Mat& A, const double* x, double* sendbuf)
{
    if (A.send_total == 0) return CEL_OK;
    int G = (A.send_total + kBlock - 1) / kBlock;
    hipLaunchKernelGGL(k_gather, dim3(G), dim3(kBlock), 0, rt().stream, A.send_total, A.send_perm.p,
                       x, sendbuf);
    CEL_HIP(hipGetLastError());
    return CEL_OK;
}

// ------------------------------------------------------------------ generators
// One thread per cell; row starts are closed-form, so no scan is needed.
// A cell's row holds itself plus one entry per interior face; summing the
// face indicators over all earlier cells separates by direction.
// edge(m, c) = sum over t in [0, c) of (t > 0) + (t < m - 1).
__device__ __forceinline__ long long edge_sum(long long m, long long c)
{
    long long a = c - 1 > 0 ? c - 1 : 0;
    long long b = c < m - 1 ? c : m - 1;
    if (b < 0) b = 0;
    return a + b;
}

__device__ __forceinline__ long long p7_entries_before(int nx, int ny, int nz, int x, int y, int z)
{
    const long long plane = (long long)nx * ny;
    const long long cells = (long long)z * plane + (long long)y * nx + x;
    const long long gy = (y > 0) + (y < ny - 1), gz = (z > 0) + (z < nz - 1);
    const long long xl = ((long long)z * ny + y) * edge_sum(nx, nx) + edge_sum(nx, x);
    const long long yl = (long long)z * nx * edge_sum(ny, ny) + (long long)nx * edge_sum(ny, y) +
                         (long long)x * gy;
    const long long zl = plane * edge_sum(nz, z) + ((long long)y * nx + x) * gz;
    return cells + xl + yl + zl;
}

__global__ void k_gen7(int nx, int ny, int nz, const double* __restrict__ kappa,
                       int* __restrict__ xadj, int* __restrict__ cols, double* __restrict__ vals)
{
    const long long n = (long long)nx * ny * nz;
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const long long sxy = (long long)nx * ny;
    const int z = (int)(i / sxy);
    const int y = (int)((i - (long long)z * sxy) / nx);
    const int x = (int)(i - (long long)z * sxy - (long long)y * nx);
    long long pos = p7_entries_before(nx, ny, nz, x, y, z);
    const bool has[6] = {z > 0, y > 0, x > 0, x < nx - 1, y < ny - 1, z < nz - 1};
    const long long nb[6] = {i - sxy, i - nx, i - 1, i + 1, i + nx, i + sxy};
    double kf[6];
    const double ki = kappa ? kappa[i] : 1.0;
    for (int f = 0; f < 6; ++f) {
        if (!kappa) {
            kf[f] = 1.0;
        } else if (has[f]) {
            double kj = kappa[nb[f]];
            double num = 2.0 * ki;
            num = num * kj;
            double den = ki + kj;
            kf[f] = num / den;
        } else {
            kf[f] = ki;
        }
    }
    double diag = 0.0;
    for (int f = 0; f < 6; ++f) diag = diag + kf[f];
    if (!kappa) diag = 6.0;
    xadj[i] = (int)pos;
    for (int f = 0; f < 3; ++f)
        if (has[f]) { cols[pos] = (int)nb[f]; vals[pos] = -kf[f]; ++pos; }
    cols[pos] = (int)i; vals[pos] = diag; ++pos;
    for (int f = 3; f < 6; ++f)
        if (has[f]) { cols[pos] = (int)nb[f]; vals[pos] = -kf[f]; ++pos; }
    if (i == n - 1) xadj[n] = (int)pos;
}

int launch_gen_poisson7(int nx, int ny, int nz, const double* kappa, int* xadj, int* cols,
                        double* vals)
{
    long long n = (long long)nx * ny * nz;
    if (n == 0) return CEL_OK;
    long long G = (n + kBlock - 1) / kBlock;
    hipLaunchKernelGGL(k_gen7, dim3((unsigned)G), dim3(kBlock), 0, rt().stream, nx, ny, nz, kappa,
                       xadj, cols, vals);
    CEL_HIP(hipGetLastError());
    return CEL_OK;
}

// ------------------------------------------------------------------ diagonal, Jacobi, scaling
// Diagonal of the local block: the entry whose (global) column equals the
// global row; last one wins, as in cel_sp_mat_distr_module.f90:197-207.
// Missing diagonals stay 0 and are reported by set_jacobi.
__global__ void k_extract_diag(int n, int row_offset, int col_offset, const int* __restrict__ xadj,
                               const int* __restrict__ cols, const double* __restrict__ vals,
                               double* __restrict__ diag)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double d = 0.0;
    const int grow = i + row_offset;  // 0-based global row
    for (int j = xadj[i]; j < xadj[i + 1]; ++j)
        if (cols[j] + col_offset == grow) d = vals[j];
    diag[i] = d;
}

int launch_extract_diag(Mat& A)
{
    CEL_TRY(A.diag.alloc(A.n > 0 ? A.n : 1));
    if (A.n == 0) return CEL_OK;
    int G = (A.n + kBlock - 1) / kBlock;
    hipLaunchKernelGGL(k_extract_diag, dim3(G), dim3(kBlock), 0, rt().stream, A.n, A.row_offset,
                       A.col_offset, A.xadj.p, A.cols.p, A.vals.p, A.diag.p);
    CEL_HIP(hipGetLastError());
    return CEL_OK;
}

__global__ void k_set_jacobi(int n, int row_scaled, const double* __restrict__ diag,
                             double* __restrict__ jac, float* __restrict__ jac32,
                             int* __restrict__ bad)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double d = diag[i];
    if (d == 0.0 || !isfinite(d)) atomicOr(bad, 1);
    double j;
    if (row_scaled) {
        double sd = sqrt(fabs(d));
        j = sd / d;
    } else {
        j = 1.0 / d;
    }
    jac[i] = j;
    jac32[i] = (float)j;
}

int launch_set_jacobi(Mat& A, int row_scaled)
{
    CEL_TRY(A.jac.alloc(A.n > 0 ? A.n : 1));
    CEL_TRY(A.jac32.alloc(A.n > 0 ? A.n : 1));
    if (A.n == 0) return CEL_OK;
    DevBuf<int> bad;
    CEL_TRY(bad.alloc(1));
    CEL_HIP(hipMemsetAsync(bad.p, 0, sizeof(int), rt().stream));
    int G = (A.n + kBlock - 1) / kBlock;
    hipLaunchKernelGGL(k_set_jacobi, dim3(G), dim3(kBlock), 0, rt().stream, A.n, row_scaled, A.diag.p,
                       A.jac.p, A.jac32.p, bad.p);
    CEL_HIP(hipGetLastError());
    int h = 0;
    CEL_HIP(hipMemcpyAsync(&h, bad.p, sizeof(int), hipMemcpyDeviceToHost, rt().stream));
    CEL_HIP(hipStreamSynchronize(rt().stream));
    if (h) return fail(CEL_ERR_BAD_DIAG, "zero, missing or non-finite diagonal entry");
    return CEL_OK;
}

__global__ void k_scale_rows(int n, const int* __restrict__ xadj, double* __restrict__ vals,
                             const double* __restrict__ diag, double* __restrict__ sdiag)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double s = sqrt(fabs(diag[i]));
    sdiag[i] = s;
    for (int j = xadj[i]; j < xadj[i + 1]; ++j) vals[j] = vals[j] / s;
}

__global__ void k_scale_halo(int nrows, const int* __restrict__ hrows, const int* __restrict__ hptr,
                             double* __restrict__ hvals, const double* __restrict__ sdiag)
{
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= nrows) return;
    const double s = sdiag[hrows[q]];
    for (int j = hptr[q]; j < hptr[q + 1]; ++j) hvals[j] = hvals[j] / s;
}

int launch_scale_rows(Mat& A)
{
    CEL_TRY(A.scaled_diag.alloc(A.n > 0 ? A.n : 1));
    if (A.n == 0) return CEL_OK;
    int G = (A.n + kBlock - 1) / kBlock;
    hipLaunchKernelGGL(k_scale_rows, dim3(G), dim3(kBlock), 0, rt().stream, A.n, A.xadj.p, A.vals.p,
                       A.diag.p, A.scaled_diag.p);
    CEL_HIP(hipGetLastError());
    if (A.halo_nrows > 0) {
        int Gh = (A.halo_nrows + kBlock - 1) / kBlock;
        hipLaunchKernelGGL(k_scale_halo, dim3(Gh), dim3(kBlock), 0, rt().stream, A.halo_nrows,
                           A.halo_rows.p, A.halo_ptr.p, A.halo_vals.p, A.scaled_diag.p);
        CEL_HIP(hipGetLastError());
    }
    return CEL_OK;
}

__global__ void k_scale_vec(int n, const double* __restrict__ diag, double* __restrict__ b)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) b[i] = b[i] / sqrt(fabs(diag[i]));
}

int launch_scale_vec(const Mat& A, double* b)
{
    if (A.n == 0) return CEL_OK;
    int G = (A.n + kBlock - 1) / kBlock;
    hipLaunchKernelGGL(k_scale_vec, dim3(G), dim3(kBlock), 0, rt().stream, A.n, A.diag.p, b);
    CEL_HIP(hipGetLastError());
    return CEL_OK;
}

__global__ void k_to_f32(long long n, const double* __restrict__ s, float* __restrict__ d)
{
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (long long)gridDim.x * blockDim.x)
        d[i] = (float)s[i];
}

int launch_to_f32(const double* src, float* dst, size_t n)
{
    if (n == 0) return CEL_OK;
    long long G = ((long long)n + kBlock - 1) / kBlock;
    if (G > 65536) G = 65536;
    hipLaunchKernelGGL(k_to_f32, dim3((unsigned)G), dim3(kBlock), 0, rt().stream, (long long)n, src, dst);
    CEL_HIP(hipGetLastError());
    return CEL_OK;
}

// ------------------------------------------------------------------ adjoint helpers
// g(j) = -lambda(i) * x(cols(j)) on the CSR pattern (sampled dense product).
__global__ void k_sddmm(int n, const int* __restrict__ xadj, const int* __restrict__ cols,
                        const double* __restrict__ lam, const double* __restrict__ x,
                        double* __restrict__ g)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double li = lam[i];
    for (int j = xadj[i]; j < xadj[i + 1]; ++j) {
        double t = li * x[cols[j]];
        g[j] = -t;
    }
}

int launch_sddmm(const Mat& A, const double* lambda, const double* x, double* g)
{
    if (A.n == 0) return CEL_OK;
    int G = (A.n + kBlock - 1) / kBlock;
    hipLaunchKernelGGL(k_sddmm, dim3(G), dim3(kBlock), 0, rt().stream, A.n, A.xadj.p, A.cols.p,
                       lambda, x, g);
    CEL_HIP(hipGetLastError());
    return CEL_OK;
}

__global__ void k_kappa_grad(int nx, int ny, int nz, const double* __restrict__ kappa,
                             const double* __restrict__ x, const double* __restrict__ lam,
                             double* __restrict__ grad)
{
    const long long n = (long long)nx * ny * nz;
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const long long sxy = (long long)nx * ny;
    const int z = (int)(i / sxy);
    const int y = (int)((i - (long long)z * sxy) / nx);
    const int xx = (int)(i - (long long)z * sxy - (long long)y * nx);
    const bool has[6] = {z > 0, y > 0, xx > 0, xx < nx - 1, y < ny - 1, z < nz - 1};
    const long long nb[6] = {i - sxy, i - nx, i - 1, i + 1, i + nx, i + sxy};
    const double ki = kappa[i];
    double g = 0.0;
    for (int f = 0; f < 6; ++f) {
        double term;
        if (has[f]) {
            const long long j = nb[f];
            const double kj = kappa[j];
            double dl = lam[i] - lam[j];
            double dx = x[i] - x[j];
            double dLdk = -(dl * dx);
            double s = ki + kj;
            double num = 2.0 * kj;
            num = num * kj;
            double den = s * s;
            double dk = num / den;
            term = dLdk * dk;
        } else {
            double t = lam[i] * x[i];
            term = -t;
        }
        g = g + term;
    }
    grad[i] = g;
}

int launch_kappa_grad(int nx, int ny, int nz, const double* kappa, const double* x,
                      const double* lambda, double* grad)
{
    long long n = (long long)nx * ny * nz;
    if (n == 0) return CEL_OK;
    long long G = (n + kBlock - 1) / kBlock;
    hipLaunchKernelGGL(k_kappa_grad, dim3((unsigned)G), dim3(kBlock), 0, rt().stream, nx, ny, nz, kappa,
                       x, lambda, grad);
    CEL_HIP(hipGetLastError());
    return CEL_OK;
}

double spmv_bytes(const Mat& A, int format, int precision)
{
    const double vb = precision == CEL_PREC_FP64 ? 8.0 : 4.0;
    const double n = A.n, nnz = A.nnz;
    if (format == CEL_FMT_DIA)  // stored slots, the row masks, x once, y
        return vb * A.dia_nd * n + (A.dia_nd <= 8 ? 1.0 : 4.0) * n + vb * A.ncols + vb * n;
    double idx = (format == CEL_FMT_JAD) ? 4.0 * (A.jad_mmax + 1) + 4.0 * n : 4.0 * (n + 1);
    return nnz * (4.0 + vb) + idx + vb * A.ncols + vb * n;
}

}  // namespace cel
