// pcg.hip -- Jacobi-PCG for gfx950: FP64 (reference op order) and the
// mixed form (FP32 inner iteration, FP64 residual replacement).
//
// Reference loop (src/fcgalg/cel_cgalg_module.f90:206-304):
//   r = b - A x ; p = jac*r ; alpha = r.p ; norm2 = ||r||
//   while k < iter_max and |alpha| > 1.1 eps_mach and norm2 > eps:
//     v = A p ; k++ ; sigma = alpha / (v.p)
//     x = x + sigma p ; r = r + (-sigma) v ; z = jac*r
//     beta = (r.z)/alpha ; alpha = r.z ; p = beta p + z ; norm2 = ||r||
//
// One iteration is three kernels, each ending in a deterministic grid
// reduction whose last block updates the scalars on the device, so the host
// never waits inside the loop:
//   K1  v = A p (CSR tile from LDS or JAD), partial v.p   -> sigma
//   K2  r = r + (-sigma) v, z = jac*r, partials r.z, r.r -> beta, alpha,
//       norm2, k, stop decision
//   K3  x = x + sigma p ; p = beta p + jac*r   (x update moved here: it
//       reads p anyway, which saves one pass over p; x never feeds back
//       into the iteration, and each element is computed by the same
//       formula as the reference)
// Per element the arithmetic is the reference's (products rounded before
// the add, compiled with -ffp-contract=off); only the reduction order of
// the dot products differs.
//
// Kernels of iteration k check device flags instead of k, so one captured
// chunk of iterations can be replayed as a hipGraph and runs past
// convergence as no-ops.
#include <algorithm>
#include <cmath>
#include <vector>

#include "cel_internal.hpp"
#include "spmv_device.hpp"

namespace cel {

struct PcgState {
    double alpha, sigma, beta, norm2, eps, norm0, bnorm, tol, small, rnorm_ref, ru_delta;
    double vp;
    int iter, iter_max, stop, status, pending_xp, ru_flag, outer, tol_mode, ru_restart, hist_cap;
    double* hist;
    unsigned ctr[4];
};

// ------------------------------------------------------------------ helpers
__device__ __forceinline__ bool loop_continues(const PcgState* s)
{
    return s->iter < s->iter_max && fabs(s->alpha) > s->small && s->norm2 > s->eps;
}

__device__ __forceinline__ double eps_from(const PcgState* s, double r0)
{
    if (s->tol_mode == CEL_TOL_REL_B) return s->tol * s->bnorm;
    if (s->tol_mode == CEL_TOL_REL_R0) return s->tol * r0;
    return s->tol;
}

// ------------------------------------------------------------------ set-up
__global__ void k_fill(long long n, double* __restrict__ x, double v)
{
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (long long)gridDim.x * blockDim.x)
        x[i] = v;
}

__global__ void k_fill_f32(long long n, float* __restrict__ x, float v)
{
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (long long)gridDim.x * blockDim.x)
        x[i] = v;
}

struct MatView {
    int n, mmax;
    const int *xadj, *cols;
    const double* vals;
    const float* vals32;
    const int *jbegin, *jcols, *jperm;
    const double* jvals;
    const float* jvals32;
    // halo (FP64 only)
    const int *hrow_tile, *hrows, *hptr, *hcols;
    const double* hvals;
    const double* ghost;
};

template <typename VT>
__device__ __forceinline__ const VT* csr_vals(const MatView& A);
template <>
__device__ __forceinline__ const double* csr_vals<double>(const MatView& A) { return A.vals; }
template <>
__device__ __forceinline__ const float* csr_vals<float>(const MatView& A) { return A.vals32; }
template <typename VT>
__device__ __forceinline__ const VT* jad_vals(const MatView& A);
template <>
__device__ __forceinline__ const double* jad_vals<double>(const MatView& A) { return A.jvals; }
template <>
__device__ __forceinline__ const float* jad_vals<float>(const MatView& A) { return A.jvals32; }

// Row sums of one 256-row tile in the requested format.  Returns the sum of
// row `row` (set here).  For CSR row = r0 + tid; for JAD row = perm[r0+tid].
// HALO adds the halo entries after the local sum ((0 + local) + halo...).
template <typename VT, typename AT, int FMT, bool HALO>
__device__ __forceinline__ AT tile_rows(const MatView& A, int t, const VT* __restrict__ x,
                                        CsrStage<VT>& sm, int* smap, int& row)
{
    const int r0 = t * kTileRows;
    AT s;
    if constexpr (FMT == CEL_FMT_CSR) {
        s = csr_tile<VT, AT>(A.n, r0, A.xadj, A.cols, csr_vals<VT>(A), x, sm);
        row = r0 + threadIdx.x;
        s = AT(0) + s;
        if constexpr (HALO)
            s = halo_tile<AT>(r0, t, s, A.hrow_tile, A.hrows, A.hptr, A.hcols, A.hvals, A.ghost, smap);
    } else {
        const int k = r0 + threadIdx.x;
        s = AT(0);
        row = A.n;
        if (k < A.n) {
            s = jad_row<VT, AT>(k, A.mmax, A.jbegin, A.jcols, jad_vals<VT>(A), x);
            row = A.jperm[k];
        }
    }
    return s;
}

// Every reduction kernel below walks the chunks it owns; a chunk is
// chunk_tiles tiles of 256 rows (tile kernels) or the same rows as plain
// elements (vector kernels), and writes one partial per chunk.

// r = b - A x (FP64 operator), p = jac*r, partials: r.p (or r32.p32), r.r, b.b.
// OUT_T = double (FP64 form: r, p stored in FP64) or float (mixed form:
// r rounded to FP32, p = jac32 * r32, the FP64 norm taken before rounding).
template <typename OUT_T, int FMT, bool HALO>
__global__ __launch_bounds__(kBlock) void k_residual(MatView A, const double* __restrict__ x,
                                                     const double* __restrict__ b,
                                                     const double* __restrict__ jac,
                                                     const float* __restrict__ jac32,
                                                     OUT_T* __restrict__ r, OUT_T* __restrict__ p,
                                                     float* __restrict__ e, double* __restrict__ part,
                                                     PcgState* __restrict__ st, int nchunks,
                                                     int chunk_tiles)
{
    __shared__ CsrStage<double> sm;
    __shared__ int smap[kBlock];
    __shared__ double red[3 * kWaves];
    __shared__ int bad;
    if (threadIdx.x == 0) bad = 0;
    __syncthreads();
    const int lb = logical_block();
    const int ntiles = (A.n + kTileRows - 1) / kTileRows;
    int c0, c1;
    chunk_range(nchunks, gridDim.x, lb, c0, c1);
    for (int c = c0; c < c1; ++c) {
        double arp = 0.0, arr = 0.0, abb = 0.0;
        const int t1 = min((c + 1) * chunk_tiles, ntiles);
        for (int t = c * chunk_tiles; t < t1; ++t) {
            int row;
            double s = tile_rows<double, double, FMT, HALO>(A, t, x, sm, smap, row);
            if (row < A.n) {
                const double bi = b[row];
                if (!isfinite(bi)) bad = 1;
                const double ri = bi - s;  // r = 0; r = r + A x; r = b - r
                double bb = bi * bi;
                abb = abb + bb;
                double rr = ri * ri;
                arr = arr + rr;
                if constexpr (sizeof(OUT_T) == 8) {
                    const double pi = ri * jac[row];
                    r[row] = ri;
                    p[row] = pi;
                    double rp = ri * pi;
                    arp = arp + rp;
                } else {
                    const float rf = (float)ri;
                    const float pf = rf * jac32[row];
                    r[row] = rf;
                    p[row] = pf;
                    e[row] = 0.0f;
                    double rp = (double)rf * (double)pf;
                    arp = arp + rp;
                }
            }
        }
        block_sum2(arp, arr, red);
        const double sbb = block_sum(abb, red + 2 * kWaves);
        if (threadIdx.x == 0) {
            part[c] = arp;
            part[nchunks + c] = arr;
            part[2 * nchunks + c] = sbb;
        }
    }
    __syncthreads();
    if (threadIdx.x == 0 && bad) st->status = CEL_ERR_NONFINITE;
    if (!last_block(&st->ctr[0])) return;
    const double rp = sum_partials(part, nchunks, red);
    const double rr = sum_partials(part + nchunks, nchunks, red);
    const double bb = sum_partials(part + 2 * nchunks, nchunks, red);
    if (threadIdx.x == 0) {
        const double rn = sqrt(rr);
        st->bnorm = sqrt(bb);
        st->alpha = rp;
        st->norm2 = rn;
        st->norm0 = rn;
        st->eps = eps_from(st, rn);
        st->rnorm_ref = rn;
        st->iter = 0;
        st->outer = 0;
        st->pending_xp = 0;
        st->ru_flag = 0;
        if (st->hist_cap > 0) st->hist[0] = rn;
        st->stop = (st->status != 0 || !loop_continues(st)) ? 1 : 0;
    }
}

// K1: v = A p, partial v.p ; last block: sigma = alpha / (v.p)
template <typename VT, int FMT, bool HALO>
__global__ __launch_bounds__(kBlock) void k_pcg_spmv(MatView A, const VT* __restrict__ p,
                                                     VT* __restrict__ v, double* __restrict__ part,
                                                     PcgState* __restrict__ st, int nchunks,
                                                     int chunk_tiles)
{
    if (st->stop) return;
    __shared__ CsrStage<VT> sm;
    __shared__ int smap[kBlock];
    __shared__ double red[kWaves];
    const int lb = logical_block();
    const int ntiles = (A.n + kTileRows - 1) / kTileRows;
    int c0, c1;
    chunk_range(nchunks, gridDim.x, lb, c0, c1);
    for (int c = c0; c < c1; ++c) {
        double acc = 0.0;
        const int t1 = min((c + 1) * chunk_tiles, ntiles);
        for (int t = c * chunk_tiles; t < t1; ++t) {
            int row;
            VT s = tile_rows<VT, VT, FMT, HALO>(A, t, p, sm, smap, row);
            if (row < A.n) {
                v[row] = s;
                double prod = double(s) * double(p[row]);
                acc = acc + prod;
            }
        }
        const double bs = block_sum(acc, red);
        if (threadIdx.x == 0) part[c] = bs;
    }
    if (!last_block(&st->ctr[1])) return;
    const double vp = sum_partials(part, nchunks, red);
    if (threadIdx.x == 0) {
        const double sigma = st->alpha / vp;
        st->vp = vp;
        st->sigma = sigma;
        if (!(vp > 0.0) || !isfinite(sigma)) {
            st->status = CEL_ERR_BREAKDOWN;
            st->stop = 1;
        }
    }
}

// K2: r = r + (-sigma) v ; z = jac*r ; partials r.z, r.r ; last block:
// beta = rz/alpha, alpha = rz, norm2 = sqrt(rr), k++, stop / replacement flag.
template <typename VT, bool MIXED>
__global__ __launch_bounds__(kBlock) void k_pcg_update_r(long long n, VT* __restrict__ r,
                                                         const VT* __restrict__ v,
                                                         const VT* __restrict__ jac,
                                                         double* __restrict__ part,
                                                         PcgState* __restrict__ st, int nchunks,
                                                         int chunk_rows)
{
    if (st->stop) return;
    __shared__ double red[2 * kWaves];
    const int lb = logical_block();
    const VT ms = VT(-st->sigma);
    int c0, c1;
    chunk_range(nchunks, gridDim.x, lb, c0, c1);
    for (int c = c0; c < c1; ++c) {
        double arz = 0.0, arr = 0.0;
        const long long hi = min((long long)(c + 1) * chunk_rows, n);
        for (long long i = (long long)c * chunk_rows + threadIdx.x; i < hi; i += kBlock) {
            VT t = ms * v[i];
            VT ri = r[i] + t;
            r[i] = ri;
            VT zi = ri * jac[i];
            double rz = double(ri) * double(zi);
            arz = arz + rz;
            double rr = double(ri) * double(ri);
            arr = arr + rr;
        }
        block_sum2(arz, arr, red);
        if (threadIdx.x == 0) {
            part[c] = arz;
            part[nchunks + c] = arr;
        }
    }
    if (!last_block(&st->ctr[2])) return;
    const double rz = sum_partials(part, nchunks, red);
    const double rr = sum_partials(part + nchunks, nchunks, red);
    if (threadIdx.x == 0) {
        const double beta = rz / st->alpha;
        st->beta = beta;
        st->alpha = rz;
        st->norm2 = sqrt(rr);
        st->iter = st->iter + 1;
        if (st->iter < st->hist_cap) st->hist[st->iter] = st->norm2;
        st->pending_xp = 1;
        if (!MIXED) {
            if (!loop_continues(st)) st->stop = 1;
        } else {
            // replace the FP32 residual by the FP64 one when it has dropped
            // by ru_delta since the last replacement, looks converged, or the
            // iteration budget is spent; the stop decision uses the FP64 norm
            const bool drop = st->norm2 < st->ru_delta * st->rnorm_ref;
            if (drop || st->norm2 <= st->eps || st->iter >= st->iter_max ||
                !(fabs(st->alpha) > st->small) || !isfinite(st->norm2))
                st->ru_flag = 1;
        }
    }
}

// K3: x = x + sigma p ; p = beta p + jac*r ; last block clears pending_xp.
template <typename VT, typename XT>
__global__ __launch_bounds__(kBlock) void k_pcg_update_xp(long long n, XT* __restrict__ x,
                                                          VT* __restrict__ p,
                                                          const VT* __restrict__ r,
                                                          const VT* __restrict__ jac,
                                                          PcgState* __restrict__ st, int nchunks,
                                                          int chunk_rows)
{
    if (!st->pending_xp) return;
    const int lb = logical_block();
    const VT sigma = VT(st->sigma), beta = VT(st->beta);
    int c0, c1;
    chunk_range(nchunks, gridDim.x, lb, c0, c1);
    const long long lo = (long long)c0 * chunk_rows;
    const long long hi = min((long long)c1 * chunk_rows, n);
    for (long long i = lo + threadIdx.x; i < hi; i += kBlock) {
        const VT pi = p[i];
        XT t = XT(sigma * pi);
        x[i] = x[i] + t;
        VT zi = r[i] * jac[i];
        VT bp = beta * pi;
        p[i] = bp + zi;
    }
    if (last_block(&st->ctr[3]) && threadIdx.x == 0) st->pending_xp = 0;
}

// Mixed form, replacement step 1: x64 = x64 + e32 ; e32 = 0.
__global__ __launch_bounds__(kBlock) void k_ru_flush(long long n, double* __restrict__ x,
                                                     float* __restrict__ e,
                                                     const PcgState* __restrict__ st, int force,
                                                     int nchunks, int chunk_rows)
{
    if (!st->ru_flag && !force) return;
    const int lb = logical_block();
    int c0, c1;
    chunk_range(nchunks, gridDim.x, lb, c0, c1);
    const long long lo = (long long)c0 * chunk_rows;
    const long long hi = min((long long)c1 * chunk_rows, n);
    for (long long i = lo + threadIdx.x; i < hi; i += kBlock) {
        double t = (double)e[i];
        x[i] = x[i] + t;
        e[i] = 0.0f;
    }
}

// Mixed form, replacement step 2: r32 = float(b - A64 x64), z = jac32 r32,
// FP64 norm of the FP64 residual; last block decides convergence on it.
template <int FMT>
__global__ __launch_bounds__(kBlock) void k_ru_residual(MatView A, const double* __restrict__ x,
                                                        const double* __restrict__ b,
                                                        const float* __restrict__ jac32,
                                                        float* __restrict__ r, float* __restrict__ p,
                                                        double* __restrict__ part,
                                                        PcgState* __restrict__ st, int nchunks,
                                                        int chunk_tiles)
{
    if (!st->ru_flag) return;
    __shared__ CsrStage<double> sm;
    __shared__ int smap[kBlock];
    __shared__ double red[2 * kWaves];
    const int lb = logical_block();
    const int restart = st->ru_restart;
    const int ntiles = (A.n + kTileRows - 1) / kTileRows;
    int c0, c1;
    chunk_range(nchunks, gridDim.x, lb, c0, c1);
    for (int c = c0; c < c1; ++c) {
        double arz = 0.0, arr = 0.0;
        const int t1 = min((c + 1) * chunk_tiles, ntiles);
        for (int t = c * chunk_tiles; t < t1; ++t) {
            int row;
            double s = tile_rows<double, double, FMT, false>(A, t, x, sm, smap, row);
            if (row < A.n) {
                const double ri = b[row] - s;
                double rr = ri * ri;
                arr = arr + rr;
                const float rf = (float)ri;
                const float zf = rf * jac32[row];
                r[row] = rf;
                if (restart) p[row] = zf;
                double rz = (double)rf * (double)zf;
                arz = arz + rz;
            }
        }
        block_sum2(arz, arr, red);
        if (threadIdx.x == 0) {
            part[c] = arz;
            part[nchunks + c] = arr;
        }
    }
    if (!last_block(&st->ctr[0])) return;
    const double rz = sum_partials(part, nchunks, red);
    const double rr = sum_partials(part + nchunks, nchunks, red);
    if (threadIdx.x == 0) {
        const double s = sqrt(rr);
        st->alpha = rz;
        st->norm2 = s;
        st->rnorm_ref = s;
        st->outer = st->outer + 1;
        st->ru_flag = 0;
        if (!(s > st->eps) || st->iter >= st->iter_max || !(fabs(rz) > st->small) ||
            !isfinite(s))
            st->stop = 1;
    }
}

// ------------------------------------------------------------------ host side
struct Workspace {
    DevBuf<double> r, p, v, x, b, part, ghost, sendbuf, hist;
    DevBuf<float> r32, p32, v32, e32;
    DevBuf<PcgState> st;
    DevBuf<int> hrow_tile;
};

static Workspace& ws_of(Mat& A)
{
    if (!A.ws) A.ws = new Workspace();
    return *A.ws;
}



// Compact halo rows per 256-row tile: hrow_tile[t] = first compact row >= t*256.
static int build_hrow_tile(const Mat& A, Workspace& W, int ntiles)
{
    std::vector<int> hrows(A.halo_nrows);
    if (A.halo_nrows)
        CEL_HIP(hipMemcpy(hrows.data(), A.halo_rows.p, sizeof(int) * A.halo_nrows,
                          hipMemcpyDeviceToHost));
    std::vector<int> ht(ntiles + 1, 0);
    int q = 0;
    for (int t = 0; t <= ntiles; ++t) {
        while (q < A.halo_nrows && hrows[q] < t * kTileRows) ++q;
        ht[t] = q;
    }
    CEL_TRY(W.hrow_tile.alloc(ntiles + 1));
    CEL_HIP(hipMemcpy(W.hrow_tile.p, ht.data(), sizeof(int) * (ntiles + 1), hipMemcpyHostToDevice));
    return CEL_OK;
}

struct Launch {
    int ntiles, Gt, Gv, Gt_res, Gru;
    int chunk_rows, chunk_tiles, nchunks;
};

int pcg_solve(Mat& A, const double* b, double* x, const cel_pcg_params& prm, cel_pcg_stats& stats,
              double* hist)
{
    CEL_TRY(ensure_init());
    hipStream_t s = rt().stream;
    const int n = A.n;
    const bool mixed = prm.precision == CEL_PREC_MIXED;
    const int fmt = prm.format;
    const bool halo = prm.use_halo != 0;
    if (fmt != CEL_FMT_CSR && fmt != CEL_FMT_JAD) return fail(CEL_ERR_ARG, "format must be CSR or JAD");
    if (fmt == CEL_FMT_JAD && (A.jad_mode < 0 || A.jad_ndef != A.jad_mmax))
        return fail(CEL_ERR_STATE, "JAD requested but the matrix has no complete JAD arrays");
    if (halo && (mixed || fmt != CEL_FMT_CSR))
        return fail(CEL_ERR_ARG, "halo blocks are applied by the FP64 CSR form only");
    if (!A.has_jac) return fail(CEL_ERR_STATE, "cel_gpu_mat_set_jacobi has not been called");
    if (A.ncols != n) return fail(CEL_ERR_ARG, "PCG needs a square local block");
    if (prm.iter_max < 0) return fail(CEL_ERR_ARG, "iter_max < 0");

    Workspace& W = ws_of(A);
    Launch L;
    L.ntiles = (n + kTileRows - 1) / kTileRows;
    if (L.ntiles == 0) L.ntiles = 1;
    {
        const void* k1;
        const void* kres;
        if (!mixed) {
            k1 = halo ? (const void*)k_pcg_spmv<double, CEL_FMT_CSR, true>
                      : fmt == CEL_FMT_CSR ? (const void*)k_pcg_spmv<double, CEL_FMT_CSR, false>
                                           : (const void*)k_pcg_spmv<double, CEL_FMT_JAD, false>;
            kres = halo ? (const void*)k_residual<double, CEL_FMT_CSR, true>
                        : fmt == CEL_FMT_CSR ? (const void*)k_residual<double, CEL_FMT_CSR, false>
                                             : (const void*)k_residual<double, CEL_FMT_JAD, false>;
        } else {
            k1 = fmt == CEL_FMT_CSR ? (const void*)k_pcg_spmv<float, CEL_FMT_CSR, false>
                                    : (const void*)k_pcg_spmv<float, CEL_FMT_JAD, false>;
            kres = fmt == CEL_FMT_CSR ? (const void*)k_residual<float, CEL_FMT_CSR, false>
                                      : (const void*)k_residual<float, CEL_FMT_JAD, false>;
        }
        const void* kru = fmt == CEL_FMT_CSR ? (const void*)k_ru_residual<CEL_FMT_CSR>
                                             : (const void*)k_ru_residual<CEL_FMT_JAD>;
        L.chunk_rows = chunk_rows_for(n);
        L.chunk_tiles = L.chunk_rows / kTileRows;
        L.nchunks = (int)(((long long)n + L.chunk_rows - 1) / L.chunk_rows);
        if (L.nchunks < 1) L.nchunks = 1;
        L.Gt = resident_grid(k1, L.nchunks);
        L.Gt_res = resident_grid(kres, L.nchunks);
        L.Gru = resident_grid(kru, L.nchunks);
        L.Gv = resident_grid((const void*)k_pcg_update_r<double, false>, L.nchunks);
    }
    const int Gmax = L.nchunks;

    // matrix data for the mixed inner operator (once per matrix)
    if (mixed) {
        if (fmt == CEL_FMT_CSR && !A.vals32.p) {
            CEL_TRY(A.vals32.alloc(A.nnz > 0 ? A.nnz : 1));
            CEL_TRY(launch_to_f32(A.vals.p, A.vals32.p, A.nnz));
        }
        if (fmt == CEL_FMT_JAD && !A.jad_vals32.p) {
            CEL_TRY(A.jad_vals32.alloc(A.jad_nnz > 0 ? A.jad_nnz : 1));
            CEL_TRY(launch_to_f32(A.jad_vals.p, A.jad_vals32.p, A.jad_nnz));
        }
    }

    // workspace
    const size_t nn = n > 0 ? n : 1;
    CEL_TRY(W.part.alloc(3 * (size_t)Gmax));
    CEL_TRY(W.st.alloc(1));
    if (mixed) {
        CEL_TRY(W.r32.alloc(nn));
        CEL_TRY(W.p32.alloc(nn));
        CEL_TRY(W.v32.alloc(nn));
        CEL_TRY(W.e32.alloc(nn));
    } else {
        CEL_TRY(W.r.alloc(nn));
        CEL_TRY(W.p.alloc(nn));
        CEL_TRY(W.v.alloc(nn));
    }
    const double* db = b;
    double* dx = x;
    if (prm.memspace == CEL_MEM_HOST) {
        CEL_TRY(W.b.alloc(nn));
        CEL_TRY(W.x.alloc(nn));
        CEL_HIP(hipMemcpyAsync(W.b.p, b, sizeof(double) * n, hipMemcpyHostToDevice, s));
        if (prm.x0_mode == CEL_X0_USER)
            CEL_HIP(hipMemcpyAsync(W.x.p, x, sizeof(double) * n, hipMemcpyHostToDevice, s));
        db = W.b.p;
        dx = W.x.p;
    }
    const int hist_cap = hist ? prm.iter_max + 1 : 0;
    if (hist_cap) CEL_TRY(W.hist.alloc(hist_cap));
    if (halo) {
        CEL_TRY(W.ghost.alloc(A.ghost_len > 0 ? A.ghost_len : 1));
        CEL_TRY(W.sendbuf.alloc(A.send_total > 0 ? A.send_total : 1));
        hipLaunchKernelGGL(k_fill, dim3(64), dim3(kBlock), 0, s, (long long)W.ghost.n, W.ghost.p, 0.0);
        CEL_TRY(build_hrow_tile(A, W, L.ntiles));
    }

    PcgState h{};
    h.tol = prm.tol;
    h.tol_mode = prm.tol_mode;
    h.iter_max = prm.iter_max;
    h.small = 1.1 * 2.220446049250313e-16;  // cel_real_small (cel_types_module.f90:33)
    h.ru_delta = prm.ru_delta > 0 ? prm.ru_delta : 0.1;
    h.ru_restart = prm.ru_restart;
    h.hist = hist_cap ? W.hist.p : nullptr;
    h.hist_cap = hist_cap;
    CEL_HIP(hipMemcpyAsync(W.st.p, &h, sizeof(PcgState), hipMemcpyHostToDevice, s));

    MatView V{};
    V.n = n;
    V.mmax = A.jad_mmax;
    V.xadj = A.xadj.p;
    V.cols = A.cols.p;
    V.vals = A.vals.p;
    V.vals32 = A.vals32.p;
    V.jbegin = A.jad_begin.p;
    V.jcols = A.jad_cols.p;
    V.jperm = A.jad_perm.p;
    V.jvals = A.jad_vals.p;
    V.jvals32 = A.jad_vals32.p;
    if (halo) {
        V.hrow_tile = W.hrow_tile.p;
        V.hrows = A.halo_rows.p;
        V.hptr = A.halo_ptr.p;
        V.hcols = A.halo_cols.p;
        V.hvals = A.halo_vals.p;
        V.ghost = W.ghost.p;
    }

    hipEvent_t e0, e1, e2;
    CEL_HIP(hipEventCreate(&e0));
    CEL_HIP(hipEventCreate(&e1));
    CEL_HIP(hipEventCreate(&e2));
    CEL_HIP(hipEventRecord(e0, s));

    // x0 (cel_cgalg_module.f90:209 uses 1.0)
    if (prm.x0_mode != CEL_X0_USER)
        hipLaunchKernelGGL(k_fill, dim3(std::max(1, std::min(L.Gv, 4096))), dim3(kBlock), 0, s,
                           (long long)n, dx, prm.x0_mode == CEL_X0_ONE ? 1.0 : 0.0);
    if (halo) CEL_TRY(launch_face_pack(A, dx, W.sendbuf.p));

    // r0, p0, alpha, norm2
#define RES_ARGS(OT, R, P)                                                                   \
    V, dx, db, A.jac.p, A.jac32.p, (OT*)(R), (OT*)(P), W.e32.p, W.part.p, W.st.p, L.nchunks, \
        L.chunk_tiles
    if (!mixed) {
        if (fmt == CEL_FMT_CSR && halo)
            hipLaunchKernelGGL((k_residual<double, CEL_FMT_CSR, true>), dim3(L.Gt_res), dim3(kBlock), 0, s,
                               RES_ARGS(double, W.r.p, W.p.p));
        else if (fmt == CEL_FMT_CSR)
            hipLaunchKernelGGL((k_residual<double, CEL_FMT_CSR, false>), dim3(L.Gt_res), dim3(kBlock), 0,
                               s, RES_ARGS(double, W.r.p, W.p.p));
        else
            hipLaunchKernelGGL((k_residual<double, CEL_FMT_JAD, false>), dim3(L.Gt_res), dim3(kBlock), 0,
                               s, RES_ARGS(double, W.r.p, W.p.p));
    } else {
        if (fmt == CEL_FMT_CSR)
            hipLaunchKernelGGL((k_residual<float, CEL_FMT_CSR, false>), dim3(L.Gt_res), dim3(kBlock), 0, s,
                               RES_ARGS(float, W.r32.p, W.p32.p));
        else
            hipLaunchKernelGGL((k_residual<float, CEL_FMT_JAD, false>), dim3(L.Gt_res), dim3(kBlock), 0, s,
                               RES_ARGS(float, W.r32.p, W.p32.p));
    }
#undef RES_ARGS
    CEL_HIP(hipGetLastError());
    CEL_HIP(hipEventRecord(e1, s));

    // profiling events around K1
    std::vector<hipEvent_t> pev;
    int pev_used = 0;
    auto prof_event = [&]() -> hipEvent_t {
        if (pev_used == (int)pev.size()) {
            hipEvent_t ev;
            if (hipEventCreate(&ev) != hipSuccess) return nullptr;
            pev.push_back(ev);
        }
        return pev[pev_used++];
    };

    auto enqueue_iter = [&](bool profile) -> int {
        hipEvent_t ea = nullptr, eb = nullptr;
        if (profile) {
            ea = prof_event();
            eb = prof_event();
            if (!ea || !eb) return fail(CEL_ERR_HIP, "hipEventCreate failed");
            CEL_HIP(hipEventRecord(ea, s));
        }
        if (!mixed) {
            if (halo) {
                CEL_TRY(launch_face_pack(A, W.p.p, W.sendbuf.p));
                hipLaunchKernelGGL((k_pcg_spmv<double, CEL_FMT_CSR, true>), dim3(L.Gt), dim3(kBlock), 0, s,
                                   V, W.p.p, W.v.p, W.part.p, W.st.p, L.nchunks, L.chunk_tiles);
            } else if (fmt == CEL_FMT_CSR) {
                hipLaunchKernelGGL((k_pcg_spmv<double, CEL_FMT_CSR, false>), dim3(L.Gt), dim3(kBlock), 0, s,
                                   V, W.p.p, W.v.p, W.part.p, W.st.p, L.nchunks, L.chunk_tiles);
            } else {
                hipLaunchKernelGGL((k_pcg_spmv<double, CEL_FMT_JAD, false>), dim3(L.Gt), dim3(kBlock), 0, s,
                                   V, W.p.p, W.v.p, W.part.p, W.st.p, L.nchunks, L.chunk_tiles);
            }
            if (profile) CEL_HIP(hipEventRecord(eb, s));
            hipLaunchKernelGGL((k_pcg_update_r<double, false>), dim3(L.Gv), dim3(kBlock), 0, s, (long long)n,
                               W.r.p, W.v.p, A.jac.p, W.part.p, W.st.p, L.nchunks, L.chunk_rows);
            hipLaunchKernelGGL((k_pcg_update_xp<double, double>), dim3(L.Gv), dim3(kBlock), 0, s,
                               (long long)n, dx, W.p.p, W.r.p, A.jac.p, W.st.p, L.nchunks, L.chunk_rows);
        } else {
            if (fmt == CEL_FMT_CSR)
                hipLaunchKernelGGL((k_pcg_spmv<float, CEL_FMT_CSR, false>), dim3(L.Gt), dim3(kBlock), 0, s,
                                   V, W.p32.p, W.v32.p, W.part.p, W.st.p, L.nchunks, L.chunk_tiles);
            else
                hipLaunchKernelGGL((k_pcg_spmv<float, CEL_FMT_JAD, false>), dim3(L.Gt), dim3(kBlock), 0, s,
                                   V, W.p32.p, W.v32.p, W.part.p, W.st.p, L.nchunks, L.chunk_tiles);
            if (profile) CEL_HIP(hipEventRecord(eb, s));
            hipLaunchKernelGGL((k_pcg_update_r<float, true>), dim3(L.Gv), dim3(kBlock), 0, s, (long long)n,
                               W.r32.p, W.v32.p, A.jac32.p, W.part.p, W.st.p, L.nchunks, L.chunk_rows);
            hipLaunchKernelGGL((k_pcg_update_xp<float, float>), dim3(L.Gv), dim3(kBlock), 0, s,
                               (long long)n, W.e32.p, W.p32.p, W.r32.p, A.jac32.p, W.st.p, L.nchunks,
                               L.chunk_rows);
            hipLaunchKernelGGL(k_ru_flush, dim3(L.Gv), dim3(kBlock), 0, s, (long long)n, dx, W.e32.p,
                               W.st.p, 0, L.nchunks, L.chunk_rows);
            if (fmt == CEL_FMT_CSR)
                hipLaunchKernelGGL(k_ru_residual<CEL_FMT_CSR>, dim3(L.Gru), dim3(kBlock), 0, s, V, dx,
                                   db, A.jac32.p, W.r32.p, W.p32.p, W.part.p, W.st.p, L.nchunks,
                                   L.chunk_tiles);
            else
                hipLaunchKernelGGL(k_ru_residual<CEL_FMT_JAD>, dim3(L.Gru), dim3(kBlock), 0, s, V, dx,
                                   db, A.jac32.p, W.r32.p, W.p32.p, W.part.p, W.st.p, L.nchunks,
                                   L.chunk_tiles);
        }
        CEL_HIP(hipGetLastError());
        return CEL_OK;
    };

    // iteration loop: chunks of `chunk` iterations, one chunk in flight
    // ahead of the host's check of the stop flag
    const int chunk = prm.check_every > 0 ? prm.check_every : 16;
    int* hstop = nullptr;
    CEL_HIP(hipHostMalloc(reinterpret_cast<void**>(&hstop), 2 * sizeof(int), hipHostMallocDefault));
    hstop[0] = hstop[1] = 0;
    hipEvent_t cev[2];
    CEL_HIP(hipEventCreateWithFlags(&cev[0], hipEventDisableTiming));
    CEL_HIP(hipEventCreateWithFlags(&cev[1], hipEventDisableTiming));

    hipGraphExec_t gexec = nullptr;
    hipGraph_t graph = nullptr;
    const bool use_graph = prm.use_graph && !prm.profile;
    int status = CEL_OK;
    long long enqueued = 0;
    int k = 0;
    // the mixed form may need iterations beyond iter_max only for the final
    // replacement, which runs inside the same chunk
    const long long budget = prm.iter_max;
    while (enqueued < budget) {
        const int c = (int)std::min<long long>(chunk, budget - enqueued);
        if (use_graph && c == chunk) {
            if (!gexec) {
                CEL_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
                for (int i = 0; i < c; ++i) {
                    int rc = enqueue_iter(false);
                    if (rc < 0) {
                        (void)hipStreamEndCapture(s, &graph);
                        return rc;
                    }
                }
                CEL_HIP(hipStreamEndCapture(s, &graph));
                CEL_HIP(hipGraphInstantiate(&gexec, graph, nullptr, nullptr, 0));
            }
            CEL_HIP(hipGraphLaunch(gexec, s));
        } else {
            for (int i = 0; i < c; ++i) {
                status = enqueue_iter(prm.profile != 0);
                if (status < 0) break;
            }
            if (status < 0) break;
        }
        enqueued += c;
        CEL_HIP(hipMemcpyAsync(&hstop[k & 1], &W.st.p->stop, sizeof(int), hipMemcpyDeviceToHost, s));
        CEL_HIP(hipEventRecord(cev[k & 1], s));
        if (k > 0) {
            CEL_HIP(hipEventSynchronize(cev[(k - 1) & 1]));
            if (hstop[(k - 1) & 1]) break;
        }
        ++k;
    }
    if (mixed) {
        // a budget that ran out mid-way leaves a correction in e: fold it in
        hipLaunchKernelGGL(k_ru_flush, dim3(L.Gv), dim3(kBlock), 0, s, (long long)n, dx, W.e32.p, W.st.p, 1,
                           L.nchunks, L.chunk_rows);
    }
    CEL_HIP(hipEventRecord(e2, s));
    CEL_HIP(hipStreamSynchronize(s));
    if (gexec) (void)hipGraphExecDestroy(gexec);
    if (graph) (void)hipGraphDestroy(graph);
    (void)hipEventDestroy(cev[0]);
    (void)hipEventDestroy(cev[1]);
    (void)hipHostFree(hstop);
    if (status < 0) return status;

    PcgState out;
    CEL_HIP(hipMemcpy(&out, W.st.p, sizeof(PcgState), hipMemcpyDeviceToHost));
    float ms_setup = 0, ms_all = 0;
    CEL_HIP(hipEventElapsedTime(&ms_setup, e0, e1));
    CEL_HIP(hipEventElapsedTime(&ms_all, e0, e2));
    double spmv_ms = 0;
    int spmv_n = 0;
    for (int i = 0; i + 1 < pev_used; i += 2) {
        float t = 0;
        if (hipEventElapsedTime(&t, pev[i], pev[i + 1]) == hipSuccess) {
            spmv_ms += t;
            ++spmv_n;
        }
    }
    for (auto ev : pev) (void)hipEventDestroy(ev);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    (void)hipEventDestroy(e2);

    if (prm.memspace == CEL_MEM_HOST)
        CEL_HIP(hipMemcpy(x, dx, sizeof(double) * n, hipMemcpyDeviceToHost));
    if (hist && hist_cap) {
        int cnt = std::min(out.iter + 1, hist_cap);
        CEL_HIP(hipMemcpy(hist, W.hist.p, sizeof(double) * cnt, hipMemcpyDeviceToHost));
    }

    stats.iters = out.iter;
    stats.outer_iters = out.outer;
    stats.norm0 = out.norm0;
    stats.final_norm = out.norm2;
    stats.eps_abs = out.eps;
    stats.alpha = out.alpha;
    stats.solve_ms = ms_all;
    stats.setup_ms = ms_setup;
    stats.spmv_ms = spmv_ms;
    stats.spmv_launches = spmv_n;
    const int prec = mixed ? CEL_PREC_MIXED : CEL_PREC_FP64;
    stats.spmv_bytes = spmv_bytes(A, fmt, prec);
    {
        const double vb = mixed ? 4.0 : 8.0;
        // K1 matrix + p gather + v write ; K2 r,v,jac read + r write ;
        // K3 p,r,jac,x read + p,x write
        stats.iter_bytes = stats.spmv_bytes + vb * n * 4.0 + vb * n * 6.0 +
                           (halo ? 20.0 * A.send_total + 12.0 * A.ghost_len : 0.0);
    }
    if (out.status != 0) {
        stats.status = out.status;
        return fail(out.status, out.status == CEL_ERR_BREAKDOWN
                                    ? "PCG breakdown: p.Ap <= 0 or not finite"
                                    : "non-finite right-hand side");
    }
    const bool converged = out.norm2 <= out.eps;
    stats.status = converged ? CEL_OK : CEL_NOT_CONVERGED;
    if (!converged && out.iter < prm.iter_max && !(std::fabs(out.alpha) > h.small))
        stats.status = CEL_OK;  // stopped on |alpha| <= 1.1 eps (reference condition)
    return stats.status;
}

void pcg_release_workspace(Mat* A)
{
    delete A->ws;
    A->ws = nullptr;
}

}  // namespace cel
