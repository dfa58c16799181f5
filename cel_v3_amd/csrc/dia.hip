// dia.hip -- DIA form of a CSR block and its SpMV (no reference counterpart).
//
// A block whose entries lie on a few diagonals (structured-grid stencils)
// is stored per diagonal: vals[d * n + i] is the entry of row i in column
// i + off[d] (0 where absent), bit d of mask[i] tells whether it exists.
// Offsets ascend, so a row sum over d visits the row's entries in the CSR
// column order; absent entries are skipped (not added as zeros: 0 * x is
// not +0 for every x), so the product is bit-identical to the CSR one,
// without column indices and with x read contiguously per diagonal.  The
// mask is one byte per row when there are at most 8 diagonals.
#include <hipcub/hipcub.hpp>

#include "cel_internal.hpp"
#include "spmv_device.hpp"

namespace cel {

constexpr int kMaxDiag = 32;

namespace {

int grid_of(long long n)
{
    long long g = (n + kBlock - 1) / kBlock;
    long long cap = (long long)rt().num_cus * 16;
    return (int)(g < 1 ? 1 : g > cap ? cap : g);
}

// off[j] = col - row for every entry; flags rows whose columns do not
// strictly ascend.
__global__ void __launch_bounds__(kBlock)
k_dia_offsets(int n, const int* __restrict__ xadj, const int* __restrict__ cols,
              int* __restrict__ off, int* __restrict__ bad)
{
    int any = 0;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (long long)gridDim.x * blockDim.x) {
        int prev = -1;
        for (int j = xadj[i]; j < xadj[i + 1]; ++j) {
            const int c = cols[j];
            any |= c <= prev;
            prev = c;
            off[j] = c - (int)i;
        }
    }
    if (any) atomicOr(bad, 1);
}

struct Offsets {
    int nd;
    int off[kMaxDiag];
};

__global__ void __launch_bounds__(kBlock)
k_dia_fill(int n, const int* __restrict__ xadj, const int* __restrict__ cols,
           const double* __restrict__ vals, Offsets O, double* __restrict__ dv,
           uint32_t* __restrict__ mask, uint8_t* __restrict__ mask8)
{
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (long long)gridDim.x * blockDim.x) {
        unsigned m = 0;
        for (int j = xadj[i]; j < xadj[i + 1]; ++j) {
            const int o = cols[j] - (int)i;
            int d = 0;
            while (d < O.nd - 1 && O.off[d] != o) ++d;
            dv[(size_t)d * n + i] = vals[j];
            m |= 1u << d;
        }
        if (mask8) mask8[i] = (uint8_t)m;
        else mask[i] = m;
    }
}

Offsets offsets_of(const Mat& A)
{
    Offsets O{};
    O.nd = A.dia_nd;
    for (int d = 0; d < A.dia_nd; ++d) O.off[d] = A.dia_off[d];
    return O;
}

// y = A x (accumulate: y += A x); for every group T of R 256-row tiles it
// owns (T = block, block + G, ...), each thread sums rows T*R*256 + tid +
// 256k, k < R, loading GD diagonals of all R rows at a time.
template <typename VT, int R, int GD>
__global__ void __launch_bounds__(kBlock)
k_dia_spmv(int n, Offsets O, const VT* __restrict__ dv, const uint32_t* __restrict__ mask,
           const uint8_t* __restrict__ mask8, const VT* __restrict__ x, VT* __restrict__ y,
           int accumulate, const int* __restrict__ skip)
{
    if (skip && *skip) return;
    const long long groups = ((long long)n + R * kBlock - 1) / (R * kBlock);
    for (long long T = blockIdx.x; T < groups; T += gridDim.x) {
        int r[R];
#pragma unroll
        for (int k = 0; k < R; ++k) r[k] = (int)(T * R * kBlock) + threadIdx.x + k * kBlock;
        if (r[0] >= n) continue;
        VT s[R];
        dia_rows<VT, R, GD>(n, O.nd, O.off, dv, mask8, mask, r, x, s);
#pragma unroll
        for (int k = 0; k < R; ++k)
            if (r[k] < n) y[r[k]] = accumulate ? y[r[k]] + s[k] : s[k];
    }
}

// Launch shapes of the plain product, measured on the 256^3 7-point block
// (20 back-to-back launches): FP32 4 rows per thread at 2 blocks per CU
// (100 us against 106-113 us for 2 rows per thread at 2-8 blocks per CU);
// FP64 8 rows per thread, 4 diagonals per load group, at 8 blocks per CU
// (225 us against 231 us).
template <typename VT>
struct DiaShape;
template <>
struct DiaShape<float> {
    static constexpr int R = 4, GD = 8, per_cu = 2;
};
template <>
struct DiaShape<double> {
    static constexpr int R = 8, GD = 4, per_cu = 8;
};

template <typename VT>
int launch_dia(const Mat& A, const VT* vals, const VT* x, VT* y, int accumulate, const int* skip)
{
    using S = DiaShape<VT>;
    const long long groups = ((long long)A.n + S::R * kBlock - 1) / (S::R * kBlock);
    const int G = (int)std::min<long long>(groups, (long long)rt().num_cus * S::per_cu);
    hipLaunchKernelGGL((k_dia_spmv<VT, S::R, S::GD>), dim3(G), dim3(kBlock), 0, rt().stream, A.n,
                       offsets_of(A), vals, A.dia_mask.p, A.dia_mask8.p, x, y, accumulate, skip);
    CEL_HIP(hipGetLastError());
    return CEL_OK;
}

// FP64 product fused with the next dot product (kFuseDot / kFuseResid,
// cel_internal.hpp): block b covers the rows of its chunk [b*chunk,
// b*chunk + chunk) in steps of R*256 rows, R rows 256 apart per thread (the
// plain FP64 shape), so each thread meets its rows in ascending order.
template <int FUSE>
__global__ void __launch_bounds__(kBlock)
k_dia_spmv_fused(int n, long long chunk, Offsets O, const double* __restrict__ dv,
                 const uint32_t* __restrict__ mask, const uint8_t* __restrict__ mask8,
                 const double* __restrict__ x, double* __restrict__ y,
                 const double* __restrict__ aux, double* __restrict__ part,
                 const int* __restrict__ skip)
{
    if (skip && *skip) return;  // uniform across the grid
    __shared__ double red[kWaves];
    constexpr int R = DiaShape<double>::R, GD = DiaShape<double>::GD;
    const long long c0 = (long long)blockIdx.x * chunk;
    const long long c1 = c0 + chunk < n ? c0 + chunk : n;
    double acc = 0.0;
    for (long long s0 = c0; s0 < c1; s0 += (long long)R * kBlock) {
        int r[R];
#pragma unroll
        for (int k = 0; k < R; ++k) {
            const long long rk = s0 + threadIdx.x + (long long)k * kBlock;
            r[k] = rk < c1 ? (int)rk : n;  // rows past the chunk are not used
        }
        if (r[0] >= n) continue;  // no barrier in this loop
        double s[R];
        dia_rows<double, R, GD>(n, O.nd, O.off, dv, mask8, mask, r, x, s);
#pragma unroll
        for (int k = 0; k < R; ++k)
            if (r[k] < n) acc = acc + fuse_row<FUSE>(s[k], r[k], aux, y);
    }
    const double bs = block_sum(acc, red);
    if (threadIdx.x == 0) part[blockIdx.x] = bs;
}


}  // namespace

int launch_dia_spmv(const Mat& A, const double* x, double* y, int accumulate, const int* skip)
{
    if (A.n == 0) return CEL_OK;
    return launch_dia<double>(A, A.dia_vals.p, x, y, accumulate, skip);
}

int launch_dia_spmv_fused(const Mat& A, int fuse, long long chunk, int G, const double* x,
                          double* y, const double* aux, double* part, const int* skip)
{
    if (A.n == 0 || A.dia_nd == 0 || A.n != A.ncols)
        return fail(CEL_ERR_ARG, "fused product: a square block with a DIA form");
    if (chunk <= 0 || chunk % kBlock != 0 || (long long)G * chunk < A.n)
        return fail(CEL_ERR_ARG, "fused product: chunks must be whole tiles covering every row");
    if (fuse == kFuseDot)
        hipLaunchKernelGGL(k_dia_spmv_fused<kFuseDot>, dim3(G), dim3(kBlock), 0, rt().stream, A.n,
                           chunk, offsets_of(A), A.dia_vals.p, A.dia_mask.p, A.dia_mask8.p, x, y,
                           aux, part, skip);
    else if (fuse == kFuseResid)
        hipLaunchKernelGGL(k_dia_spmv_fused<kFuseResid>, dim3(G), dim3(kBlock), 0, rt().stream, A.n,
                           chunk, offsets_of(A), A.dia_vals.p, A.dia_mask.p, A.dia_mask8.p, x, y,
                           aux, part, skip);
    else
        return fail(CEL_ERR_ARG, "fused product: unknown epilogue");
    CEL_HIP(hipGetLastError());
    return CEL_OK;
}

int dia_prepare_f32(Mat& A)
{
    if (A.dia_vals32.p || A.dia_nd == 0) return CEL_OK;
    const size_t cnt = (size_t)A.dia_nd * (size_t)A.n;
    CEL_TRY(A.dia_vals32.alloc(cnt > 0 ? cnt : 1));
    return launch_to_f32(A.dia_vals.p, A.dia_vals32.p, cnt);
}

int launch_dia_spmv_f32(Mat& A, const float* x, float* y, int accumulate)
{
    if (A.n == 0) return CEL_OK;
    CEL_TRY(dia_prepare_f32(A));
    return launch_dia<float>(A, A.dia_vals32.p, x, y, accumulate, nullptr);
}

}  // namespace cel

using namespace cel;

extern "C" int cel_gpu_mat_to_dia(cel_gpu_mat* A)
{
    CEL_TRY(ensure_init());
    if (!A) return fail(CEL_ERR_ARG, "null matrix");
    Mat& M = A->m;
    if (M.dia_nd > 0) return CEL_OK;
    if (M.n == 0 || M.nnz == 0) return fail(CEL_ERR_BAD_MATRIX, "an empty block has no DIA form");
    // the DIA row sums read x through a resource of n entries (the row
    // count), so a column index must also be a row index
    if (M.ncols != M.n) return fail(CEL_ERR_BAD_MATRIX, "a non-square block has no DIA form");
    // the DIA kernels address a diagonal with 32-bit byte offsets
    if ((long long)M.n * (long long)sizeof(double) >= (1LL << 31))
        return fail(CEL_ERR_BAD_MATRIX, "more than 2^28 rows: no DIA form");
    hipStream_t st = rt().stream;
    const long long nnz = M.nnz;
    DevBuf<int> off, sorted, uniq, counts, nruns, flag;
    CEL_TRY(off.alloc(nnz));
    CEL_TRY(flag.alloc(1));
    CEL_HIP(hipMemsetAsync(flag.p, 0, sizeof(int), st));
    hipLaunchKernelGGL(k_dia_offsets, dim3(grid_of(M.n)), dim3(kBlock), 0, st, M.n, M.xadj.p,
                       M.cols.p, off.p, flag.p);
    CEL_HIP(hipGetLastError());
    int bad = 0;
    CEL_HIP(hipMemcpyAsync(&bad, flag.p, sizeof(int), hipMemcpyDeviceToHost, st));
    CEL_HIP(hipStreamSynchronize(st));
    if (bad) return fail(CEL_ERR_BAD_MATRIX, "columns not strictly ascending in every row: no DIA form");
    // distinct offsets: sort, then run-length encode
    CEL_TRY(sorted.alloc(nnz));
    CEL_TRY(uniq.alloc(nnz));
    CEL_TRY(counts.alloc(nnz));
    CEL_TRY(nruns.alloc(1));
    size_t b1 = 0, b2 = 0;
    CEL_HIP(hipcub::DeviceRadixSort::SortKeys(nullptr, b1, off.p, sorted.p, (int)nnz, 0, 32, st));
    CEL_HIP(hipcub::DeviceRunLengthEncode::Encode(nullptr, b2, sorted.p, uniq.p, counts.p, nruns.p,
                                                  (int)nnz, st));
    DevBuf<unsigned char> temp;
    CEL_TRY(temp.alloc(std::max(b1, b2) > 0 ? std::max(b1, b2) : 1));
    // signed keys: radix sort orders int by value (two's complement handled by hipCUB)
    CEL_HIP(hipcub::DeviceRadixSort::SortKeys(temp.p, b1, off.p, sorted.p, (int)nnz, 0, 32, st));
    CEL_HIP(hipcub::DeviceRunLengthEncode::Encode(temp.p, b2, sorted.p, uniq.p, counts.p, nruns.p,
                                                  (int)nnz, st));
    int nd = 0;
    CEL_HIP(hipMemcpyAsync(&nd, nruns.p, sizeof(int), hipMemcpyDeviceToHost, st));
    CEL_HIP(hipStreamSynchronize(st));
    if (nd > kMaxDiag)
        return fail(CEL_ERR_BAD_MATRIX, std::to_string(nd) + " diagonals: more than the DIA form holds (32)");
    if ((long long)nd * M.n > 2 * nnz)
        return fail(CEL_ERR_BAD_MATRIX, "DIA would store more than twice nnz slots");
    int hoff[kMaxDiag] = {};
    CEL_HIP(hipMemcpyAsync(hoff, uniq.p, sizeof(int) * nd, hipMemcpyDeviceToHost, st));
    CEL_HIP(hipStreamSynchronize(st));
    const size_t cnt = (size_t)nd * (size_t)M.n;
    CEL_TRY(M.dia_vals.alloc(cnt));
    // exactly one mask array: the kernels read both (the other reads 0)
    if (nd <= 8) {
        M.dia_mask.release();
        CEL_TRY(M.dia_mask8.alloc(M.n));
    } else {
        M.dia_mask8.release();
        CEL_TRY(M.dia_mask.alloc(M.n));
    }
    CEL_HIP(hipMemsetAsync(M.dia_vals.p, 0, sizeof(double) * cnt, st));
    Offsets O{};
    O.nd = nd;
    for (int d = 0; d < nd; ++d) O.off[d] = hoff[d];
    hipLaunchKernelGGL(k_dia_fill, dim3(grid_of(M.n)), dim3(kBlock), 0, st, M.n, M.xadj.p, M.cols.p,
                       M.vals.p, O, M.dia_vals.p, nd <= 8 ? nullptr : M.dia_mask.p,
                       nd <= 8 ? M.dia_mask8.p : nullptr);
    CEL_HIP(hipGetLastError());
    CEL_HIP(hipStreamSynchronize(st));
    M.dia_nd = nd;
    for (int d = 0; d < nd; ++d) M.dia_off[d] = hoff[d];
    M.dia_vals32.release();
    return CEL_OK;
}

extern "C" int cel_gpu_mat_dia_info(const cel_gpu_mat* A, int* ndiag, int* offsets)
{
    if (!A) return fail(CEL_ERR_ARG, "null matrix");
    if (ndiag) *ndiag = A->m.dia_nd;
    if (offsets)
        for (int d = 0; d < A->m.dia_nd; ++d) offsets[d] = A->m.dia_off[d];
    return CEL_OK;
}
