// spmv_device.hpp -- per-tile SpMV bodies shared by the plain SpMV kernels
// and the fused PCG kernels.
//
// CSR tile (CSR-stream): one 256-row tile per block iteration.  The tile's
// entries [xadj[r0], xadj[r0+256]) are contiguous, so the block first
// streams values and column indices into LDS with coalesced loads, then
// every thread sums its own row from LDS in column order.  That keeps the
// reference's per-row summation order (src/fop/cel_fop_mv_module.f90:215-229)
// while the HBM stream stays fully coalesced.  Tiles whose entries do not
// fit the LDS stage fall back to direct per-row loads (same order).
//
// JAD tile: thread k of the tile owns permuted row k; diagonal d stores
// entry k at begin[d] + k, so consecutive threads read consecutive
// addresses.  Diagonal order = column order of the CSR row, so the sum is
// bit-identical to the CSR sum.
#pragma once

#include "device_common.hpp"

namespace cel {

constexpr int kTileRows = kBlock;
constexpr int kStageCap = 2048;  // staged entries per tile (8 per row on average)

template <typename VT>
struct CsrStage {
    VT val[kStageCap];
    int col[kStageCap];
    int ptr[kTileRows + 1];
};

// Returns the row sum of row r0 + threadIdx.x (0 for threads past n).
// Contains __syncthreads(); every thread of the block must call it.
template <typename VT, typename AT>
__device__ __forceinline__ AT csr_tile(int n, int r0, const int* __restrict__ xadj,
                                       const int* __restrict__ cols,
                                       const VT* __restrict__ vals, const VT* __restrict__ x,
                                       CsrStage<VT>& sm)
{
    const int tid = threadIdx.x;
    const int nrows = min(kTileRows, n - r0);
    // nrows + 1 row pointers: a full tile needs 257 of them, one more than
    // the block has threads
    for (int k = tid; k <= nrows; k += kBlock) sm.ptr[k] = xadj[r0 + k];
    __syncthreads();
    const int e0 = sm.ptr[0];
    const int cnt = sm.ptr[nrows] - e0;
    AT sum = AT(0);
    if (cnt <= kStageCap) {
        for (int k = tid; k < cnt; k += kBlock) {
            sm.val[k] = ld_stream(vals + e0 + k);
            sm.col[k] = ld_stream(cols + e0 + k);
        }
        __syncthreads();
        if (tid < nrows) {
            const int j1 = sm.ptr[tid + 1] - e0;
            for (int j = sm.ptr[tid] - e0; j < j1; ++j) {
                AT prod = AT(sm.val[j]) * AT(x[sm.col[j]]);
                sum = sum + prod;
            }
        }
    } else if (tid < nrows) {
        const int j1 = sm.ptr[tid + 1];
        for (int j = sm.ptr[tid]; j < j1; ++j) {
            AT prod = AT(ld_stream(vals + j)) * AT(x[ld_stream(cols + j)]);
            sum = sum + prod;
        }
    }
    __syncthreads();
    return sum;
}

// JAD: permuted row k = k0 + threadIdx.x.  dlen[d] = entries of diagonal d
// (non-increasing), begin[d] = start of diagonal d (0-based).
template <typename VT, typename AT>
__device__ __forceinline__ AT jad_row(int k, int mmax, const int* __restrict__ begin,
                                      const int* __restrict__ jcols,
                                      const VT* __restrict__ jvals, const VT* __restrict__ x)
{
    AT sum = AT(0);
    for (int d = 0; d < mmax; ++d) {
        const int b = begin[d];
        if (k >= begin[d + 1] - b) break;
        AT prod = AT(ld_stream(jvals + b + k)) * AT(x[ld_stream(jcols + b + k)]);
        sum = sum + prod;
    }
    return sum;
}

// Halo entries of one tile.  hrow_tile[t]..hrow_tile[t+1] index the compact
// list of rows with halo entries (ascending); entries of compact row q are
// hptr[q]..hptr[q+1] in block order, columns are ghost slots.  Adds them to
// `sum` of the calling thread's row in that order.  Contains __syncthreads().
template <typename AT>
__device__ __forceinline__ AT halo_tile(int r0, int t, AT sum, const int* __restrict__ hrow_tile,
                                        const int* __restrict__ hrows,
                                        const int* __restrict__ hptr,
                                        const int* __restrict__ hcols,
                                        const double* __restrict__ hvals,
                                        const double* __restrict__ ghost, int* smap)
{
    const int q0 = hrow_tile[t], q1 = hrow_tile[t + 1];
    if (q0 == q1) return sum;  // uniform across the block
    smap[threadIdx.x] = -1;
    __syncthreads();
    for (int q = q0 + threadIdx.x; q < q1; q += kBlock) smap[hrows[q] - r0] = q;
    __syncthreads();
    const int q = smap[threadIdx.x];
    if (q >= 0) {
        for (int j = hptr[q]; j < hptr[q + 1]; ++j) {
            AT prod = AT(hvals[j]) * AT(ghost[hcols[j]]);
            sum = sum + prod;
        }
    }
    __syncthreads();
    return sum;
}

}  // namespace cel
