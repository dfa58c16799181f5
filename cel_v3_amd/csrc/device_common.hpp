// device_common.hpp -- gfx950 device helpers shared by every kernel file.
//
//  * 256-thread blocks = 4 wavefronts of 64 lanes.
//  * Grids are "persistent": G blocks (a multiple of the 8 XCDs) each own
//    a contiguous range of tiles.  The hardware hands block b to XCD b % 8,
//    so the logical block lb = (b % 8) * (G / 8) + b / 8 gives every XCD one
//    contiguous slab of rows: stencil neighbours of a row (the +-nx and
//    +-nx*ny columns) are then read by the same XCD and hit its L2.
//  * Reductions are deterministic: a fixed per-thread order, a fixed
//    butterfly inside the wavefront, a fixed order across waves, one partial
//    per logical block; the consuming kernel sums the partials in
//    logical-block order (sum_partials), in every block that needs them.
#pragma once

#include <hip/hip_runtime.h>

namespace cel {

constexpr int kBlock = 256;
constexpr int kWaves = kBlock / 64;
constexpr int kXcds = 8;

__device__ __forceinline__ int logical_block()
{
    const int G = gridDim.x;
    const int b = blockIdx.x;
    if (G % kXcds) return b;
    return (b % kXcds) * (G / kXcds) + b / kXcds;
}

// Streaming loads of the matrix: read once per SpMV, keep them out of L2 so
// the gathered vector stays resident.
template <typename T>
__device__ __forceinline__ T ld_stream(const T* p)
{
    return __builtin_nontemporal_load(p);
}

__device__ __forceinline__ double wave_sum(double v)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v = v + __shfl_xor(v, off, 64);
    return v;
}

// Sum over the block; valid in thread 0.  Order: butterfly in each wave,
// then waves 0..3 in order.
__device__ __forceinline__ double block_sum(double v, double* smem /* kWaves */)
{
    v = wave_sum(v);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (lane == 0) smem[wave] = v;
    __syncthreads();
    double s = 0.0;
    if (threadIdx.x == 0) {
        for (int w = 0; w < kWaves; ++w) s = s + smem[w];
    }
    __syncthreads();
    return s;
}

// Two sums at once (same order as block_sum).
__device__ __forceinline__ void block_sum2(double& a, double& b, double* smem /* 2*kWaves */)
{
    a = wave_sum(a);
    b = wave_sum(b);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (lane == 0) {
        smem[wave] = a;
        smem[kWaves + wave] = b;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double sa = 0.0, sb = 0.0;
        for (int w = 0; w < kWaves; ++w) {
            sa = sa + smem[w];
            sb = sb + smem[kWaves + w];
        }
        a = sa;
        b = sb;
    }
    __syncthreads();
}

// Sum of G partials in logical-block order by the whole block: thread t
// sums t, t+256, ... sequentially, then block_sum.  Valid in thread 0.
// Every block of the consuming kernel reads the same partials, so they are
// plain (cached) loads: after the first block on an XCD the rest hit its L2.
__device__ __forceinline__ double sum_partials(const double* partials, int G, double* smem)
{
    double s = 0.0;
    for (int i = threadIdx.x; i < G; i += kBlock) s = s + partials[i];
    return block_sum(s, smem);
}

}  // namespace cel
