// spmv_device.hpp -- per-tile SpMV bodies shared by the plain SpMV kernels
// and the fused PCG kernels.
//
// CSR tile (CSR-stream): one 256-row tile per block iteration.  The tile's
// entries [xadj[r0], xadj[r0+256]) are contiguous, so the block streams
// values and column indices with coalesced 16-byte loads into registers (8
// entries per thread, all loads issued before the first wait; measured 6-7%
// faster than 4- and 8-byte loads on the 256^3 7-point block), parks them in LDS,
// and every thread then sums *its own row* from LDS in column order.  That
// keeps the reference's per-row summation order
// (src/fop/cel_fop_mv_module.f90:215-229) while the HBM stream stays fully
// coalesced.  The loads of tile t+1 are issued before the row sums of tile
// t, so they are in flight while tile t's gathers of x wait.  Tiles whose
// entries do not fit the LDS stage fall back to direct per-row loads (same
// order).
//
// Row sums gather up to 8 x values at once before adding them in column
// order, so a 7-point row costs one round trip to L2, not seven.
//
// JAD tile: thread k of the tile owns permuted row k; diagonal d stores
// entry k at begin[d] + k, so consecutive threads read consecutive
// addresses.  Diagonal order = column order of the CSR row, so the sum is
// bit-identical to the CSR sum.
#pragma once

#include "cel_internal.hpp"
#include "device_common.hpp"

namespace cel {

constexpr int kTileRows = kBlock;
constexpr int kStageCap = 2048;            // staged entries per tile (8 per row on average)
constexpr int kPer = kStageCap / kBlock;   // staged entries per thread
// entries of a tile that can be staged: the 16-byte loads of a tile start
// up to 3 entries before it (at a 16-byte boundary)
constexpr int kStageMax = kStageCap - 4;
constexpr int kGather = 8;                 // x values gathered per round trip

// Tile order of the kernels that walk whole 256-row tiles (the plain CSR
// product, the PCG K1s): tiles go in groups of kTileGroup
// consecutive tiles, group q to block q mod G; step k of a block covers
// tile kTileGroup*q + k mod kTileGroup of its (k / kTileGroup)-th group
// (-1: past the end).  Neighbouring blocks stream neighbouring groups:
// measured faster on the 256^3 7-point block than a contiguous tile range
// per block (PCG K1s 3-7%, plain CSR 7-8%).
constexpr int kTileGroup = 2;
__device__ __forceinline__ long long group_tile(long long k, long long ntiles)
{
    const long long ngroups = (ntiles + kTileGroup - 1) / kTileGroup;
    const long long q = (long long)blockIdx.x + (k / kTileGroup) * gridDim.x;
    const long long t = kTileGroup * q + k % kTileGroup;
    return (q < ngroups && t < ntiles) ? t : -1;
}

// Epilogue of a fused FP64 product row (kFuseDot / kFuseResid): stores the
// row's value (s = the plain product's value) and returns the row's term of
// the block's dot-product partial.
template <int FUSE>
__device__ __forceinline__ double fuse_row(double s, long long row, const double* __restrict__ aux,
                                           double* __restrict__ y)
{
    if constexpr (FUSE == kFuseDot) {
        y[row] = s;
        const double p = aux[row] * s;
        return p;
    } else {
        const double ri = aux[row] - s;  // r = b - A x
        y[row] = ri;
        const double p = ri * ri;
        return p;
    }
}

typedef unsigned int U32x4 __attribute__((ext_vector_type(4)));  // one 16-byte load

template <typename VT>
struct CsrStage {
    VT val[kStageCap];
    int col[kStageCap];
};

// sum + sum_{start <= j < len} a[j] * x[c[j]] in j order, one entry at a
// time: the path of rows longer than kGather and of tiles larger than the
// stage.  Kept lean on purpose -- its registers would count against the
// occupancy of the common path.  LDS or global a/c (NT: non-temporal
// global loads).
template <typename AT, typename VT, bool NT>
__device__ __forceinline__ AT row_sum_from(AT sum, const VT* a, const int* c, int start, int len,
                                           const VT* __restrict__ x)
{
#pragma unroll 1
    for (int j = start; j < len; ++j) {
        const int col = NT ? ld_stream(c + j) : c[j];
        const VT av = NT ? ld_stream(a + j) : a[j];
        AT prod = AT(av) * AT(x[col]);
        sum = sum + prod;
    }
    return sum;
}

// Buffer resource over `bytes` bytes at `base` (gfx9 raw buffer, stride 0):
// loads at offsets >= bytes return 0.
__device__ __forceinline__ __amdgpu_buffer_rsrc_t tile_rsrc(const void* base, int bytes)
{
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(base), (short)0, bytes,
                                             0x00020000);
}

// Non-temporal buffer load (aux bit 1 = nt on gfx950).
template <typename T>
__device__ __forceinline__ T buf_load_nt(__amdgpu_buffer_rsrc_t r, int voff, int soff);
template <>
__device__ __forceinline__ double buf_load_nt<double>(__amdgpu_buffer_rsrc_t r, int voff, int soff)
{
    return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, voff, soff, 2));
}
template <>
__device__ __forceinline__ float buf_load_nt<float>(__amdgpu_buffer_rsrc_t r, int voff, int soff)
{
    return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, voff, soff, 2));
}
template <>
__device__ __forceinline__ int buf_load_nt<int>(__amdgpu_buffer_rsrc_t r, int voff, int soff)
{
    return (int)__builtin_amdgcn_raw_buffer_load_b32(r, voff, soff, 2);
}

// Cached buffer load; an offset outside the resource returns 0.
template <typename T>
__device__ __forceinline__ T buf_load(__amdgpu_buffer_rsrc_t r, int voff);
template <>
__device__ __forceinline__ double buf_load<double>(__amdgpu_buffer_rsrc_t r, int voff)
{
    return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, voff, 0, 0));
}
template <>
__device__ __forceinline__ float buf_load<float>(__amdgpu_buffer_rsrc_t r, int voff)
{
    return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, voff, 0, 0));
}

// DIA row sums of R rows r[0..R) (r[k] >= n: not used; r[0] < n) of a block
// stored per diagonal (vals[d*n + i]; offsets ascending; bit d of the row
// mask marks an entry): each row's entries in ascending offset (= CSR
// column) order, absent ones skipped, so the sums are the CSR row sums bit
// for bit.  A group of GD diagonals of all R rows is loaded before any of
// their products is added; values through non-temporal buffer loads (one
// resource per diagonal), x through one buffer resource, so the x of an
// absent boundary entry reads 0 from outside the resource instead of
// needing a clamp.  n * sizeof(VT) must stay below 2^31 (32-bit byte
// offsets).  With xr, x[r[k]] is returned in xr[k] from the loads of the
// diagonal with offset 0 (index dzero >= 0, present in the caller's
// layout), taken after the products' wait, so it costs no load and no
// round trip of its own.
template <typename VT, int R, int GD>
__device__ __forceinline__ void dia_rows(int n, int nd, const int* doff, const VT* vals,
                                         const uint8_t* m8, const uint32_t* m32, const int (&r)[R],
                                         const VT* __restrict__ x, VT (&s)[R], VT* xr = nullptr,
                                         int dzero = -1)
{
    constexpr int B = (int)sizeof(VT);
    int rc[R];
    unsigned m[R];
#pragma unroll
    for (int k = 0; k < R; ++k) {
        rc[k] = r[k] < n ? r[k] : r[0];
        m[k] = m8 ? (unsigned)m8[rc[k]] : m32[rc[k]];  // m8: uniform
        s[k] = VT(0);
    }
    const auto rx = tile_rsrc(x, n * B);
    for (int d0 = 0; d0 < nd; d0 += GD) {
        VT v[GD][R], xv[GD][R];
#pragma unroll
        for (int g = 0; g < GD; ++g) {
            const int d = d0 + g;
            if (d < nd) {
                const auto rv = tile_rsrc(vals + (size_t)d * (size_t)n, n * B);
                const int o = doff[d];
#pragma unroll
                for (int k = 0; k < R; ++k) {
                    v[g][k] = buf_load_nt<VT>(rv, rc[k] * B, 0);
                    // an absent entry's column may lie outside [0, n): its
                    // byte offset wraps modulo 2^32 (unsigned, defined) and
                    // the load returns 0 or a value the mask discards
                    xv[g][k] = buf_load<VT>(rx, (int)((unsigned)(rc[k] + o) * (unsigned)B));
                }
            }
        }
#pragma unroll
        for (int g = 0; g < GD; ++g) {
            const int d = d0 + g;
            if (d < nd) {
#pragma unroll
                for (int k = 0; k < R; ++k) {
                    const VT prod = v[g][k] * xv[g][k];
                    s[k] = ((m[k] >> d) & 1u) ? s[k] + prod : s[k];
                }
                if (xr && d == dzero) {
#pragma unroll
                    for (int k = 0; k < R; ++k) xr[k] = xv[g][k];
                }
            }
        }
    }
}

struct NoLoads {
    __device__ __forceinline__ void operator()() const {}
};

// Software-pipelined CSR tile stream.  bounds(t) loads the entry range of
// tile t (two row pointers); fetch(t) issues the loads of tile t (its
// entries into registers when the tile fits the stage, and the calling
// thread's row bounds) from the range bounds() loaded one step earlier;
// stage() parks the entries in LDS; rows() sums the calling thread's row of
// the staged tile while the next tile loads:
//   bounds(t0); fetch(t0); bounds(t1); stage; sync
//   for t: s = rows(t, next = t1, after = t2); sync; stage; sync
// rows() issues, in this order: the gathers of x for tile t, the loads of
// tile t1, the row bounds of t1 and the entry range of t2.  Vector memory
// operations complete in issue order, so nothing in a step waits for a load
// issued later in it: the sum waits for the gathers only, the stage for the
// tile's entries only.  (With the entry range of t1 loaded in the same step
// as its entries, the tile's loads had to wait for the gathers issued before
// them, which put one L2 round trip and one HBM round trip in series on
// every step.)
template <typename VT>
struct CsrTileStream {
    static constexpr int EV = 16 / (int)sizeof(VT);  // values per 16-byte load
    VT v[kPer];
    int c[kPer];
    int e0, cnt;  // first entry and entries of the fetched tile (uniform)
    int lo, hi;   // the calling thread's row bounds (lo = hi = e0 past n)
    int sv, sc;   // entries the 16-byte value / column loads start before e0
    int pb0, pb1; // entry range of the tile bounds() was called for

    // Entry range of tile t (t < 0: tile 0, a valid load whose value the
    // next fetch uses with want = false).
    __device__ __forceinline__ void bounds(int n, int t, const int* __restrict__ xadj)
    {
        const int r0 = (t >= 0 ? t : 0) * kTileRows;
        pb0 = xadj[r0];
        pb1 = xadj[r0 + min(kTileRows, n - r0)];
    }

    // Loads of tile t, whose entry range the last bounds() call loaded.
    // Unconditional by design: with want = false (no next tile) or a tile
    // larger than the stage, the buffer covers 0 bytes and the loads return
    // 0 without touching memory, so the registers never merge loaded and
    // stale values (such merges force a wait for the loads before the
    // current tile's work).
    __device__ __forceinline__ void fetch(int n, int t, bool want, const int* __restrict__ xadj,
                                          const int* __restrict__ cols,
                                          const VT* __restrict__ vals)
    {
        const int tid = threadIdx.x;
        const int r0 = t * kTileRows;
        const int nrows = min(kTileRows, n - r0);
        // every lane loaded the same two words; take them as scalars so the
        // buffer resources are built in scalar registers
        e0 = __builtin_amdgcn_readfirstlane(pb0);
        cnt = __builtin_amdgcn_readfirstlane(pb1) - e0;
        const int m = (want && cnt <= kStageMax) ? cnt : 0;
        // 16-byte loads from the 16-byte boundary at or before e0: thread
        // tid holds the EV (values) or 4 (columns) consecutive entries at
        // EV*tid + EV*256*i (4*tid + 1024*i) counted from that boundary
        sv = e0 % EV;
        sc = e0 % 4;
        const auto rv = tile_rsrc(vals + (e0 - sv), m > 0 ? (m + sv) * (int)sizeof(VT) : 0);
        const auto rc = tile_rsrc(cols + (e0 - sc), m > 0 ? (m + sc) * (int)sizeof(int) : 0);
#pragma unroll
        for (int i = 0; i < kPer / EV; ++i) {
            const auto w = __builtin_amdgcn_raw_buffer_load_b128(rv, 16 * tid, 16 * kBlock * i, 2);
#pragma unroll
            for (int j = 0; j < EV; ++j) {
                if constexpr (sizeof(VT) == 8) {
                    const unsigned long long u =
                        (unsigned long long)(unsigned)w[2 * j] | ((unsigned long long)(unsigned)w[2 * j + 1] << 32);
                    v[i * EV + j] = __builtin_bit_cast(VT, u);
                } else {
                    const unsigned u = (unsigned)w[j];  // a plain value, not the vector element
                    v[i * EV + j] = __builtin_bit_cast(VT, u);
                }
            }
        }
#pragma unroll
        for (int i = 0; i < kPer / 4; ++i) {
            const auto w = __builtin_amdgcn_raw_buffer_load_b128(rc, 16 * tid, 16 * kBlock * i, 2);
#pragma unroll
            for (int j = 0; j < 4; ++j) c[i * 4 + j] = (int)w[j];
        }
        // threads past the last row get lo = hi (an empty row)
        lo = xadj[r0 + min(tid, nrows)];
        hi = xadj[r0 + min(tid + 1, nrows)];
    }

    // First tile of a walk: clears the staged columns (rows() gathers
    // through stale slots, which must hold valid columns), then its range,
    // its loads and the range of the tile after it.  Every thread of the
    // block calls it (it synchronises); the caller stages and synchronises.
    __device__ __forceinline__ void start(int n, int t, int next, const int* __restrict__ xadj,
                                          const int* __restrict__ cols,
                                          const VT* __restrict__ vals, CsrStage<VT>& sm)
    {
#pragma unroll
        for (int i = 0; i < kPer; ++i) sm.col[threadIdx.x + kBlock * i] = 0;
        __syncthreads();
        bounds(n, t, xadj);
        fetch(n, t, true, xadj, cols, vals);
        bounds(n, next, xadj);
    }

    // Parks the fetched entries in LDS (tiles that fit the stage).
    __device__ __forceinline__ void stage(CsrStage<VT>& sm)
    {
        if (cnt <= kStageMax) {  // uniform across the block
#pragma unroll
            for (int i = 0; i < kPer / EV; ++i)
#pragma unroll
                for (int j = 0; j < EV; ++j) {
                    const int k = EV * (int)threadIdx.x + EV * kBlock * i + j - sv;
                    if (k >= 0 && k < cnt) sm.val[k] = v[i * EV + j];
                }
#pragma unroll
            for (int i = 0; i < kPer / 4; ++i)
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    const int k = 4 * (int)threadIdx.x + 4 * kBlock * i + j - sc;
                    if (k >= 0 && k < cnt) sm.col[k] = c[i * 4 + j];
                }
        }
    }

    // Row sum of the staged tile t, init + a_1 x_1 + a_2 x_2 + ... in column
    // order (init = 0: the CSR row sum; init = y[row]: the COO product
    // y(row) += a x entry by entry); starts fetching tile `next` (< 0:
    // none) and loads the entry range of tile `after` (the tile following
    // next; < 0: none).  The gathers of x (and x[row] into *xrow when xrow
    // is not null) are issued first (see above).  The fetched registers are
    // consumed by the caller's stage() in the same iteration, never carried
    // across it.
    template <typename AT, typename Extra = NoLoads>
    __device__ __forceinline__ AT rows(int n, int t, int next, int after,
                                       const int* __restrict__ xadj, const int* __restrict__ cols,
                                       const VT* __restrict__ vals, const VT* __restrict__ x,
                                       const CsrStage<VT>& sm, int row, VT* xrow, AT init = AT(0),
                                       Extra&& extra = Extra{})
    {
        const int b0 = e0, l = lo - e0, len = hi - lo;
        const bool staged = cnt <= kStageMax;  // uniform across the block
        // Branch-free gathers and first kGather products: entries past the
        // row read a clamped LDS slot and x[0], and a tile that is not staged
        // reads stale slots (valid columns: start() clears them); those
        // products are not kept.  With a branch on `staged` here the
        // compiler moves loads into it, and the two paths then leave
        // different loads in flight, so the waits below (for the next
        // tile's entry range, for the gathers) would have to cover every
        // load issued, the next tile's included.
        VT xv[kGather];
#pragma unroll
        for (int i = 0; i < kGather; ++i) {
            const int ci = sm.col[min(l + i, kStageCap - 1)];
            xv[i] = x[i < len ? ci : 0];
        }
        VT xr = VT(0);
        if (xrow) xr = x[row < n ? row : 0];
        // the caller's loads for this step: behind the gathers (waiting for
        // those does not wait for these), ahead of the next tile's entries
        // (whose wait, for the stage, covers these)
        extra();
        // keep the issue order (gathers, then the next tile): the scheduler
        // would otherwise move the gathers next to their uses
        __builtin_amdgcn_sched_barrier(0);
        fetch(n, next >= 0 ? next : t, next >= 0, xadj, cols, vals);
        bounds(n, after, xadj);
        __builtin_amdgcn_sched_barrier(0);
        if (xrow) *xrow = xr;
        AT sum = init;
#pragma unroll
        for (int i = 0; i < kGather; ++i) {
            const VT a = sm.val[min(l + i, kStageCap - 1)];
            AT prod = AT(a) * AT(xv[i]);
            AT tt = sum + prod;
            sum = i < len ? tt : sum;
        }
        // Rows longer than kGather continue from LDS in column order; a tile
        // that is not staged sums every row from global memory.  Both take
        // the products above as their input (init for the second), so those
        // stay needed on every path (a path on which they were dead would
        // let the compiler move them, gathers included, into the other).
        sum = staged ? sum : init;
        const int j0 = staged ? min(len, kGather) : 0;
        if (j0 < len) {
            const VT* av = staged ? (const VT*)(sm.val + l) : vals + b0 + l;
            const int* cv = staged ? (const int*)(sm.col + l) : cols + b0 + l;
            sum = row_sum_from<AT, VT, false>(sum, av, cv, j0, len, x);
        }
        return sum;
    }
};

// JAD: permuted row k = k0 + threadIdx.x.  begin[d+1] - begin[d] = entries
// of diagonal d (non-increasing), begin[d] = start of diagonal d (0-based).
// kGather diagonals per round trip; the sum runs in diagonal order.
template <typename VT, typename AT>
__device__ __forceinline__ AT jad_row(int k, int mmax, const int* __restrict__ begin,
                                      const int* __restrict__ jcols,
                                      const VT* __restrict__ jvals, const VT* __restrict__ x)
{
    AT sum = AT(0);
    for (int d0 = 0; d0 < mmax; d0 += kGather) {
        VT av[kGather], xv[kGather];
        bool on[kGather];
#pragma unroll
        for (int i = 0; i < kGather; ++i) {
            const int d = d0 + i;
            on[i] = false;
            if (d < mmax) {
                const int b = begin[d];
                on[i] = k < begin[d + 1] - b;
                if (on[i]) {
                    const int col = ld_stream(jcols + b + k);
                    av[i] = ld_stream(jvals + b + k);
                    xv[i] = x[col];
                }
            }
        }
#pragma unroll
        for (int i = 0; i < kGather; ++i) {
            if (on[i]) {
                AT prod = AT(av[i]) * AT(xv[i]);
                sum = sum + prod;
            }
        }
        if (!on[kGather - 1]) break;
    }
    return sum;
}

// JAD rows in the permuted row order, one per thread and 256-row tile,
// software-pipelined like the CSR tile stream: row() issues the gathers of
// the current row (its columns and values were loaded by the previous
// call), then the loads of the next row, then sums.  The entry of permuted
// row k on diagonal d sits at begin[d] + k, so every load is coalesced.
// The lengths of the first kGather diagonals stay in scalar registers for
// the whole walk; the columns, values and x go through one buffer resource
// each, and an entry a row does not have is loaded from an offset outside
// the resource (it reads 0, is gathered from x[0] and is not added).
// Diagonals past kGather are summed after the first kGather, in order,
// from global memory.  jad_nnz * sizeof(VT) and n * sizeof(VT) must stay
// below 2^32 - 16 (unsigned 32-bit byte offsets).
template <typename VT>
struct JadStream {
    int bg[kGather];  // starts of diagonals 0..kGather-1 (clamped)
    int ln[kGather];  // their lengths (0 past the last diagonal)
    int c[kGather];   // the fetched row's columns and values
    VT a[kGather];
    __amdgpu_buffer_rsrc_t rc, rv, rx;

    __device__ __forceinline__ void init(int n, int mmax, long long nnz,
                                         const int* __restrict__ begin, const int* cols,
                                         const VT* vals, const VT* x)
    {
        constexpr unsigned B = sizeof(VT);
        const int nd = min(mmax, kGather);
#pragma unroll
        for (int i = 0; i < kGather; ++i) {
            bg[i] = __builtin_amdgcn_readfirstlane(begin[min(i, mmax)]);
            const int e = __builtin_amdgcn_readfirstlane(begin[min(i + 1, mmax)]);
            ln[i] = i < nd ? e - bg[i] : 0;
        }
        rc = __builtin_amdgcn_make_buffer_rsrc(const_cast<int*>(cols), (short)0,
                                               (int)((unsigned)nnz * 4u), 0x00020000);
        rv = __builtin_amdgcn_make_buffer_rsrc(const_cast<VT*>(vals), (short)0,
                                               (int)((unsigned)nnz * B), 0x00020000);
        rx = __builtin_amdgcn_make_buffer_rsrc(const_cast<VT*>(x), (short)0,
                                               (int)((unsigned)n * B), 0x00020000);
    }

    // every load is issued (entries a row lacks read from outside the
    // resources), so the loads in flight do not depend on the data and the
    // waits after them stay exact
    __device__ __forceinline__ void fetch(int k)
    {
        constexpr unsigned B = sizeof(VT);
#pragma unroll
        for (int i = 0; i < kGather; ++i) {
            const unsigned e = (unsigned)(bg[i] + k);
            const bool on = k < ln[i];
            c[i] = (int)__builtin_amdgcn_raw_buffer_load_b32(rc, on ? e * 4u : 0xFFFFFFFFu, 0, 2);
            a[i] = buf_load_nt<VT>(rv, (int)(on ? e * B : 0xFFFFFFFFu), 0);
        }
    }

    // Sum of permuted row k (fetched by the previous call) in diagonal
    // order; fetches row kn (kn >= n: loads that return 0) and returns
    // x[k] (0 for k >= n) in xr.
    template <typename AT>
    __device__ __forceinline__ AT row(int k, int kn, int mmax, const int* __restrict__ begin,
                                      const int* __restrict__ cols, const VT* __restrict__ vals,
                                      const VT* __restrict__ x, VT& xr)
    {
        constexpr unsigned B = sizeof(VT);
        VT xv[kGather];
#pragma unroll
        for (int i = 0; i < kGather; ++i) xv[i] = buf_load<VT>(rx, (int)((unsigned)c[i] * B));
        xr = buf_load<VT>(rx, (int)((unsigned)k * B));
        VT ac[kGather];
#pragma unroll
        for (int i = 0; i < kGather; ++i) ac[i] = a[i];
        __builtin_amdgcn_sched_barrier(0);
        fetch(kn);
        __builtin_amdgcn_sched_barrier(0);
        AT sum = AT(0);
#pragma unroll
        for (int i = 0; i < kGather; ++i) {
            AT prod = AT(ac[i]) * AT(xv[i]);
            AT tt = sum + prod;
            sum = k < ln[i] ? tt : sum;
        }
        if (mmax > kGather) {  // uniform; rare (more than kGather diagonals)
            for (int d = kGather; d < mmax; ++d) {
                const int b = begin[d];
                if (k >= begin[d + 1] - b) break;
                AT prod = AT(ld_stream(vals + b + k)) * AT(x[ld_stream(cols + b + k)]);
                sum = sum + prod;
            }
        }
        return sum;
    }
};

// Halo record of a row with halo entries: the products a*ghost of its
// first two halo entries in block order, 16 bytes at its compact index q
// (rows with halo entries in ascending order), written every SpMV by
// k_halo_rec.  The row's entry count comes from its halo flag (hflag[row]
// = min(count, 255)) and q from the wavefront: the first compact index of
// its 64-row group (wq[row / 64]) plus the flagged lanes below it.  A K1
// adds the products to the local sum one by one; a row with more than two
// halo entries takes them from the compact lists instead.  The records are
// read through a buffer resource over the compact rows (16 bytes each,
// below 2^31 - 64 in all): a row without halo entries reads from outside
// it, so the 16-byte load touches no memory and returns 0s.  Records of
// neighbouring rows sit next to each other, so a tile's records stream
// rather than touching one memory page per row.
constexpr int kHaloRec = 2;

struct HaloRec {
    U32x4 a;
};

// q0: the compact index of the wavefront's first flagged row (wq of its
// 64-row group); c: this lane's halo entry count
__device__ __forceinline__ HaloRec halo_rec_load(__amdgpu_buffer_rsrc_t rr, int q0, int c)
{
    const unsigned long long b = __builtin_amdgcn_ballot_w64(c != 0);
    const unsigned below = __builtin_amdgcn_mbcnt_hi((unsigned)(b >> 32), __builtin_amdgcn_mbcnt_lo((unsigned)b, 0u));
    const unsigned o = c != 0 ? (unsigned)(q0 + (int)below) * 16u : 0xFFFFFFC0u;
    HaloRec h;
    h.a = __builtin_amdgcn_raw_buffer_load_b128(rr, (int)o, 0, 0);
    return h;
}

__device__ __forceinline__ double u2d(unsigned lo, unsigned hi)
{
    return __builtin_bit_cast(double, (unsigned long long)lo | ((unsigned long long)hi << 32));
}

// sum + the record's products in order for a row with c halo entries;
// overflow: more than kHaloRec entries (the caller adds them from the
// compact lists instead)
template <typename AT>
__device__ __forceinline__ AT halo_rec_add(AT sum, const HaloRec& h, int c, bool& overflow)
{
    overflow = c > kHaloRec;
    if (c > 0) sum = sum + AT(u2d(h.a[0], h.a[1]));
    if (c > 1) sum = sum + AT(u2d(h.a[2], h.a[3]));
    return sum;
}

// Halo entries (rows of this rank with columns owned by a neighbour).
// hrow_tile[t]..hrow_tile[t+1] index the compact, ascending list of rows of
// tile t that have halo entries; entries of compact row q are
// hptr[q]..hptr[q+1] in block order, and their columns are ghost slots.

// Halo entries of row `row` of tile t found by binary search in the tile's
// compact rows (no LDS map; used where a kernel runs once per solve).
template <typename AT>
__device__ __forceinline__ AT halo_row(int row, int t, AT sum, const int* __restrict__ hrow_tile,
                                       const int* __restrict__ hrows,
                                       const int* __restrict__ hptr,
                                       const int* __restrict__ hcols,
                                       const double* __restrict__ hvals,
                                       const double* __restrict__ ghost)
{
    int lo = hrow_tile[t];
    const int end = hrow_tile[t + 1];
    if (lo == end) return sum;
    int hi = end;
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (hrows[mid] < row) lo = mid + 1;
        else hi = mid;
    }
    if (lo < end && hrows[lo] == row) {
        for (int j = hptr[lo]; j < hptr[lo + 1]; ++j) {
            AT prod = AT(hvals[j]) * AT(ghost[hcols[j]]);
            sum = sum + prod;
        }
    }
    return sum;
}

}  // namespace cel
