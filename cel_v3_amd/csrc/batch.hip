// batch.hip -- many independent Jacobi-PCG systems in one kernel ring:
// systems of uneven size and multi-right-hand-side blocks, FP64, CSR.
//
// No reference counterpart: the reference solves one system per call of
// cel_cgalg_jacobi (src/fcgalg/cel_cgalg_module.f90:178-305).  Every
// right-hand side of every system is a *lane* with its own scalars
// (alpha, sigma, norm, iteration count, stop decision), and each lane runs
// exactly the reference's op sequence.
//
// Layout.  The vectors of all systems are concatenated, each starting at
// a multiple of 64 doubles; system s owns n_s * kp_s doubles, stored
// row-major [row][kp_s] with kp_s = the rhs count rounded up to a power of
// two (padding lanes get b = 0 and stop at once).  In the SpMV the kp lanes
// of one row read x[col][0..kp) as one contiguous segment (kp/2 threads,
// two lanes each) and the matrix entries of a row are read once for all kp
// lanes; the vector kernels cover 256 elements per block and step.
//
// Work tables.  Every block of a kernel's grid is assigned to one system
// (a contiguous block range per system, sized by the system's work) and
// owns a contiguous range of that system's tiles/elements.  K1 and the
// set-up share one table (ranges b0/b1); K2 and K3 share one of their own,
// over a grid of their own (ranges c0/c1).  Partial dot products are per
// (block, lane); the consumer kernel sums the producer's partials of the
// system's block range in block order, so all lane scalars are
// deterministic for given tables.  The tables live in device memory:
// between chunks of iterations the host re-deals the blocks of systems that
// have stopped to the systems still running, without re-capturing the
// hipGraph that replays the chunk.
//
// Token ring per lane, as in the single-system solver: K1 reads slot 0 and
// writes slot 1, K2 1 -> 2, K3 2 -> 0; the first block of a system writes
// its lanes' tokens.  Stopped lanes are forwarded unchanged and do no work.
#include <algorithm>
#include <cmath>
#include <map>
#include <memory>
#include <numeric>
#include <vector>

#include "cel_internal.hpp"
#include "device_common.hpp"
#include "spmv_device.hpp"

namespace cel {

constexpr int kMaxRhs = 64;
// multi-rhs SpMV: row steps whose gathers are issued together (one: 80
// VGPRs, four waves per SIMD; two or four steps measured slower)
constexpr int kPairSteps = 1;
typedef double Dbl2 __attribute__((ext_vector_type(2)));
constexpr int kUnroll = 4;  // elements per thread and step in the vector kernels
constexpr int kBStage = 2048;  // staged entries per tile
// every system's vectors start at a multiple of 64 doubles (512 bytes): the
// kp-lane rows of a multi-rhs system then never straddle a 128-byte line
// and the lane pairs of its SpMV are 16-byte aligned
constexpr int kVecAlign = 64;

struct BCtrl {
    double alpha, sigma, norm2, eps, norm0, bnorm, pad0, pad1;
    int run, iter, status, reduce;
    int pad2[4];
};

struct BSys {
    const int* xadj;
    const int* cols;
    const double* vals;
    const double* jac;
    long long voff;  // first element of the system's vectors
    int n, kp, ksh;  // rows, padded lanes (power of two), log2(kp)
    int lane0;       // first lane of the system
    int b0, b1;      // current block range of K1 and the set-up (their table)
    int c0, c1;      // current block range of K2 and K3 (their own table and grid)
    int ntiles;      // tiles of 256/kp rows
    int period;      // multi-rhs: band period in 256-row chunks (0: none)
};

struct BBlk {
    int sys;         // -1: idle
    int t0, t1;      // K1 tiles [t0, t1)
    long long u0, u1;  // K2/K3 units of 256 elements [u0, u1)
};

struct BPrm {
    double tol, small;
    int tol_mode, iter_max, alpha_stop, kpmax;
};

enum { kBReduceNone = 0, kBReduceSetup = 1 };

__device__ __forceinline__ bool b_alpha_ok(const BPrm& P, double a)
{
    return !P.alpha_stop || fabs(a) > P.small;
}

// Per-lane block sum: lanes j = tid & (kp-1) of every row are summed in a
// fixed order (butterfly over the row bits inside each wave, then waves in
// order).  Valid in threads tid < kp.
__device__ __forceinline__ double lane_block_sum(double v, int kp, double* smem /* 4*64 */)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1)
        if (off >= kp) v = v + __shfl_xor(v, off, 64);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (lane < kp) smem[wave * 64 + lane] = v;
    __syncthreads();
    double s = 0.0;
    if (threadIdx.x < kp) {
        for (int w = 0; w < kWaves; ++w) s = s + smem[w * 64 + threadIdx.x];
    }
    __syncthreads();
    return s;
}

// The same for threads that hold two adjacent lanes each (kp >= 2): thread
// tid holds lanes 2 (tid & (kp/2 - 1)) and that + 1.  Valid in threads
// tid < kp, which get lane tid.
__device__ __forceinline__ double lane_pair_block_sum(double v0, double v1, int kp,
                                                      double* smem /* 4*64 */)
{
    const int h = kp >> 1;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1)
        if (off >= h) {
            v0 = v0 + __shfl_xor(v0, off, 64);
            v1 = v1 + __shfl_xor(v1, off, 64);
        }
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (lane < h) {
        smem[wave * 64 + 2 * lane] = v0;
        smem[wave * 64 + 2 * lane + 1] = v1;
    }
    __syncthreads();
    double s = 0.0;
    if (threadIdx.x < kp) {
        for (int w = 0; w < kWaves; ++w) s = s + smem[w * 64 + threadIdx.x];
    }
    __syncthreads();
    return s;
}

// Sum over blocks [b0, b1) of part[b * kpmax + j] for every lane j < kp,
// in a fixed order; result in out[j] (LDS), visible to the whole block.
// Plain loads: every block of the system reads the same partials, which
// after the first read on an XCD sit in its L2.
__device__ __forceinline__ void lane_partials(const double* __restrict__ part, int kpmax, int b0,
                                              int b1, int kp, int ksh, double* out, double* smem)
{
    const int j = threadIdx.x & (kp - 1);
    const int step = kBlock >> ksh;
    double acc = 0.0;
    for (int b = b0 + (threadIdx.x >> ksh); b < b1; b += step)
        acc = acc + part[(size_t)b * kpmax + j];
    const double s = lane_block_sum(acc, kp, smem);
    if (threadIdx.x < kp) out[threadIdx.x] = s;
    __syncthreads();
}

__device__ __forceinline__ double b_eps(const BPrm& P, double bnorm, double r0)
{
    if (P.tol_mode == CEL_TOL_REL_B) return P.tol * bnorm;
    if (P.tol_mode == CEL_TOL_REL_R0) return P.tol * r0;
    return P.tol;
}

struct BStage {
    double val[kBStage];
    int col[kBStage];
};

// Row sum of row r0 + (tid >> ksh), lane tid & (kp-1), of tile [r0, r0+R):
// entries staged in LDS when they fit.  x points at the system's vectors.
// Contains __syncthreads().
__device__ __forceinline__ double btile(const BSys& S, int r0, const double* __restrict__ x,
                                        BStage& sm, bool lane_on, int& row)
{
    const int tid = threadIdx.x;
    const int R = kBlock >> S.ksh;
    const int nrows = min(R, S.n - r0);
    const int e0 = S.xadj[r0];
    const int cnt = S.xadj[r0 + nrows] - e0;
    const int my = tid >> S.ksh, j = tid & (S.kp - 1);
    row = r0 + my;
    int lo = e0, hi = e0;
    if (my < nrows) {
        lo = S.xadj[row];
        hi = S.xadj[row + 1];
    }
    const bool staged = cnt <= kBStage;
    if (staged) {
        for (int k = tid; k < cnt; k += kBlock) {
            sm.val[k] = __builtin_nontemporal_load(S.vals + e0 + k);
            sm.col[k] = __builtin_nontemporal_load(S.cols + e0 + k);
        }
    }
    __syncthreads();
    double sum = 0.0;
    if (my < nrows && lane_on) {
        if (staged) {
            for (int k = lo - e0; k < hi - e0; ++k) {
                double prod = sm.val[k] * x[((size_t)sm.col[k] << S.ksh) + j];
                sum = sum + prod;
            }
        } else {
            for (int k = lo; k < hi; ++k) {
                double prod = S.vals[k] * x[((size_t)S.cols[k] << S.ksh) + j];
                sum = sum + prod;
            }
        }
    }
    __syncthreads();
    return sum;
}

// Reads the lane tokens of system S into LDS (thread j < kp reads lane j).
__device__ __forceinline__ void read_lanes(const BCtrl* __restrict__ cin, const BSys& S, BCtrl* lc)
{
    if (threadIdx.x < S.kp) lc[threadIdx.x] = cin[S.lane0 + threadIdx.x];
    __syncthreads();
}

// The first block of the system's range (b0 for K1, c0 for K2/K3) writes
// its lanes' tokens.
__device__ __forceinline__ void write_lanes(BCtrl* __restrict__ cout, const BSys& S, bool first,
                                            const BCtrl* lc)
{
    if (first && threadIdx.x < S.kp) cout[S.lane0 + threadIdx.x] = lc[threadIdx.x];
}

// ---------------------------------------------------------------- set-up
// r = b - A x ; p = jac * r ; partials of r.z, r.r, b.b per lane.
__global__ __launch_bounds__(kBlock) void k_bresidual(const BSys* __restrict__ sys,
                                                      const BBlk* __restrict__ tbl,
                                                      const double* __restrict__ x,
                                                      const double* __restrict__ bv,
                                                      double* __restrict__ r, double* __restrict__ p,
                                                      double* __restrict__ part, int kpmax,
                                                      int G, BCtrl* __restrict__ c0)
{
    __shared__ BStage sm;
    __shared__ double red[4 * 64];
    const int b = blockIdx.x;  // hardware order: a block range has equal shares of every XCD
    const BBlk blk = tbl[b];
    if (blk.sys < 0) return;
    const BSys S = sys[blk.sys];
    const double* xs = x + S.voff;
    const int j = threadIdx.x & (S.kp - 1);
    double arz = 0.0, arr = 0.0, abb = 0.0;
    for (int t = blk.t0; t < blk.t1; ++t) {
        int row;
        const double s = btile(S, t * (kBlock >> S.ksh), xs, sm, true, row);
        if (row < S.n) {
            const size_t e = S.voff + ((size_t)row << S.ksh) + j;
            const double bi = bv[e];
            const double ri = bi - (0.0 + s);
            const double zi = ri * S.jac[row];
            r[e] = ri;
            p[e] = zi;
            double rz = ri * zi;
            arz = arz + rz;
            double rr = ri * ri;
            arr = arr + rr;
            double bb = bi * bi;
            abb = abb + bb;
        }
    }
    const double s1 = lane_block_sum(arz, S.kp, red);
    const double s2 = lane_block_sum(arr, S.kp, red);
    const double s3 = lane_block_sum(abb, S.kp, red);
    if (threadIdx.x < S.kp) {
        part[(size_t)b * kpmax + threadIdx.x] = s1;
        part[((size_t)G + b) * kpmax + threadIdx.x] = s2;
        part[((size_t)2 * G + b) * kpmax + threadIdx.x] = s3;
    }
    if (b == S.b0 && threadIdx.x < S.kp) {
        BCtrl c{};
        c.run = 1;
        c.reduce = kBReduceSetup;
        c0[S.lane0 + threadIdx.x] = c;
    }
}

// ---------------------------------------------------------------- K1
// v = A p ; partial of v.p per lane.  Folds the set-up partials first.
__global__ __launch_bounds__(kBlock) void k_bspmv(const BSys* __restrict__ sys,
                                                  const BBlk* __restrict__ tbl,
                                                  const double* __restrict__ p,
                                                  double* __restrict__ v, double* __restrict__ part,
                                                  const double* __restrict__ rpart, int kpmax, int G,
                                                  const BCtrl* __restrict__ cin,
                                                  BCtrl* __restrict__ cout, BPrm P)
{
    __shared__ CsrStage<double> sm;
    __shared__ double red[4 * 64];
    __shared__ double s1[kMaxRhs], s2[kMaxRhs], s3[kMaxRhs];
    __shared__ BCtrl lc[kMaxRhs];
    __shared__ int any;
    const int b = blockIdx.x;  // hardware order: a block range has equal shares of every XCD
    const BBlk blk = tbl[b];
    if (blk.sys < 0) return;
    const BSys S = sys[blk.sys];
    read_lanes(cin, S, lc);
    if (lc[0].reduce == kBReduceSetup) {  // the same for every lane
        lane_partials(rpart, kpmax, S.b0, S.b1, S.kp, S.ksh, s1, red);
        lane_partials(rpart + (size_t)G * kpmax, kpmax, S.b0, S.b1, S.kp, S.ksh, s2, red);
        lane_partials(rpart + (size_t)2 * G * kpmax, kpmax, S.b0, S.b1, S.kp, S.ksh, s3, red);
        if (threadIdx.x < S.kp) {
            BCtrl& c = lc[threadIdx.x];
            const double rn = sqrt(s2[threadIdx.x]);
            c.bnorm = sqrt(s3[threadIdx.x]);
            c.norm0 = rn;
            c.norm2 = rn;
            c.alpha = s1[threadIdx.x];
            c.eps = b_eps(P, c.bnorm, rn);
            c.iter = 0;
            c.reduce = kBReduceNone;
            if (!isfinite(s3[threadIdx.x]) && c.status == 0) c.status = CEL_ERR_NONFINITE;
            c.run = (c.status == 0 && c.iter < P.iter_max && b_alpha_ok(P, c.alpha) &&
                     c.norm2 > c.eps)
                        ? 1
                        : 0;
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        int a = 0;
        for (int l = 0; l < S.kp; ++l) a |= lc[l].run;
        any = a;
    }
    __syncthreads();
    if (!any) {
        write_lanes(cout, S, b == S.b0, lc);
        return;
    }
    const double* ps = p + S.voff;
    double* vs = v + S.voff;
    double acc = 0.0;
    if (S.kp == 1) {
        // one right-hand side: the pipelined LDS tile of the single-system K1
        // (256 rows per tile, the next tile's loads in flight meanwhile)
        CsrTileStream<double> cs;
        const int t0 = blk.t0, t1 = blk.t1;
        if (t0 < t1) {
            cs.start(S.n, t0, t0 + 1 < t1 ? t0 + 1 : -1, S.xadj, S.cols, S.vals, sm);
            cs.stage(sm);
            __syncthreads();
        }
        for (int t = t0; t < t1; ++t) {
            const int nxt = t + 1 < t1 ? t + 1 : -1;
            const int aft = t + 2 < t1 ? t + 2 : -1;
            const int row = t * kTileRows + threadIdx.x;
            double pr = 0.0;
            double sum = cs.template rows<double>(S.n, t, nxt, aft, S.xadj, S.cols, S.vals, ps, sm,
                                                  row, &pr);
            __syncthreads();
            if (nxt >= 0) {
                cs.stage(sm);
                __syncthreads();
            }
            if (row < S.n) {
                const double vi = 0.0 + sum;
                vs[row] = vi;
                double prod = vi * pr;
                acc = acc + prod;
            }
        }
    } else {
        // kp right-hand sides, two adjacent lanes per thread: the kp/2
        // threads of a row gather x[col][0..kp) contiguously, 16 bytes each
        // (one load instruction serves two lanes).  The block's steps go in
        // chunks of 256 rows whose row pointers and entries are staged in
        // LDS with coalesced loads first, so a row's only global loads are
        // its gathers (a row read straight from HBM waits on its row bounds,
        // then its entries, then the gathers).  Every lane sums its row in
        // entry order; the rows of a thread are added to its partials in
        // row order.  A chunk whose entries do not fit the stage reads them
        // from HBM.
        __shared__ int sx[kTileRows + 1];
        const int kh = S.ksh - 1;                              // kp/2 = 2^kh threads per row
        const int R2 = kBlock >> kh;                           // rows per step
        const int my = threadIdx.x >> kh;
        const int jl = 2 * (threadIdx.x & ((S.kp >> 1) - 1));  // the thread's first lane
        const bool on2 = lc[jl].run != 0 || lc[jl + 1].run != 0;
        // (voff and kp are even, so the pairs are 16-byte aligned)
        const double* pl = ps + jl;
        double* vl = vs + jl;
        double acc1 = 0.0;
        // Chunk order.  A system with a band period P (its bandwidth in
        // chunks, P >= 8) on at least 8 blocks: XCD x takes the chunks at
        // positions [P x / 8, P (x + 1) / 8) of every period, its blocks of
        // the system dealing them round-robin, so rows a band apart --
        // which gather the same x rows -- run on the same XCD and share its
        // L2, and the XCD's blocks sweep its chunks together.  Otherwise
        // chunks go round-robin over the system's blocks (chunk c to block
        // b0 + c mod nb).
        const int nb = S.b1 - S.b0;
        const int first = b - S.b0;  // the table keeps b in [b0, b1)
        const int nchunks = (first >= 0 && first < nb) ? (S.ntiles + S.kp - 1) / S.kp : 0;
        // (hardware block b runs on XCD b mod 8; the system's blocks on XCD x
        // are b0 <= b < b1 with b = x mod 8)
        const bool band = S.period >= kXcds && nb >= kXcds;
        const int xcd = b % kXcds;
        const int bx0 = S.b0 + ((xcd - S.b0 % kXcds) + kXcds) % kXcds;  // first of them
        const int gx = band ? (S.b1 - 1 - bx0) / kXcds + 1 : 1;       // their number
        const int jb = band ? (b - bx0) / kXcds : 0;
        const int p0 = band ? S.period * xcd / kXcds : 0;
        const int pc = band ? S.period * (xcd + 1) / kXcds - p0 : 1;  // positions of this XCD
        for (int q = 0;; ++q) {
            int c;
            if (band) {
                const long long rank = (long long)jb + (long long)q * gx;
                c = (int)((rank / pc) * S.period + p0 + rank % pc);
            } else {
                c = first + q * nb;
            }
            if (c >= nchunks) break;  // uniform; chunk indices grow with q
            // chunk c = tiles [c kp, c kp + kp) of 256/kp rows = rows [r0, r1)
            const int r0 = c * kTileRows;
            if (r0 >= S.n) break;  // uniform
            const int r1 = min(r0 + kTileRows, S.n);
            const int e0 = S.xadj[r0];
            const int cnt = S.xadj[r1] - e0;
            const bool staged = cnt <= kStageCap;  // uniform across the block
            __syncthreads();  // the previous chunk's stage is consumed
            // the matrix is read once per product: non-temporal loads keep
            // it from evicting the gathered x lines from L2
            // (all of a thread's stage loads are issued before any is
            // stored: a loop that stored each before loading the next paid
            // one memory round trip per 256 entries)
            {
                constexpr int kSx = (kTileRows + 1 + kBlock - 1) / kBlock;
                int xa[kSx];
#pragma unroll
                for (int i = 0; i < kSx; ++i) {
                    const int q = threadIdx.x + i * kBlock;
                    xa[i] = ld_stream(S.xadj + r0 + min(q, r1 - r0));
                }
#pragma unroll
                for (int i = 0; i < kSx; ++i) {
                    const int q = threadIdx.x + i * kBlock;
                    if (q <= r1 - r0) sx[q] = xa[i] - e0;
                }
            }
            if (staged && cnt > 0) {
                constexpr int kPerT = kStageCap / kBlock;
                double va[kPerT];
                int ca[kPerT];
#pragma unroll
                for (int i = 0; i < kPerT; ++i) {
                    const int k = min((int)threadIdx.x + i * kBlock, cnt - 1);
                    va[i] = ld_stream(S.vals + e0 + k);
                    ca[i] = ld_stream(S.cols + e0 + k);
                }
#pragma unroll
                for (int i = 0; i < kPerT; ++i) {
                    const int k = (int)threadIdx.x + i * kBlock;
                    if (k < cnt) {
                        sm.val[k] = va[i];
                        sm.col[k] = ca[i];
                    }
                }
            }
            __syncthreads();
            if (!on2) continue;
            const int nst = (r1 - r0 + R2 - 1) / R2;  // steps of this chunk
            if (staged) {
                // kPairSteps steps at a time: the gathers of all their rows
                // (the first 8 entries of each, branch-free: an entry past
                // the row reads x row 0 and is not added) and the rows' own
                // p values are in flight together; the sums then run in
                // entry order, rows longer than 8 continuing one entry at a
                // time
                for (int s0 = 0; s0 < nst; s0 += kPairSteps) {
                    int rw[kPairSteps], lo[kPairSteps], len[kPairSteps];
                    Dbl2 xv[kPairSteps][8], pr[kPairSteps];
#pragma unroll
                    for (int u = 0; u < kPairSteps; ++u) {
                        rw[u] = r0 + (s0 + u) * R2 + my;
                        const bool ok = s0 + u < nst && rw[u] < r1;
                        const int li = ok ? rw[u] - r0 : 0;
                        lo[u] = ok ? sx[li] : 0;
                        len[u] = ok ? sx[li + 1] - lo[u] : 0;
#pragma unroll
                        for (int i = 0; i < 8; ++i) {
                            const int c = sm.col[min(lo[u] + i, kStageCap - 1)];
                            xv[u][i] = *(const Dbl2*)(pl + ((size_t)(i < len[u] ? c : 0) << S.ksh));
                        }
                        pr[u] = *(const Dbl2*)(pl + ((size_t)(ok ? rw[u] : 0) << S.ksh));
                    }
#pragma unroll
                    for (int u = 0; u < kPairSteps; ++u) {
                        if (s0 + u < nst && rw[u] < r1) {
                            double sum0 = 0.0, sum1 = 0.0;
#pragma unroll
                            for (int i = 0; i < 8; ++i) {
                                const double a = sm.val[min(lo[u] + i, kStageCap - 1)];
                                double prod0 = a * xv[u][i].x;
                                double prod1 = a * xv[u][i].y;
                                const double t0 = sum0 + prod0, t1 = sum1 + prod1;
                                sum0 = i < len[u] ? t0 : sum0;
                                sum1 = i < len[u] ? t1 : sum1;
                            }
                            for (int k = lo[u] + 8; k < lo[u] + len[u]; ++k) {
                                const double a = sm.val[k];
                                const Dbl2 xk = *(const Dbl2*)(pl + ((size_t)sm.col[k] << S.ksh));
                                double prod0 = a * xk.x;
                                double prod1 = a * xk.y;
                                sum0 = sum0 + prod0;
                                sum1 = sum1 + prod1;
                            }
                            Dbl2 vi;
                            vi.x = 0.0 + sum0;
                            vi.y = 0.0 + sum1;
                            // read by the next kernel only
                            __builtin_nontemporal_store(vi, (Dbl2*)(vl + ((size_t)rw[u] << S.ksh)));
                            double d0 = vi.x * pr[u].x;
                            acc = acc + d0;
                            double d1 = vi.y * pr[u].y;
                            acc1 = acc1 + d1;
                        }
                    }
                }
                continue;
            }
            for (int st = 0; st < nst; ++st) {
                const int row = r0 + st * R2 + my;
                if (row < r1) {
                    const int lo = sx[row - r0], hi = sx[row - r0 + 1];
                    double sum0 = 0.0, sum1 = 0.0;  // a chunk too large for the stage: entries from HBM
                    for (int k = lo; k < hi; ++k) {
                        const double a = ld_stream(S.vals + e0 + k);
                        const Dbl2 xk =
                            *(const Dbl2*)(pl + ((size_t)ld_stream(S.cols + e0 + k) << S.ksh));
                        double prod0 = a * xk.x;
                        double prod1 = a * xk.y;
                        sum0 = sum0 + prod0;
                        sum1 = sum1 + prod1;
                    }
                    Dbl2 vi;
                    vi.x = 0.0 + sum0;
                    vi.y = 0.0 + sum1;
                    __builtin_nontemporal_store(vi, (Dbl2*)(vl + ((size_t)row << S.ksh)));
                    const Dbl2 pv = *(const Dbl2*)(pl + ((size_t)row << S.ksh));
                    double d0 = vi.x * pv.x;
                    acc = acc + d0;
                    double d1 = vi.y * pv.y;
                    acc1 = acc1 + d1;
                }
            }
        }
        const double bs = lane_pair_block_sum(acc, acc1, S.kp, red);
        if (threadIdx.x < S.kp) part[(size_t)b * kpmax + threadIdx.x] = bs;
        write_lanes(cout, S, b == S.b0, lc);
        return;
    }
    const double bs = lane_block_sum(acc, S.kp, red);
    if (threadIdx.x < S.kp) part[(size_t)b * kpmax + threadIdx.x] = bs;
    write_lanes(cout, S, b == S.b0, lc);
}

// ---------------------------------------------------------------- K2
// sigma = alpha / sum(v.p) ; r = r + (-sigma) v ; z = jac r ; partials r.z, r.r.
__global__ __launch_bounds__(kBlock) void k_bupdate_r(const BSys* __restrict__ sys,
                                                      const BBlk* __restrict__ tbl,
                                                      double* __restrict__ r,
                                                      const double* __restrict__ v,
                                                      double* __restrict__ part,
                                                      const double* __restrict__ vpart, int kpmax,
                                                      int G, const BCtrl* __restrict__ cin,
                                                      BCtrl* __restrict__ cout)
{
    __shared__ double red[4 * 64];
    __shared__ double s1[kMaxRhs];
    __shared__ BCtrl lc[kMaxRhs];
    __shared__ int any;
    const int b = blockIdx.x;  // hardware order: a block range has equal shares of every XCD
    const BBlk blk = tbl[b];
    if (blk.sys < 0) return;
    const BSys S = sys[blk.sys];
    read_lanes(cin, S, lc);
    if (threadIdx.x == 0) {
        int a = 0;
        for (int l = 0; l < S.kp; ++l) a |= lc[l].run;
        any = a;
    }
    __syncthreads();
    if (!any) {
        write_lanes(cout, S, b == S.c0, lc);
        return;
    }
    // K1's partials: its blocks [b0, b1) of the K1 table
    lane_partials(vpart, kpmax, S.b0, S.b1, S.kp, S.ksh, s1, red);
    if (threadIdx.x < S.kp) {
        BCtrl& c = lc[threadIdx.x];
        if (c.run) {
            const double vp = s1[threadIdx.x];
            const double sigma = c.alpha / vp;
            if (!(vp > 0.0) || !isfinite(sigma)) {
                c.status = CEL_ERR_BREAKDOWN;
                c.run = 0;
            }
            c.sigma = sigma;
        }
    }
    __syncthreads();
    const int j = threadIdx.x & (S.kp - 1);
    const bool on = lc[j].run != 0;
    const double ms = -lc[j].sigma;
    const long long ne = (long long)S.n << S.ksh;
    double arz = 0.0, arr = 0.0;
    if (on) {
        const long long lo = blk.u0 * kBlock, hi = min(blk.u1 * (long long)kBlock, ne);
        // kUnroll elements per step (e, e+256, ...): their loads are in
        // flight together; the sums still run in ascending e
        long long e = lo + threadIdx.x;
        for (; e + (kUnroll - 1) * kBlock < hi; e += kUnroll * kBlock) {
            double vv[kUnroll], rv[kUnroll], jv[kUnroll];
#pragma unroll
            for (int u = 0; u < kUnroll; ++u) {
                const long long eu = e + u * kBlock;
                vv[u] = v[S.voff + eu];
                rv[u] = r[S.voff + eu];
                jv[u] = S.jac[eu >> S.ksh];
            }
#pragma unroll
            for (int u = 0; u < kUnroll; ++u) {
                double t = ms * vv[u];
                const double ri = rv[u] + t;
                r[S.voff + e + u * kBlock] = ri;
                const double zi = ri * jv[u];
                double rz = ri * zi;
                arz = arz + rz;
                double rr = ri * ri;
                arr = arr + rr;
            }
        }
        for (; e < hi; e += kBlock) {
            const size_t g = S.voff + e;
            double t = ms * v[g];
            const double ri = r[g] + t;
            r[g] = ri;
            const double zi = ri * S.jac[e >> S.ksh];
            double rz = ri * zi;
            arz = arz + rz;
            double rr = ri * ri;
            arr = arr + rr;
        }
    }
    const double a1 = lane_block_sum(arz, S.kp, red);
    const double a2 = lane_block_sum(arr, S.kp, red);
    if (threadIdx.x < S.kp) {
        part[(size_t)b * kpmax + threadIdx.x] = a1;
        part[((size_t)G + b) * kpmax + threadIdx.x] = a2;
    }
    write_lanes(cout, S, b == S.c0, lc);
}

// ---------------------------------------------------------------- K3
// beta = sum(r.z)/alpha ; x = x + sigma p ; p = beta p + jac r ; norm, k++,
// stop decision.
__global__ __launch_bounds__(kBlock) void k_bupdate_xp(const BSys* __restrict__ sys,
                                                       const BBlk* __restrict__ tbl,
                                                       double* __restrict__ x,
                                                       double* __restrict__ p,
                                                       const double* __restrict__ r,
                                                       const double* __restrict__ rpart, int kpmax,
                                                       int G, const BCtrl* __restrict__ cin,
                                                       BCtrl* __restrict__ cout, BPrm P)
{
    __shared__ double red[4 * 64];
    __shared__ double s1[kMaxRhs], s2[kMaxRhs], sb[kMaxRhs];
    __shared__ BCtrl lc[kMaxRhs];
    __shared__ int any;
    const int b = blockIdx.x;  // hardware order: a block range has equal shares of every XCD
    const BBlk blk = tbl[b];
    if (blk.sys < 0) return;
    const BSys S = sys[blk.sys];
    read_lanes(cin, S, lc);
    if (threadIdx.x == 0) {
        int a = 0;
        for (int l = 0; l < S.kp; ++l) a |= lc[l].run;
        any = a;
    }
    __syncthreads();
    if (!any) {
        write_lanes(cout, S, b == S.c0, lc);
        return;
    }
    // K2's partials: its blocks [c0, c1) of its table (G of them)
    lane_partials(rpart, kpmax, S.c0, S.c1, S.kp, S.ksh, s1, red);
    lane_partials(rpart + (size_t)G * kpmax, kpmax, S.c0, S.c1, S.kp, S.ksh, s2, red);
    if (threadIdx.x < S.kp) sb[threadIdx.x] = s1[threadIdx.x] / lc[threadIdx.x].alpha;
    __syncthreads();
    const int j = threadIdx.x & (S.kp - 1);
    const bool on = lc[j].run != 0;
    const double sigma = lc[j].sigma, beta = sb[j];
    const long long ne = (long long)S.n << S.ksh;
    if (on) {
        const long long lo = blk.u0 * kBlock, hi = min(blk.u1 * (long long)kBlock, ne);
        long long e = lo + threadIdx.x;
        for (; e + (kUnroll - 1) * kBlock < hi; e += kUnroll * kBlock) {
            double pv[kUnroll], xv[kUnroll], rv[kUnroll], jv[kUnroll];
#pragma unroll
            for (int u = 0; u < kUnroll; ++u) {
                const long long eu = e + u * kBlock;
                pv[u] = p[S.voff + eu];
                xv[u] = x[S.voff + eu];
                rv[u] = r[S.voff + eu];
                jv[u] = S.jac[eu >> S.ksh];
            }
#pragma unroll
            for (int u = 0; u < kUnroll; ++u) {
                const size_t g = S.voff + e + u * kBlock;
                double t = sigma * pv[u];
                x[g] = xv[u] + t;
                const double zi = rv[u] * jv[u];
                double bp = beta * pv[u];
                p[g] = bp + zi;
            }
        }
        for (; e < hi; e += kBlock) {
            const size_t g = S.voff + e;
            const double pi = p[g];
            double t = sigma * pi;
            x[g] = x[g] + t;
            const double zi = r[g] * S.jac[e >> S.ksh];
            double bp = beta * pi;
            p[g] = bp + zi;
        }
    }
    if (threadIdx.x < S.kp) {
        BCtrl& c = lc[threadIdx.x];
        if (c.run) {
            c.alpha = s1[threadIdx.x];
            c.norm2 = sqrt(s2[threadIdx.x]);
            c.iter = c.iter + 1;
            c.run = (c.iter < P.iter_max && b_alpha_ok(P, c.alpha) && c.norm2 > c.eps) ? 1 : 0;
        }
    }
    __syncthreads();
    write_lanes(cout, S, b == S.c0, lc);
}

// ---------------------------------------------------------------- layout kernels
// dst[voff + row*kp + j] = src[j*n + row] for j < k, 0 for padding lanes.
__global__ void k_binterleave(const double* __restrict__ src, double* __restrict__ dst, int n,
                              int k, int kp, int ksh, double fill_pad)
{
    const long long ne = (long long)n << ksh;
    for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < ne;
         e += (long long)gridDim.x * blockDim.x) {
        const int j = (int)(e & (kp - 1));
        const long long row = e >> ksh;
        dst[e] = j < k ? src[(long long)j * n + row] : fill_pad;
    }
}

__global__ void k_bdeinterleave(const double* __restrict__ src, double* __restrict__ dst, int n,
                                int k, int ksh)
{
    const long long ne = (long long)n * k;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < ne;
         i += (long long)gridDim.x * blockDim.x) {
        const long long j = i / n, row = i - j * n;
        dst[i] = src[(row << ksh) + j];
    }
}

// dst[row*kp + j] = v for the k real lanes, 0 for padding lanes.
__global__ void k_bfill(double* __restrict__ dst, int n, int k, int kp, int ksh, double v)
{
    const long long ne = (long long)n << ksh;
    for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < ne;
         e += (long long)gridDim.x * blockDim.x)
        dst[e] = (int)(e & (kp - 1)) < k ? v : 0.0;
}

// ---------------------------------------------------------------- host side
struct Tuning {
    int grid;     // blocks of K1 and the set-up
    int chunk;    // iterations per graph replay
    int weight;   // 0: blocks by entries x lanes, 1: by rows x lanes
    int rebalance;
    int grid23 = 0;  // blocks of K2 and K3 (0: grid)
    int g23() const { return grid23 > 0 ? grid23 : grid; }
};

struct Batch {
    std::vector<Mat*> mats;
    std::vector<int> k, kp, ksh, lane0;
    std::vector<long long> voff;
    long long total = 0;  // vector elements, each system's padded to kVecAlign
    long long elems = 0;  // vector elements of the systems
    int lanes = 0;        // padded lanes
    int kpmax = 1;
    DevBuf<BSys> d_sys;
    DevBuf<BBlk> d_tbl, d_tbl2;  // K1 and set-up; K2 and K3
    DevBuf<double> r, p, v, x, b, part1, part2, rpart;
    DevBuf<BCtrl> ctrl;  // 3 slots of `lanes` tokens
    std::vector<BSys> h_sys;
    bool tuned = false;
    Tuning tune{0, 16, 0, 1};
    double tuned_ms_per_iter = 0.0;
};

static int grid_max()
{
    int g = resident_grid((const void*)k_bspmv, 1 << 30);
    g = std::min(g, resident_grid((const void*)k_bresidual, 1 << 30));
    g = std::min(g, resident_grid((const void*)k_bupdate_r, 1 << 30));
    g = std::min(g, resident_grid((const void*)k_bupdate_xp, 1 << 30));
    return g;
}

// Deals G blocks to the systems with active[s] (at least one each, when
// there are enough blocks), proportional to their work, and fills the work
// table.  Blocks of the table are indexed by logical block.
// Deals G blocks to the active systems (ranges b0/b1 of the K1 table, or
// c0/c1 of the K2/K3 table when k23).
static int build_table(Batch& B, int G, int weight, const std::vector<char>& active,
                       std::vector<BBlk>& tbl, bool k23)
{
    const int S = (int)B.mats.size();
    std::vector<double> w(S, 0.0);
    double W = 0.0;
    int nact = 0;
    for (int s = 0; s < S; ++s) {
        if (!active[s]) continue;
        const Mat& A = *B.mats[s];
        w[s] = (weight == 0 ? (double)A.nnz + A.n : (double)A.n) * B.kp[s];
        W += w[s];
        ++nact;
    }
    tbl.assign(G, BBlk{-1, 0, 0, 0, 0});
    for (int s = 0; s < S; ++s) {
        if (k23)
            B.h_sys[s].c0 = B.h_sys[s].c1 = 0;
        else
            B.h_sys[s].b0 = B.h_sys[s].b1 = 0;
    }
    if (nact == 0) return CEL_OK;
    if (nact > G) return fail(CEL_ERR_ARG, "more systems than resident blocks in one batch");
    // largest-remainder apportionment with a floor of one block
    std::vector<int> nb(S, 0);
    int left = G - nact;
    std::vector<std::pair<double, int>> rem;
    for (int s = 0; s < S; ++s) {
        if (!active[s]) continue;
        const double share = (double)left * w[s] / W;
        const int tiles = B.h_sys[s].ntiles;
        int extra = (int)std::floor(share);
        extra = std::min(extra, std::max(0, tiles - 1));
        nb[s] = 1 + extra;
        rem.push_back({share - std::floor(share), s});
    }
    int used = std::accumulate(nb.begin(), nb.end(), 0);
    std::sort(rem.begin(), rem.end(), [](auto& a, auto& b) {
        return a.first != b.first ? a.first > b.first : a.second < b.second;
    });
    for (auto& pr : rem) {
        if (used >= G) break;
        const int s = pr.second;
        if (nb[s] < B.h_sys[s].ntiles) {
            ++nb[s];
            ++used;
        }
    }
    int cur = 0;
    for (int s = 0; s < S; ++s) {
        if (!active[s]) continue;
        BSys& H = B.h_sys[s];
        (k23 ? H.c0 : H.b0) = cur;
        (k23 ? H.c1 : H.b1) = cur + nb[s];
        const long long units = (((long long)H.n << H.ksh) + kBlock - 1) / kBlock;
        for (int i = 0; i < nb[s]; ++i) {
            BBlk& e = tbl[cur + i];
            e.sys = s;
            e.t0 = (int)((long long)H.ntiles * i / nb[s]);
            e.t1 = (int)((long long)H.ntiles * (i + 1) / nb[s]);
            e.u0 = units * i / nb[s];
            e.u1 = units * (i + 1) / nb[s];
        }
        cur += nb[s];
    }
    return CEL_OK;
}

static int upload_table(Batch& B, const std::vector<BBlk>& tbl, const std::vector<BBlk>& tbl2,
                        hipStream_t s)
{
    CEL_TRY(B.d_tbl.alloc(tbl.size()));
    CEL_TRY(B.d_tbl2.alloc(tbl2.size()));
    CEL_HIP(hipMemcpyAsync(B.d_tbl.p, tbl.data(), sizeof(BBlk) * tbl.size(), hipMemcpyHostToDevice, s));
    CEL_HIP(hipMemcpyAsync(B.d_tbl2.p, tbl2.data(), sizeof(BBlk) * tbl2.size(), hipMemcpyHostToDevice,
                           s));
    CEL_HIP(hipMemcpyAsync(B.d_sys.p, B.h_sys.data(), sizeof(BSys) * B.h_sys.size(),
                           hipMemcpyHostToDevice, s));
    // the copies read pageable host memory: finish them before the host
    // vectors change
    CEL_HIP(hipStreamSynchronize(s));
    return CEL_OK;
}

struct RunResult {
    std::vector<BCtrl> lanes;
    float ms = 0.f;
    int chunks = 0;
};

// One batched solve with tuning T from the current x (x0 already in B.x).
static int run_batch(Batch& B, const Tuning& T, const cel_pcg_params& prm, int iter_cap,
                     RunResult& out)
{
    hipStream_t s = rt().stream;
    const int G = T.grid;
    const int G2 = T.g23();
    const int S = (int)B.mats.size();
    CEL_TRY(B.part1.alloc((size_t)G * B.kpmax));
    CEL_TRY(B.part2.alloc((size_t)2 * G2 * B.kpmax));
    CEL_TRY(B.rpart.alloc((size_t)3 * G * B.kpmax));
    CEL_TRY(B.ctrl.alloc((size_t)3 * B.lanes));
    std::vector<char> active(S, 1);
    std::vector<BBlk> tbl, tbl2;
    CEL_TRY(build_table(B, G, T.weight, active, tbl, false));
    CEL_TRY(build_table(B, G2, T.weight, active, tbl2, true));
    CEL_TRY(upload_table(B, tbl, tbl2, s));

    BPrm P{};
    P.tol = prm.tol;
    P.small = 1.1 * 2.220446049250313e-16;
    P.tol_mode = prm.tol_mode;
    P.iter_max = iter_cap;
    P.alpha_stop = prm.alpha_stop;
    P.kpmax = B.kpmax;
    BCtrl* C0 = B.ctrl.p;
    BCtrl* C1 = B.ctrl.p + B.lanes;
    BCtrl* C2 = B.ctrl.p + 2 * B.lanes;

    StreamScope R(s);
    hipEvent_t e0, e1;
    CEL_TRY(R.event(e0));
    CEL_TRY(R.event(e1));
    CEL_HIP(hipEventRecord(e0, s));
    hipLaunchKernelGGL(k_bresidual, dim3(G), dim3(kBlock), 0, s, B.d_sys.p, B.d_tbl.p, B.x.p, B.b.p,
                       B.r.p, B.p.p, B.rpart.p, B.kpmax, G, C0);
    CEL_HIP(hipGetLastError());

    auto enqueue_iter = [&]() -> int {
        hipLaunchKernelGGL(k_bspmv, dim3(G), dim3(kBlock), 0, s, B.d_sys.p, B.d_tbl.p, B.p.p, B.v.p,
                           B.part1.p, B.rpart.p, B.kpmax, G, C0, C1, P);
        hipLaunchKernelGGL(k_bupdate_r, dim3(G2), dim3(kBlock), 0, s, B.d_sys.p, B.d_tbl2.p, B.r.p,
                           B.v.p, B.part2.p, B.part1.p, B.kpmax, G2, C1, C2);
        hipLaunchKernelGGL(k_bupdate_xp, dim3(G2), dim3(kBlock), 0, s, B.d_sys.p, B.d_tbl2.p, B.x.p,
                           B.p.p, B.r.p, B.part2.p, B.kpmax, G2, C2, C0, P);
        CEL_HIP(hipGetLastError());
        return CEL_OK;
    };

    // chunks of T.chunk iterations replayed from one captured graph; the
    // host checks the lanes one chunk behind and re-deals the blocks of
    // stopped systems
    CEL_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    for (int i = 0; i < T.chunk; ++i) {
        int rc = enqueue_iter();
        if (rc < 0) {
            (void)hipStreamEndCapture(s, &R.graph);
            return rc;
        }
    }
    CEL_HIP(hipStreamEndCapture(s, &R.graph));
    CEL_HIP(hipGraphInstantiate(&R.exec, R.graph, nullptr, nullptr, 0));
    hipGraphExec_t gexec = R.exec;

    BCtrl* hc = nullptr;
    CEL_TRY(R.host_alloc(hc, 2 * (size_t)B.lanes));
    hipEvent_t cev[2];
    CEL_TRY(R.event(cev[0], hipEventDisableTiming));
    CEL_TRY(R.event(cev[1], hipEventDisableTiming));
    const long long budget = (long long)iter_cap + 1;
    long long enq = 0;
    int k = 0;
    int rc = CEL_OK;
    while (enq < budget) {
        CEL_HIP(hipGraphLaunch(gexec, s));
        enq += T.chunk;
        BCtrl* slot = hc + (size_t)(k & 1) * B.lanes;
        CEL_HIP(hipMemcpyAsync(slot, C0, sizeof(BCtrl) * B.lanes, hipMemcpyDeviceToHost, s));
        CEL_HIP(hipEventRecord(cev[k & 1], s));
        if (k > 0) {
            CEL_HIP(hipEventSynchronize(cev[(k - 1) & 1]));
            const BCtrl* prev = hc + (size_t)((k - 1) & 1) * B.lanes;
            std::vector<char> act(S, 0);
            int nrun = 0;
            for (int q = 0; q < S; ++q)
                for (int l = 0; l < B.kp[q]; ++l)
                    if (prev[B.lane0[q] + l].run) act[q] = 1;
            for (int q = 0; q < S; ++q) nrun += act[q];
            if (nrun == 0) break;
            if (T.rebalance && act != active) {
                // every lane of a dropped system has stored its final token
                // in slot 0; the chunk in flight does not touch it again
                CEL_HIP(hipEventSynchronize(cev[k & 1]));
                active = act;
                rc = build_table(B, G, T.weight, active, tbl, false);
                if (rc >= 0) rc = build_table(B, G2, T.weight, active, tbl2, true);
                if (rc < 0) break;
                rc = upload_table(B, tbl, tbl2, s);
                if (rc < 0) break;
            }
        }
        ++k;
    }
    CEL_HIP(hipEventRecord(e1, s));
    CEL_HIP(hipStreamSynchronize(s));
    out.chunks = k + 1;
    CEL_HIP(hipEventElapsedTime(&out.ms, e0, e1));
    out.lanes.resize(B.lanes);
    CEL_HIP(hipMemcpy(out.lanes.data(), C0, sizeof(BCtrl) * B.lanes, hipMemcpyDeviceToHost));
    return rc;
}

static int load_rhs(Batch& B, const double* const* b, int memspace)
{
    hipStream_t s = rt().stream;
    const int S = (int)B.mats.size();
    DevBuf<double> tmp;
    for (int q = 0; q < S; ++q) {
        const int n = B.mats[q]->n;
        const double* src = b[q];
        if (!src) return fail(CEL_ERR_ARG, "null right-hand side");
        if (memspace == CEL_MEM_HOST) {
            CEL_TRY(tmp.alloc((size_t)n * B.k[q]));
            CEL_HIP(hipMemcpyAsync(tmp.p, src, sizeof(double) * n * B.k[q], hipMemcpyHostToDevice, s));
            CEL_HIP(hipStreamSynchronize(s));
            src = tmp.p;
        }
        hipLaunchKernelGGL(k_binterleave, dim3(1024), dim3(kBlock), 0, s, src, B.b.p + B.voff[q], n,
                           B.k[q], B.kp[q], B.ksh[q], 0.0);
        CEL_HIP(hipGetLastError());
    }
    CEL_HIP(hipStreamSynchronize(s));
    return CEL_OK;
}

static int init_x(Batch& B, double* const* x, const cel_pcg_params& prm)
{
    hipStream_t s = rt().stream;
    if (prm.x0_mode == CEL_X0_USER) {
        DevBuf<double> tmp;
        for (size_t q = 0; q < B.mats.size(); ++q) {
            const int n = B.mats[q]->n;
            const double* src = x[q];
            if (prm.memspace == CEL_MEM_HOST) {
                CEL_TRY(tmp.alloc((size_t)n * B.k[q]));
                CEL_HIP(hipMemcpyAsync(tmp.p, src, sizeof(double) * n * B.k[q], hipMemcpyHostToDevice, s));
                CEL_HIP(hipStreamSynchronize(s));
                src = tmp.p;
            }
            hipLaunchKernelGGL(k_binterleave, dim3(1024), dim3(kBlock), 0, s, src, B.x.p + B.voff[q], n,
                               B.k[q], B.kp[q], B.ksh[q], 0.0);
        }
    } else {
        for (size_t q = 0; q < B.mats.size(); ++q)
            hipLaunchKernelGGL(k_bfill, dim3(1024), dim3(kBlock), 0, s, B.x.p + B.voff[q],
                               B.mats[q]->n, B.k[q], B.kp[q], B.ksh[q],
                               prm.x0_mode == CEL_X0_ONE ? 1.0 : 0.0);
    }
    CEL_HIP(hipGetLastError());
    CEL_HIP(hipStreamSynchronize(s));
    return CEL_OK;
}

}  // namespace cel

using namespace cel;

struct cel_gpu_batch {
    Batch b;
};

// Bandwidth max |col - row| of a square block (0-based rows and columns):
// per-block maxima, the host takes the largest.
__global__ __launch_bounds__(kBlock) void k_bandwidth(int n, const int* __restrict__ xadj,
                                                      const int* __restrict__ cols,
                                                      int* __restrict__ out)
{
    __shared__ int wm[kWaves];
    int m = 0;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (long long)gridDim.x * blockDim.x)
        for (int k = xadj[i]; k < xadj[i + 1]; ++k) m = max(m, abs(cols[k] - (int)i));
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) m = max(m, __shfl_xor(m, off, 64));
    if ((threadIdx.x & 63) == 0) wm[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        int b = 0;
        for (int w = 0; w < kWaves; ++w) b = max(b, wm[w]);
        out[blockIdx.x] = b;
    }
}

static int block_bandwidth(const Mat& A, int& bw)
{
    bw = 0;
    if (A.n == 0) return CEL_OK;
    constexpr int kG = 256;
    DevBuf<int> part;
    CEL_TRY(part.alloc(kG));
    hipStream_t s = rt().stream;
    hipLaunchKernelGGL(k_bandwidth, dim3(kG), dim3(kBlock), 0, s, A.n, A.xadj.p, A.cols.p, part.p);
    CEL_HIP(hipGetLastError());
    std::vector<int> h(kG);
    CEL_HIP(hipMemcpyAsync(h.data(), part.p, sizeof(int) * kG, hipMemcpyDeviceToHost, s));
    CEL_HIP(hipStreamSynchronize(s));
    for (int v : h) bw = std::max(bw, v);
    return CEL_OK;
}

extern "C" int cel_gpu_batch_create(cel_gpu_batch** out, int nsys, cel_gpu_mat* const* mats,
                                    const int* nrhs)
{
    if (!out || nsys <= 0 || !mats) return fail(CEL_ERR_ARG, "bad arguments to cel_gpu_batch_create");
    CEL_TRY(ensure_init());
    auto h = std::make_unique<cel_gpu_batch>();
    Batch& B = h->b;
    long long off = 0;
    int lane = 0;
    for (int q = 0; q < nsys; ++q) {
        Mat* A = mats[q] ? &mats[q]->m : nullptr;
        if (!A) return fail(CEL_ERR_ARG, "null matrix in batch");
        if (A->n != A->ncols || A->n <= 0) return fail(CEL_ERR_ARG, "batch systems must be square and non-empty");
        if (!A->has_jac) return fail(CEL_ERR_STATE, "cel_gpu_mat_set_jacobi has not been called for a batch system");
        if (!A->halo.empty()) return fail(CEL_ERR_ARG, "batch systems carry no halo blocks");
        const int k = nrhs ? nrhs[q] : 1;
        if (k < 1 || k > kMaxRhs) return fail(CEL_ERR_ARG, "nrhs must be in 1..64");
        int kp = 1, ksh = 0;
        while (kp < k) {
            kp <<= 1;
            ++ksh;
        }
        if ((long long)A->n * kp > (1ll << 31)) return fail(CEL_ERR_ARG, "system too large for one batch lane block");
        B.mats.push_back(A);
        B.k.push_back(k);
        B.kp.push_back(kp);
        B.ksh.push_back(ksh);
        B.lane0.push_back(lane);
        B.voff.push_back(off);
        BSys S{};
        S.xadj = A->xadj.p;
        S.cols = A->cols.p;
        S.vals = A->vals.p;
        S.jac = A->jac.p;
        S.voff = off;
        S.n = A->n;
        S.kp = kp;
        S.ksh = ksh;
        S.lane0 = lane;
        S.ntiles = (int)(((long long)A->n + (kBlock >> ksh) - 1) / (kBlock >> ksh));
        if (kp > 1) {  // band period of the multi-rhs chunk order, in 256-row chunks
            int bw = 0;
            CEL_TRY(block_bandwidth(*A, bw));
            S.period = (bw + kTileRows / 2) / kTileRows;
        }
        B.h_sys.push_back(S);
        off += ((long long)A->n * kp + kVecAlign - 1) / kVecAlign * kVecAlign;
        B.elems += (long long)A->n * kp;
        lane += kp;
        B.kpmax = std::max(B.kpmax, kp);
    }
    B.total = off;
    B.lanes = lane;
    CEL_TRY(B.d_sys.alloc(nsys));
    for (auto* buf : {&B.r, &B.p, &B.v, &B.x, &B.b}) CEL_TRY(buf->alloc((size_t)off));
    B.tune.grid = grid_max();
    if (B.tune.grid >= kXcds) B.tune.grid = B.tune.grid / kXcds * kXcds;
    if (B.tune.grid < nsys) return fail(CEL_ERR_ARG, "more systems than resident blocks in one batch");
    *out = h.release();
    return CEL_OK;
}

extern "C" int cel_gpu_batch_destroy(cel_gpu_batch* B)
{
    delete B;
    return CEL_OK;
}

extern "C" int cel_gpu_batch_info(const cel_gpu_batch* h, int* lanes, long long* elements,
                                  int* grid, int* chunk, int* weight, int* rebalance,
                                  double* ms_per_iter)
{
    if (!h) return fail(CEL_ERR_ARG, "null batch");
    const Batch& B = h->b;
    int nl = 0;
    for (int k : B.k) nl += k;
    if (lanes) *lanes = nl;
    if (elements) *elements = B.elems;
    if (grid) *grid = B.tune.grid;
    if (chunk) *chunk = B.tune.chunk;
    if (weight) *weight = B.tune.weight;
    if (rebalance) *rebalance = B.tune.rebalance;
    if (ms_per_iter) *ms_per_iter = B.tuned_ms_per_iter;
    return CEL_OK;
}

extern "C" int cel_gpu_batch_get_config(const cel_gpu_batch* h, int* grid, int* grid23, int* chunk,
                                        int* weight)
{
    if (!h) return fail(CEL_ERR_ARG, "null batch");
    const Tuning& T = h->b.tune;
    if (grid) *grid = T.grid;
    if (grid23) *grid23 = T.g23();
    if (chunk) *chunk = T.chunk;
    if (weight) *weight = T.weight;
    return CEL_OK;
}

extern "C" int cel_gpu_batch_set_config(cel_gpu_batch* h, int grid, int grid23, int chunk, int weight)
{
    if (!h) return fail(CEL_ERR_ARG, "null batch");
    CEL_TRY(ensure_init());
    Batch& B = h->b;
    const int nsys = (int)B.mats.size();
    const int gmax = grid_max();
    if (grid23 == 0) grid23 = grid;
    for (int g : {grid, grid23})
        if (g < nsys || g < kXcds || g % kXcds || g > gmax)
            return fail(CEL_ERR_ARG, "batch grids must be multiples of 8 between the number of "
                                     "systems and " + std::to_string(gmax));
    if (chunk < 1 || chunk > 1024 || (weight != 0 && weight != 1))
        return fail(CEL_ERR_ARG, "batch chunk must be 1..1024 and weight 0 or 1");
    Tuning T{grid, chunk, weight, 1};
    T.grid23 = grid23;
    B.tune = T;
    B.tuned = false;
    B.tuned_ms_per_iter = 0.0;
    return CEL_OK;
}

extern "C" int cel_gpu_batch_tune(cel_gpu_batch* h, const double* const* b, int memspace, int iters)
{
    if (!h || !b || iters < 1) return fail(CEL_ERR_ARG, "bad arguments to cel_gpu_batch_tune");
    CEL_TRY(ensure_init());
    Batch& B = h->b;
    CEL_TRY(load_rhs(B, b, memspace));
    cel_pcg_params prm;
    cel_gpu_pcg_params_default(&prm);
    prm.tol = 0.0;  // run the full sample
    prm.tol_mode = CEL_TOL_ABS;
    prm.x0_mode = CEL_X0_ZERO;
    prm.alpha_stop = 0;
    const int gmax = std::max(kXcds, grid_max() / kXcds * kXcds);
    // K1 on all, three quarters or half of the resident grid; K2 and K3
    // on the same grid or on half of the resident grid
    std::vector<Tuning> cand;
    for (int gq : {4, 3, 2})
        for (int g2q : {0, 2})
            for (int weight : {0, 1})
                for (int chunk : {8, 32}) {
                    const int g = std::max(kXcds, gmax * gq / 4 / kXcds * kXcds);
                    const int g2 = g2q ? std::max(kXcds, gmax * g2q / 4 / kXcds * kXcds) : g;
                    if (g < (int)B.mats.size() || g2 < (int)B.mats.size()) continue;
                    if (g2q && g2 == g) continue;  // the same as g2q = 0
                    Tuning T{g, chunk, weight, 1};
                    T.grid23 = g2;
                    cand.push_back(T);
                }
    double best = 1e300;
    Tuning bt = B.tune;
    for (const Tuning& T : cand) {
        CEL_TRY(init_x(B, nullptr, prm));
        RunResult rr;
        // the sample is a whole number of chunks
        const int cap = std::max(T.chunk, (iters + T.chunk - 1) / T.chunk * T.chunk) - 1;
        CEL_TRY(run_batch(B, T, prm, cap, rr));
        int it = 0;
        for (const BCtrl& c : rr.lanes) it = std::max(it, c.iter);
        const double per = rr.ms / std::max(1, it);
        if (per < best) {
            best = per;
            bt = T;
        }
    }
    B.tune = bt;
    B.tuned = true;
    B.tuned_ms_per_iter = best;
    return CEL_OK;
}

extern "C" int cel_gpu_batch_solve(cel_gpu_batch* h, const double* const* b, double* const* x,
                                   const cel_pcg_params* params, cel_pcg_stats* lane_stats,
                                   cel_pcg_stats* total)
{
    if (!h || !b || !x || !params) return fail(CEL_ERR_ARG, "bad arguments to cel_gpu_batch_solve");
    CEL_TRY(ensure_init());
    Batch& B = h->b;
    const cel_pcg_params& prm = *params;
    if (prm.precision != CEL_PREC_FP64 || prm.format != CEL_FMT_CSR)
        return fail(CEL_ERR_ARG, "the batched solver is FP64 CSR");
    if (prm.iter_max < 0) return fail(CEL_ERR_ARG, "iter_max < 0");
    CEL_TRY(load_rhs(B, b, prm.memspace));
    CEL_TRY(init_x(B, x, prm));
    RunResult rr;
    CEL_TRY(run_batch(B, B.tune, prm, prm.iter_max, rr));
    hipStream_t s = rt().stream;
    DevBuf<double> tmp;
    for (size_t q = 0; q < B.mats.size(); ++q) {
        const int n = B.mats[q]->n;
        double* dst = x[q];
        if (prm.memspace == CEL_MEM_HOST) {
            CEL_TRY(tmp.alloc((size_t)n * B.k[q]));
            dst = tmp.p;
        }
        hipLaunchKernelGGL(k_bdeinterleave, dim3(1024), dim3(kBlock), 0, s, B.x.p + B.voff[q], dst, n,
                           B.k[q], B.ksh[q]);
        CEL_HIP(hipGetLastError());
        if (prm.memspace == CEL_MEM_HOST) {
            CEL_HIP(hipMemcpyAsync(x[q], tmp.p, sizeof(double) * n * B.k[q], hipMemcpyDeviceToHost, s));
            CEL_HIP(hipStreamSynchronize(s));
        }
    }
    CEL_HIP(hipStreamSynchronize(s));
    int err = 0, notconv = 0, maxit = 0, lane = 0;
    for (size_t q = 0; q < B.mats.size(); ++q) {
        for (int j = 0; j < B.k[q]; ++j, ++lane) {
            const BCtrl& c = rr.lanes[B.lane0[q] + j];
            int st = c.status;
            if (st == 0 && !(c.norm2 <= c.eps)) st = CEL_NOT_CONVERGED;
            if (st < 0) err = std::min(err, st);
            if (st == CEL_NOT_CONVERGED) notconv = 1;
            maxit = std::max(maxit, c.iter);
            if (lane_stats) {
                cel_pcg_stats& o = lane_stats[lane];
                o = cel_pcg_stats{};
                o.status = st;
                o.iters = c.iter;
                o.norm0 = c.norm0;
                o.final_norm = c.norm2;
                o.eps_abs = c.eps;
                o.alpha = c.alpha;
            }
        }
    }
    const int worst = err < 0 ? err : (notconv ? CEL_NOT_CONVERGED : CEL_OK);
    if (total) {
        *total = cel_pcg_stats{};
        total->status = worst;
        total->iters = maxit;
        total->outer_iters = rr.chunks;
        total->solve_ms = rr.ms;
        double bytes = 0.0;
        for (size_t q = 0; q < B.mats.size(); ++q) {
            const Mat& A = *B.mats[q];
            // per iteration: matrix once for all lanes, p gather + v write,
            // K2 r,v,jac read + r write, K3 p,r,jac,x read + p,x write
            bytes += 12.0 * A.nnz + 4.0 * (A.n + 1) + 8.0 * A.n * B.kp[q] * 10.0;
        }
        total->iter_bytes = bytes;
    }
    if (err < 0) set_error("a lane of the batch stopped with an error (see its status)");
    return worst;
}
