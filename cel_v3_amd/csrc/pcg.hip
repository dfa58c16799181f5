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
// One iteration is three streaming kernels:
//   K1  v = A p (LDS-staged CSR tile, DIA rows or JAD), one partial of v.p
//       per block
//   K2  sigma = alpha / sum(v.p) ; r = r + (-sigma) v ; z = jac*r ;
//       partials of r.z and r.r
//   K3  beta = sum(r.z)/alpha ; x = x + sigma p ; p = beta p + jac*r ;
//       alpha, norm2, k, loop decision
// (the x update sits in K3 because K3 reads p anyway; x never feeds back
// into the iteration and each element is the reference's formula).
//
// No kernel waits on another block of the same launch: there are no grid
// barriers, atomics or device-scope fences in the loop.  The scalars travel
// as a small token along the ring of kernels -- every kernel reads the token
// its predecessor wrote and block 0 writes the token for its successor -- and
// a kernel that needs a dot product sums the predecessor's per-block
// partials itself, in every block, in the same fixed order.  All blocks
// therefore hold bit-identical scalars, the sums do not depend on which
// block finished first, and the host never waits inside the loop.  A
// stopped token is forwarded unchanged, so a captured chunk of iterations
// can be replayed as a hipGraph and runs past convergence as no-ops.
//
// Per element the arithmetic is the reference's (products rounded before the
// add, -ffp-contract=off); only the summation order of the dot products
// differs from the reference (fixed tree here, worker chunks there).
#include <algorithm>
#include <chrono>
#include <cmath>
#include <initializer_list>
#include <vector>

#include "cel_internal.hpp"
#include "spmv_device.hpp"

namespace cel {

// Token passed along the kernel ring.
struct Ctrl {
    double alpha, sigma, beta, norm2, eps, norm0, bnorm, rnorm_ref;
    int run;     // 1 while the loop continues
    int iter;    // completed iterations
    int status;  // 0 or a CEL_ERR_* code
    int reduce;  // partials the next K1 must fold in first: 0 none, 1 set-up, 2 replacement
    int ru;      // mixed: replace the residual at the end of this iteration
    int outer;   // mixed: replacements done
    int since;   // mixed: iterations since the last replacement
    int stall;   // mixed: consecutive replacements that did not halve the FP64 residual
};

// Constant solver parameters.
struct Prm {
    double tol, small, ru_delta;
    int tol_mode, iter_max, ru_restart, hist_cap, alpha_stop, ru_max_inner;
    double* hist;
};

// the reference's absolute test |alpha| > 1.1 eps_mach (optional)
__device__ __forceinline__ bool alpha_ok(const Prm& P, double alpha)
{
    return !P.alpha_stop || fabs(alpha) > P.small;
}

enum { kReduceNone = 0, kReduceSetup = 1, kReduceReplace = 2 };

// ------------------------------------------------------------------ helpers
__device__ __forceinline__ double eps_from(const Prm& P, double bnorm, double r0)
{
    if (P.tol_mode == CEL_TOL_REL_B) return P.tol * bnorm;
    if (P.tol_mode == CEL_TOL_REL_R0) return P.tol * r0;
    return P.tol;
}

// Every thread of the block reads the token (written by the previous kernel).
__device__ __forceinline__ Ctrl read_ctrl(const Ctrl* c) { return *c; }

__device__ __forceinline__ void write_ctrl(Ctrl* dst, const Ctrl& c)
{
    if (blockIdx.x == 0 && threadIdx.x == 0) *dst = c;
}

// Sum of G partials in fixed order, returned to every thread of the block.
__device__ __forceinline__ double all_sum(const double* __restrict__ part, int G, double* red,
                                          double* bcast)
{
    const double s = sum_partials(part, G, red);
    if (threadIdx.x == 0) *bcast = s;
    __syncthreads();
    const double v = *bcast;
    __syncthreads();
    return v;
}

// Two folds at once, each in all_sum's order (thread-strided sums, then the
// block tree of block_sum): one pass over both partial arrays.
__device__ __forceinline__ void all_sum2(const double* __restrict__ pa, const double* __restrict__ pb,
                                         int G, double* red /* 2*kWaves */, double* bcast /* 2 */,
                                         double& a, double& b)
{
    double sa = 0.0, sb = 0.0;
    for (int i = threadIdx.x; i < G; i += kBlock) {
        sa = sa + pa[i];
        sb = sb + pb[i];
    }
    block_sum2(sa, sb, red);
    if (threadIdx.x == 0) {
        bcast[0] = sa;
        bcast[1] = sb;
    }
    __syncthreads();
    a = bcast[0];
    b = bcast[1];
    __syncthreads();
}

// Balanced contiguous ranges: logical block lb of G owns units [u0, u1).
__device__ __forceinline__ void unit_range(long long nunits, int G, int lb, long long& u0,
                                           long long& u1)
{
    u0 = nunits * lb / G;
    u1 = nunits * (lb + 1) / G;
}

// ------------------------------------------------------------------ set-up
__global__ void k_fill(long long n, double* __restrict__ x, double v)
{
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (long long)gridDim.x * blockDim.x)
        x[i] = v;
}

// dst[k] = src[perm[k]] (to the permuted row order) and back
template <typename T>
__global__ void k_perm_gather(long long n, const int* __restrict__ perm, const T* __restrict__ src,
                              T* __restrict__ dst)
{
    for (long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x; k < n;
         k += (long long)gridDim.x * blockDim.x)
        dst[k] = src[perm[k]];
}

template <typename T>
__global__ void k_perm_scatter(long long n, const int* __restrict__ perm, const T* __restrict__ src,
                               T* __restrict__ dst)
{
    for (long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x; k < n;
         k += (long long)gridDim.x * blockDim.x)
        dst[perm[k]] = src[k];
}

struct MatView {
    int n, mmax;
    int coo;  // the local block is applied as COO: v(row) += a p entry by entry from v = 0
    long long jnnz;
    const int *xadj, *cols;
    const double* vals;
    const float* vals32;
    const int *jbegin, *jcols, *jperm;
    const double* jvals;
    const float* jvals32;
    // DIA: values [nd][n], presence mask, ascending offsets
    const double* dvals;
    const float* dvals32;
    const uint32_t* dmask;   // presence bits, more than 8 diagonals
    const uint8_t* dmask8;   // presence bits, at most 8 diagonals
    int nd;
    int dzero;               // index of the diagonal with offset 0, or -1
    int doff[32];
    // halo (FP64 only)
    const int *hrow_tile, *hrows, *hrowq, *hptr, *hcols;
    const double* hvals;
    const double* ghost;
    const double* hrec;     // per compact halo row: {p1, p2}
    const uint8_t* hflag;   // per row: min(halo entries, 255); hflag[n] = 0
    const int* hwq;         // per 64-row group: compact index of its first halo row
    int hnq;                // rows with halo entries
};

// The halo records of this SpMV (after the ghost values arrived; here
// after the face pack): one thread per row with halo entries, in compact
// order.
__global__ __launch_bounds__(kBlock) void k_halo_rec(int nq, const int* __restrict__ hptr,
                                                     const int* __restrict__ hcols,
                                                     const double* __restrict__ hvals,
                                                     const double* __restrict__ ghost,
                                                     double* __restrict__ rec,
                                                     const Ctrl* __restrict__ cin)
{
    if (!cin->run) return;  // uniform: the iteration does not run
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= nq) return;
    const int j0 = hptr[q];
    const int c = hptr[q + 1] - j0;
    double w[kHaloRec] = {0.0, 0.0};
#pragma unroll
    for (int k = 0; k < kHaloRec; ++k)
        if (k < c) w[k] = hvals[j0 + k] * ghost[hcols[j0 + k]];
    double* o = rec + 2 * (size_t)q;  // compact: the K1s find q from their wavefront
    o[0] = w[0];
    o[1] = w[1];
}

template <typename VT>
__device__ __forceinline__ const VT* csr_vals(const MatView& A);
template <>
__device__ __forceinline__ const double* csr_vals<double>(const MatView& A) { return A.vals; }
template <>
__device__ __forceinline__ const float* csr_vals<float>(const MatView& A) { return A.vals32; }
template <typename VT>
__device__ __forceinline__ const VT* dia_vals(const MatView& A);
template <>
__device__ __forceinline__ const double* dia_vals<double>(const MatView& A) { return A.dvals; }
template <>
__device__ __forceinline__ const float* dia_vals<float>(const MatView& A) { return A.dvals32; }

template <typename VT>
__device__ __forceinline__ const VT* jad_vals(const MatView& A);
template <>
__device__ __forceinline__ const double* jad_vals<double>(const MatView& A) { return A.jvals; }
template <>
__device__ __forceinline__ const float* jad_vals<float>(const MatView& A) { return A.jvals32; }

// Walks the tiles [t0, t1) of one block in the requested format and
// returns, per tile, the calling thread's row sum and its row.  For CSR
// row = r0 + tid (the next tile's loads are issued while this tile's
// gathers wait); for JAD row = perm[r0 + tid].  HALO adds the halo entries
// after the local sum ((0 + local) + halo...).  With xrow, x[row] is
// loaded too (before the next tile's loads).  next() contains
// __syncthreads() for CSR: every thread of the block calls it.
template <typename VT, typename AT, int FMT, bool HALO>
struct TileWalk {
    CsrTileStream<VT> cs;

    // Contains __syncthreads() for CSR: every thread of the block calls it.
    __device__ __forceinline__ void start(const MatView& A, long long t0, long long t1,
                                          CsrStage<VT>& sm)
    {
        if constexpr (FMT == CEL_FMT_CSR) {
            if (t0 < t1) {  // uniform across the block
                cs.start(A.n, (int)t0, t0 + 1 < t1 ? (int)t0 + 1 : -1, A.xadj, A.cols,
                         csr_vals<VT>(A), sm);
                cs.stage(sm);
                __syncthreads();
            }
        }
    }

    __device__ __forceinline__ AT next(const MatView& A, long long t, long long t1,
                                       const VT* __restrict__ x, CsrStage<VT>& sm, int& row,
                                       VT* xrow)
    {
        const int r0 = (int)t * kTileRows;
        AT s;
        if constexpr (FMT == CEL_FMT_CSR) {
            const int nxt = (t + 1 < t1) ? (int)(t + 1) : -1;
            const int aft = (t + 2 < t1) ? (int)(t + 2) : -1;
            row = r0 + threadIdx.x;
            s = cs.template rows<AT>(A.n, (int)t, nxt, aft, A.xadj, A.cols, csr_vals<VT>(A), x, sm,
                                     row, xrow);
            __syncthreads();  // every row of the tile is summed
            if (nxt >= 0) {   // uniform across the block
                cs.stage(sm);
                __syncthreads();
            }
            // CSR: r = 0; r = r + tmp.  COO: the entries were added to 0 one
            // by one (the same bits but for the sign of a zero sum)
            s = A.coo ? s : AT(0) + s;
            if constexpr (HALO) {
                if (row < A.n)
                    s = halo_row<AT>(row, (int)t, s, A.hrow_tile, A.hrows, A.hptr, A.hcols,
                                     A.hvals, A.ghost);
            }
        } else {
            // JAD solves keep their vectors in the permuted row order, with
            // the columns renumbered into it: row k of the permuted system
            const int k = r0 + threadIdx.x;
            s = AT(0);
            row = A.n;
            if (k < A.n) {
                VT xr = VT(0);
                if (xrow) xr = x[k];
                s = jad_row<VT, AT>(k, A.mmax, A.jbegin, A.jcols, jad_vals<VT>(A), x);
                row = k;
                if (xrow) *xrow = xr;
            }
        }
        return s;
    }
};

// ------------------------------------------------------------------ residual kernels
// r = b - A x with the FP64 operator; per-block partials: r.z, r.r, b.b.
// FP64 form: r and p = z = jac*r stored in FP64.  Mixed form (OUT_T =
// float): r rounded to FP32, z = jac32 * r32 (p = z at the start or on a
// restart), e = 0 at the start, the norm taken in FP64 before rounding.
// The next K1 folds the partials into the token.
template <typename OUT_T, int FMT, bool HALO, bool REPLACE>
__global__ __launch_bounds__(kBlock) void k_residual(MatView A, const double* __restrict__ x,
                                                     const double* __restrict__ b,
                                                     const double* __restrict__ jac,
                                                     const float* __restrict__ jac32,
                                                     OUT_T* __restrict__ r, OUT_T* __restrict__ p,
                                                     float* __restrict__ e,
                                                     double* __restrict__ part,
                                                     const Ctrl* __restrict__ cin,
                                                     Ctrl* __restrict__ cout, Prm P)
{
    Ctrl c = read_ctrl(cin);
    if (REPLACE) {
        if (!c.run || !c.ru) {
            write_ctrl(cout, c);
            return;
        }
    }
    __shared__ CsrStage<double> sm;
    __shared__ double red[3 * kWaves];
    const int lb = logical_block();
    const int G = gridDim.x;
    const long long ntiles = (A.n + kTileRows - 1) / kTileRows;
    long long t0, t1;
    unit_range(ntiles, G, lb, t0, t1);
    const bool restart = REPLACE && P.ru_restart;
    double arz = 0.0, arr = 0.0, abb = 0.0;
    TileWalk<double, double, FMT, HALO> walk;
    walk.start(A, t0, t1, sm);
    for (long long t = t0; t < t1; ++t) {
        int row;
        double s = walk.next(A, t, t1, x, sm, row, nullptr);
        if (row < A.n) {
            const double bi = b[row];
            const double ri = bi - s;  // r = 0; r = r + A x; r = b - r
            double bb = bi * bi;  // a non-finite b makes this sum non-finite
            abb = abb + bb;
            double rr = ri * ri;
            arr = arr + rr;
            if constexpr (sizeof(OUT_T) == 8) {
                const double zi = ri * jac[row];
                r[row] = ri;
                p[row] = zi;
                double rz = ri * zi;
                arz = arz + rz;
            } else {
                const float rf = (float)ri;
                const float zf = rf * jac32[row];
                r[row] = rf;
                if (!REPLACE || restart) p[row] = zf;
                if (!REPLACE) e[row] = 0.0f;
                double rz = (double)rf * (double)zf;
                arz = arz + rz;
            }
        }
    }
    block_sum2(arz, arr, red);
    const double sbb = block_sum(abb, red + 2 * kWaves);
    if (threadIdx.x == 0) {
        part[lb] = arz;
        part[G + lb] = arr;
        part[2 * G + lb] = sbb;
    }
    c.reduce = REPLACE ? kReduceReplace : kReduceSetup;
    c.ru = 0;
    write_ctrl(cout, c);
}

// The CSR, DIA and JAD K1s walk their tiles in group_tile's order
// (spmv_device.hpp).  The CSR and DIA K1s add a thread's rows to the block
// partial in that order, so their partials and solves are bit-identical;
// the JAD K1's tiles hold permuted rows, so it shares only the order of the
// tiles, and its solves agree with the others to rounding.  (Groups of 4
// tiles measured: FP64 DIA solve 1% faster, mixed DIA 1% slower.)

// K1: v = A p, per-block partial of v.p.  Folds pending set-up or
// replacement partials into the token first (every block, same order) and
// decides whether the iteration runs.
template <typename VT, int FMT, bool HALO>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(sizeof(VT) == 4 ? 5 : 4, 8))) void k_pcg_spmv(MatView A, const VT* __restrict__ p,
                                                     VT* __restrict__ v, double* __restrict__ part,
                                                     const double* __restrict__ rpart, int Gr,
                                                     const Ctrl* __restrict__ cin,
                                                     Ctrl* __restrict__ cout, Prm P)
{
    __shared__ double red[kWaves];
    __shared__ double bcast;
    Ctrl c = read_ctrl(cin);
    if (c.reduce != kReduceNone) {
        const double rz = all_sum(rpart, Gr, red, &bcast);
        const double rr = all_sum(rpart + Gr, Gr, red, &bcast);
        const double rn = sqrt(rr);
        if (c.reduce == kReduceSetup) {
            const double bb = all_sum(rpart + 2 * Gr, Gr, red, &bcast);
            c.bnorm = sqrt(bb);
            c.norm0 = rn;
            c.eps = eps_from(P, c.bnorm, rn);
            c.iter = 0;
            c.outer = 0;
            if (!isfinite(bb) && c.status == 0) c.status = CEL_ERR_NONFINITE;
            if (blockIdx.x == 0 && threadIdx.x == 0 && P.hist_cap > 0) P.hist[0] = rn;
        } else {
            c.outer = c.outer + 1;
            // attainable accuracy: two cycles in a row that fail to halve the
            // FP64 residual end the solve (not converged)
            c.stall = (rn > 0.5 * c.rnorm_ref) ? c.stall + 1 : 0;
        }
        c.since = 0;
        c.alpha = rz;
        c.norm2 = rn;
        c.rnorm_ref = rn;
        c.reduce = kReduceNone;
        c.run = (c.status == 0 && c.iter < P.iter_max && alpha_ok(P, c.alpha) &&
                 c.norm2 > c.eps && c.stall < 2)
                    ? 1
                    : 0;
    }
    if (!c.run) {
        write_ctrl(cout, c);
        return;
    }
    __shared__ CsrStage<VT> sm;
    const int lb = logical_block();
    const long long ntiles = (A.n + kTileRows - 1) / kTileRows;
    double acc = 0.0;
    if constexpr (FMT == CEL_FMT_CSR) {
        // halo rows (HALO): a row's halo record (count and products, in
        // block order) is added after its local sum.  Its flag loads one
        // step ahead, behind the previous tile's entries; the record loads
        // in the row's own step, behind the gathers (so the wait for the
        // gathers does not include it) and ahead of the next tile's entries
        // (whose wait, for the stage, covers it).  Unflagged rows load no
        // record and read count 0.  Each loaded value is used once in the
        // step after it, in place: a copy of a register still being loaded
        // would wait for every load issued before it.
        CsrTileStream<VT> cs;
        const VT* vals = csr_vals<VT>(A);
        const long long tfirst = group_tile(0, ntiles);
        const auto rrec = tile_rsrc(HALO ? A.hrec : nullptr, HALO ? A.hnq * 16 : 0);
        // flag index of this thread's row of tile t (rows past n read hflag[n] = 0)
        auto fidx = [&](long long t) {
            const long long r = t * kTileRows + threadIdx.x;
            return t >= 0 && r < A.n ? (int)r : A.n;
        };
        // the halo flag (entry count) of this thread's row of the tile being
        // summed, and its wavefront's first compact halo index
        int hf = 0, hq = 0;
        if (tfirst >= 0) {  // uniform across the block
            cs.start(A.n, (int)tfirst, (int)group_tile(1, ntiles), A.xadj, A.cols, vals, sm);
            if constexpr (HALO) {
                hf = A.hflag[fidx(tfirst)];
                hq = A.hwq[fidx(tfirst) >> 6];
            }
            cs.stage(sm);
            __syncthreads();
        }
        long long k = 0;
        for (long long t = tfirst; t >= 0; t = group_tile(++k, ntiles)) {
            const int nxt = (int)group_tile(k + 1, ntiles);
            const int aft = (int)group_tile(k + 2, ntiles);
            const int row = (int)t * kTileRows + threadIdx.x;
            VT pr = VT(0);
            HaloRec hw{};
            int hc = 0;
            auto recs = [&]() {
                if constexpr (HALO) {
                    hc = hf;
                    hw = halo_rec_load(rrec, hq, hc);
                    hf = A.hflag[fidx(nxt)];
                    hq = A.hwq[fidx(nxt) >> 6];
                }
            };
            VT s = cs.template rows<VT>(A.n, (int)t, nxt, aft, A.xadj, A.cols, vals, p, sm, row,
                                        &pr, VT(0), recs);
            s = A.coo ? s : VT(0) + s;
            if constexpr (HALO) {
                bool over = false;
                const VT sh = halo_rec_add<VT>(s, hw, hc, over);
                s = over ? halo_row<VT>(row, (int)t, s, A.hrow_tile, A.hrows, A.hptr, A.hcols, A.hvals,
                                        A.ghost)
                         : sh;
            }
            __syncthreads();  // the tile's LDS stage is consumed
            if (nxt >= 0) {
                cs.stage(sm);
                __syncthreads();
            }
            if (row < A.n) {
                v[row] = s;
                double prod = double(s) * double(pr);
                acc = acc + prod;
            }
        }
    } else if constexpr (FMT == CEL_FMT_DIA) {
        // all tiles of a group per step, one row of each per thread; the
        // rows are added to the partial in tile order, as in the CSR K1.
        // Halo rows (HALO): as in the CSR K1, a row's record is added after
        // its local sum.  The records of a step load at the end of the step
        // before (after its products, into the registers those just freed)
        // and the flags one step earlier still, so both are in by the time
        // the step's products are and neither is copied while in flight.
        constexpr int R = kTileGroup;
        const long long ngroups = (ntiles + R - 1) / R;
        const auto rrec = tile_rsrc(HALO ? A.hrec : nullptr, HALO ? A.hnq * 16 : 0);
        // flag index of this thread's row of tile t (rows past n read hflag[n] = 0)
        auto fidx = [&](long long t) {
            const long long r = t * kTileRows + threadIdx.x;
            return t < ntiles && r < A.n ? (int)r : A.n;
        };
        // flags and first compact halo indices of the next group, entry
        // counts and records of this group
        int hf[R], hq[R], hc[R];
        HaloRec hr[R];
        auto load_flags = [&](long long q) {
#pragma unroll
            for (int k = 0; k < R; ++k) {
                hf[k] = A.hflag[fidx(R * q + k)];
                hq[k] = A.hwq[fidx(R * q + k) >> 6];
            }
        };
        // a wave without halo rows in a tile skips that tile's record load
        auto load_recs = [&]() {
#pragma unroll
            for (int k = 0; k < R; ++k) {
                hc[k] = hf[k];
                hr[k] = HaloRec{};
                if (__builtin_amdgcn_ballot_w64(hc[k] != 0)) hr[k] = halo_rec_load(rrec, hq[k], hc[k]);
            }
        };
        if constexpr (HALO) {
            const long long q0 = blockIdx.x;
            load_flags(q0);
            load_recs();
            load_flags(q0 + gridDim.x);
        }
        for (long long q = blockIdx.x; q < ngroups; q += gridDim.x) {
            int r[R];
#pragma unroll
            for (int k = 0; k < R; ++k) {
                const long long t = R * q + k;
                r[k] = t < ntiles ? (int)t * kTileRows + threadIdx.x : A.n;
            }
            if (r[0] < A.n) {
                VT s[R], pr[R];
                // p[row] comes from the loads of the diagonal with offset 0
                dia_rows<VT, R, (R >= 4 ? 4 : 8)>(A.n, A.nd, A.doff, dia_vals<VT>(A), A.dmask8,
                                                   A.dmask, r, p, s, pr, A.dzero);
#pragma unroll
                for (int k = 0; k < R; ++k) {
                    if (r[k] < A.n) {
                        const VT pk = A.dzero >= 0 ? pr[k] : p[r[k]];
                        VT sk = VT(0) + s[k];
                        if constexpr (HALO) {
                            bool over = false;
                            sk = halo_rec_add<VT>(sk, hr[k], hc[k], over);
                            if (over) {  // more than kHaloRec entries: the compact lists
                                VT s0 = VT(0) + s[k];
                                const int hq = A.hrowq[r[k]];
                                for (int j = A.hptr[hq]; j < A.hptr[hq + 1]; ++j) {
                                    VT prod = VT(A.hvals[j]) * VT(A.ghost[A.hcols[j]]);
                                    s0 = s0 + prod;
                                }
                                sk = s0;
                            }
                        }
                        v[r[k]] = sk;
                        double prod = double(sk) * double(pk);
                        acc = acc + prod;
                    }
                }
            }
            if constexpr (HALO) {  // the next group's records, the flags of the one after
                load_recs();
                load_flags(q + 2 * (long long)gridDim.x);
            }
        }
    } else {
        // JAD in the permuted row order: the CSR K1's tile-pair order, one
        // row per thread and tile, the next tile's row loaded while this
        // one's gathers are in flight
        JadStream<VT> js;
        js.init(A.n, A.mmax, A.jnnz, A.jbegin, A.jcols, jad_vals<VT>(A), p);
        long long k = 0;
        long long t = group_tile(0, ntiles);
        if (t >= 0) js.fetch((int)t * kTileRows + threadIdx.x);
        for (; t >= 0; t = group_tile(++k, ntiles)) {
            const long long tn = group_tile(k + 1, ntiles);
            const int row = (int)t * kTileRows + threadIdx.x;
            const int rn = tn >= 0 ? (int)tn * kTileRows + threadIdx.x : A.n;
            VT pr;
            VT s = js.template row<VT>(row, rn, A.mmax, A.jbegin, A.jcols, jad_vals<VT>(A), p, pr);
            if (row < A.n) {
                // the reference's JAD product overwrites v (y(perm(k)) = tmp(k))
                v[row] = s;
                double prod = double(s) * double(pr);
                acc = acc + prod;
            }
        }
    }
    const double bs = block_sum(acc, red);
    if (threadIdx.x == 0) part[lb] = bs;
    write_ctrl(cout, c);
}

// 16-byte vectors for the streaming kernels
template <typename VT>
struct Vec;
template <>
struct Vec<double> {
    using T = double2;
    static constexpr int W = 2;
};
template <>
struct Vec<float> {
    using T = float4;
    static constexpr int W = 4;
};

template <typename VT>
__device__ __forceinline__ VT lane(const typename Vec<VT>::T& a, int k)
{
    if constexpr (Vec<VT>::W == 2) {
        return k == 0 ? a.x : a.y;
    } else {
        return k == 0 ? a.x : k == 1 ? a.y : k == 2 ? a.z : a.w;
    }
}

template <typename VT>
__device__ __forceinline__ void set_lane(typename Vec<VT>::T& a, int k, VT v)
{
    if constexpr (Vec<VT>::W == 2) {
        if (k == 0) a.x = v;
        else a.y = v;
    } else {
        if (k == 0) a.x = v;
        else if (k == 1) a.y = v;
        else if (k == 2) a.z = v;
        else a.w = v;
    }
}

// Whole 16-byte units per thread and loop step in the vector kernels (all
// their loads in flight together).  Measured on the 256^3 solves (two
// interleaved repetitions, one box): FP32 4 (mixed CSR 304-305 ms against
// 309-310 with 1, 308-310 with 2), FP64 2 (CSR 497-501 against 503-507
// with 1, 501-506 with 4).
template <typename VT>
constexpr int kVecUnrollOf = sizeof(VT) == 4 ? 4 : 2;

// K2: sigma = alpha / sum(v.p) ; r = r + (-sigma) v ; z = jac*r ;
// per-block partials of r.z and r.r.
template <typename VT>
__global__ __launch_bounds__(kBlock) void k_pcg_update_r(long long n, VT* __restrict__ r,
                                                         const VT* __restrict__ v,
                                                         const VT* __restrict__ jac,
                                                         double* __restrict__ part,
                                                         const double* __restrict__ vpart, int G1,
                                                         const Ctrl* __restrict__ cin,
                                                         Ctrl* __restrict__ cout)
{
    __shared__ double red[2 * kWaves];
    __shared__ double bcast;
    Ctrl c = read_ctrl(cin);
    if (!c.run) {
        write_ctrl(cout, c);
        return;
    }
    const double vp = all_sum(vpart, G1, red, &bcast);
    const double sigma = c.alpha / vp;
    if (!(vp > 0.0) || !isfinite(sigma)) {
        c.status = CEL_ERR_BREAKDOWN;
        c.run = 0;
        write_ctrl(cout, c);
        return;
    }
    c.sigma = sigma;
    using V = typename Vec<VT>::T;
    constexpr int W = Vec<VT>::W;
    const int lb = logical_block();
    const int G = gridDim.x;
    const long long unit = (long long)kBlock * W;
    long long u0, u1;
    unit_range((n + unit - 1) / unit, G, lb, u0, u1);
    const long long lo = u0 * unit, hi = min(u1 * unit, n);
    const VT ms = VT(-sigma);
    double arz = 0.0, arr = 0.0;
    constexpr int kVecUnroll = kVecUnrollOf<VT>;
    long long i = lo + (long long)threadIdx.x * W;
    // kVecUnroll units per step while they are whole vectors (all their
    // loads in flight together), added to the partials in unit order
    for (; i + (kVecUnroll - 1) * unit + W <= hi; i += kVecUnroll * unit) {
        V vv[kVecUnroll], rv[kVecUnroll], jv[kVecUnroll];
#pragma unroll
        for (int u = 0; u < kVecUnroll; ++u) {
            vv[u] = *reinterpret_cast<const V*>(v + i + u * unit);
            rv[u] = *reinterpret_cast<const V*>(r + i + u * unit);
            jv[u] = *reinterpret_cast<const V*>(jac + i + u * unit);
        }
#pragma unroll
        for (int u = 0; u < kVecUnroll; ++u) {
#pragma unroll
            for (int k = 0; k < W; ++k) {
                VT t = ms * lane<VT>(vv[u], k);
                VT ri = lane<VT>(rv[u], k) + t;
                set_lane<VT>(rv[u], k, ri);
                VT zi = ri * lane<VT>(jv[u], k);
                double rz = double(ri) * double(zi);
                arz = arz + rz;
                double rr = double(ri) * double(ri);
                arr = arr + rr;
            }
            *reinterpret_cast<V*>(r + i + u * unit) = rv[u];
        }
    }
    for (; i < hi; i += unit) {
        if (i + W <= n) {
            const V vv = *reinterpret_cast<const V*>(v + i);
            V rv = *reinterpret_cast<const V*>(r + i);
            const V jv = *reinterpret_cast<const V*>(jac + i);
#pragma unroll
            for (int k = 0; k < W; ++k) {
                VT t = ms * lane<VT>(vv, k);
                VT ri = lane<VT>(rv, k) + t;
                set_lane<VT>(rv, k, ri);
                VT zi = ri * lane<VT>(jv, k);
                double rz = double(ri) * double(zi);
                arz = arz + rz;
                double rr = double(ri) * double(ri);
                arr = arr + rr;
            }
            *reinterpret_cast<V*>(r + i) = rv;
        } else {
            for (long long j = i; j < hi; ++j) {
                VT t = ms * v[j];
                VT ri = r[j] + t;
                r[j] = ri;
                VT zi = ri * jac[j];
                double rz = double(ri) * double(zi);
                arz = arz + rz;
                double rr = double(ri) * double(ri);
                arr = arr + rr;
            }
        }
    }
    block_sum2(arz, arr, red);
    if (threadIdx.x == 0) {
        part[lb] = arz;
        part[G + lb] = arr;
    }
    write_ctrl(cout, c);
}

// K3: beta = sum(r.z)/alpha ; x = x + sigma p ; p = beta p + jac*r ;
// alpha = r.z, norm2, k++, history, loop decision (FP64) or replacement
// request (mixed).
template <typename VT, typename XT, bool MIXED>
__global__ __launch_bounds__(kBlock) void k_pcg_update_xp(long long n, XT* __restrict__ x,
                                                          VT* __restrict__ p,
                                                          const VT* __restrict__ r,
                                                          const VT* __restrict__ jac,
                                                          double* __restrict__ x64,
                                                          const double* __restrict__ rpart, int G2,
                                                          const Ctrl* __restrict__ cin,
                                                          Ctrl* __restrict__ cout, Prm P)
{
    __shared__ double red[2 * kWaves];
    __shared__ double bcast[2];
    Ctrl c = read_ctrl(cin);
    if (!c.run) {
        write_ctrl(cout, c);
        return;
    }
    double rz, rr;
    all_sum2(rpart, rpart + G2, G2, red, bcast, rz, rr);  // r.z and r.r in one pass
    const double beta_d = rz / c.alpha;
    const VT sigma = VT(c.sigma), beta = VT(beta_d);
    // scalars of the next iteration (they do not depend on the vectors)
    c.beta = beta_d;
    c.alpha = rz;
    c.norm2 = sqrt(rr);
    c.iter = c.iter + 1;
    if (blockIdx.x == 0 && threadIdx.x == 0 && c.iter < P.hist_cap) P.hist[c.iter] = c.norm2;
    if (!MIXED) {
        c.run = (c.iter < P.iter_max && alpha_ok(P, c.alpha) && c.norm2 > c.eps) ? 1 : 0;
    } else {
        // replace the FP32 residual by the FP64 one when it has dropped by
        // ru_delta since the last replacement, looks converged, the budget
        // is spent, or the FP32 recurrence has stopped tracking the FP64
        // residual (grown past it, or a cycle far longer than expected);
        // the stop decision is taken on the FP64 norm
        c.since = c.since + 1;
        const bool drop = c.norm2 < P.ru_delta * c.rnorm_ref;
        const bool astray = c.norm2 > 2.0 * c.rnorm_ref || c.since >= P.ru_max_inner;
        c.ru = (drop || astray || c.norm2 <= c.eps || c.iter >= P.iter_max ||
                !alpha_ok(P, c.alpha) || !isfinite(c.norm2))
                   ? 1
                   : 0;
    }
    // mixed, replacement iteration: the FP32 correction is folded into the
    // FP64 solution here (x64 = x64 + e, e = 0) instead of in a kernel of
    // its own; same arithmetic
    const bool flush = MIXED && c.ru;
    using V = typename Vec<VT>::T;
    constexpr int W = Vec<VT>::W;
    const int lb = logical_block();
    const long long unit = (long long)kBlock * W;
    long long u0, u1;
    unit_range((n + unit - 1) / unit, gridDim.x, lb, u0, u1);
    const long long lo = u0 * unit, hi = min(u1 * unit, n);
    constexpr int kVecUnroll = kVecUnrollOf<VT>;
    long long i = lo + (long long)threadIdx.x * W;
    // kVecUnroll units per step while they are whole vectors
    for (; i + (kVecUnroll - 1) * unit + W <= hi; i += kVecUnroll * unit) {
        V pv[kVecUnroll], rv[kVecUnroll], jv[kVecUnroll];
        XT xs[kVecUnroll][W];
#pragma unroll
        for (int u = 0; u < kVecUnroll; ++u) {
            pv[u] = *reinterpret_cast<const V*>(p + i + u * unit);
            rv[u] = *reinterpret_cast<const V*>(r + i + u * unit);
            jv[u] = *reinterpret_cast<const V*>(jac + i + u * unit);
#pragma unroll
            for (int k = 0; k < W; ++k) xs[u][k] = x[i + u * unit + k];
        }
#pragma unroll
        for (int u = 0; u < kVecUnroll; ++u) {
#pragma unroll
            for (int k = 0; k < W; ++k) {
                const VT pi = lane<VT>(pv[u], k);
                XT t = XT(sigma * pi);
                xs[u][k] = xs[u][k] + t;
                VT zi = lane<VT>(rv[u], k) * lane<VT>(jv[u], k);
                VT bp = beta * pi;
                set_lane<VT>(pv[u], k, bp + zi);
            }
            if (flush) {
#pragma unroll
                for (int k = 0; k < W; ++k) {
                    double t = (double)xs[u][k];
                    x64[i + u * unit + k] = x64[i + u * unit + k] + t;
                    xs[u][k] = XT(0);
                }
            }
#pragma unroll
            for (int k = 0; k < W; ++k) x[i + u * unit + k] = xs[u][k];
            *reinterpret_cast<V*>(p + i + u * unit) = pv[u];
        }
    }
    for (; i < hi; i += unit) {
        if (i + W <= n) {
            V pv = *reinterpret_cast<const V*>(p + i);
            const V rv = *reinterpret_cast<const V*>(r + i);
            const V jv = *reinterpret_cast<const V*>(jac + i);
            XT xs[W];
#pragma unroll
            for (int k = 0; k < W; ++k) xs[k] = x[i + k];
#pragma unroll
            for (int k = 0; k < W; ++k) {
                const VT pi = lane<VT>(pv, k);
                XT t = XT(sigma * pi);
                xs[k] = xs[k] + t;
                VT zi = lane<VT>(rv, k) * lane<VT>(jv, k);
                VT bp = beta * pi;
                set_lane<VT>(pv, k, bp + zi);
            }
            if (flush) {
#pragma unroll
                for (int k = 0; k < W; ++k) {
                    double t = (double)xs[k];
                    x64[i + k] = x64[i + k] + t;
                    xs[k] = XT(0);
                }
            }
#pragma unroll
            for (int k = 0; k < W; ++k) x[i + k] = xs[k];
            *reinterpret_cast<V*>(p + i) = pv;
        } else {
            for (long long j = i; j < hi; ++j) {
                const VT pi = p[j];
                XT t = XT(sigma * pi);
                XT xj = x[j] + t;
                if (flush) {
                    double tt = (double)xj;
                    x64[j] = x64[j] + tt;
                    xj = XT(0);
                }
                x[j] = xj;
                VT zi = r[j] * jac[j];
                VT bp = beta * pi;
                p[j] = bp + zi;
            }
        }
    }
    write_ctrl(cout, c);
}

// Mixed form, replacement step 1: x64 = x64 + e32 ; e32 = 0.
__global__ __launch_bounds__(kBlock) void k_ru_flush(long long n, double* __restrict__ x,
                                                     float* __restrict__ e,
                                                     const Ctrl* __restrict__ cin,
                                                     Ctrl* __restrict__ cout, int force)
{
    const Ctrl c = read_ctrl(cin);
    write_ctrl(cout, c);
    if (!force && (!c.run || !c.ru)) return;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (long long)gridDim.x * blockDim.x) {
        double t = (double)e[i];
        x[i] = x[i] + t;
        e[i] = 0.0f;
    }
}

// ------------------------------------------------------------------ host side
struct Workspace {
    DevBuf<double> r, p, v, x, b, part1, part2, rpart, ghost, sendbuf, hist;
    DevBuf<float> r32, p32, v32, e32;
    // JAD solves: b, x and the Jacobi vector in the permuted row order
    DevBuf<double> bp, xp, jacp;
    // halo solves: per-row halo records {count, p1, p2, p3}
    DevBuf<double> hrec;
    DevBuf<float> jacp32;
    DevBuf<Ctrl> ctrl;
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
    CEL_TRY(W.hrow_tile.reserve(ntiles + 1));
    CEL_HIP(hipMemcpy(W.hrow_tile.p, ht.data(), sizeof(int) * (ntiles + 1), hipMemcpyHostToDevice));
    return CEL_OK;
}

// Resident blocks of several kernels at once: the smallest, so every
// variant that feeds the same consumer uses the same grid (and therefore
// the same partials and the same sums).
// Blocks per CU of the vector kernels (K2, K3).  Measured on the 256^3
// solve: larger grids of these kernels, and K1 grids larger than the
// resident blocks, were slower -- every consumer block sums all producer
// partials, so more blocks cost more reduction work.
constexpr int kVecBlocksPerCu = 4;

// Blocks per CU of the K1s (at most the resident blocks).  Measured on the
// 256^3 block (mean K1 time over 100-iteration solves, three interleaved
// repetitions on one box): FP32 4 per CU (CSR 206-223 us, DIA 121-127 us,
// against 223-230 / 136-145 us with every resident block), FP64 2 per CU
// (CSR 349-352 against 376 us, DIA 260-264 against 266-271 us).
constexpr int kK1BlocksPerCuF32 = 4;
constexpr int kK1BlocksPerCuF64 = 2;

static int common_grid(std::initializer_list<const void*> kernels, long long work)
{
    int g = 1 << 30;
    for (const void* k : kernels) g = std::min(g, resident_grid(k, work));
    return g;
}

// The K1 grid: at most kK1BlocksPerCu* blocks per CU, a multiple of the XCDs.
static int k1_cap(int G1, bool mixed)
{
    const int cap = rt().num_cus * (mixed ? kK1BlocksPerCuF32 : kK1BlocksPerCuF64);
    if (G1 > cap) G1 = cap >= kXcds ? cap / kXcds * kXcds : cap;
    return G1;
}

// The allocation-free entry point's device memory: a file-scope array
// (one copy per GPU, resolved to a device address once per device), laid
// out as r, p, v (kNoAllocMaxRows doubles each), the three partial-sum
// arrays (kNoAllocMaxGrid, 2x and 3x that) and the five token slots.
constexpr size_t kNoAllocMaxRows = size_t(1) << 20;
constexpr size_t kNoAllocMaxGrid = 8192;
constexpr size_t kNoAllocDoubles = 3 * kNoAllocMaxRows + 6 * kNoAllocMaxGrid;
constexpr size_t kNoAllocBytes = kNoAllocDoubles * sizeof(double) + 5 * 128;
static_assert(sizeof(Ctrl) <= 128, "token slot");
static __device__ __align__(256) unsigned char scratch[kNoAllocBytes];
constexpr int kMaxDevices = 64;
static void* g_scratch[kMaxDevices];

int pcg_noalloc_setup(int device)
{
    if (device < 0 || device >= kMaxDevices) return fail(CEL_ERR_ARG, "device index beyond the scratch table");
    if (g_scratch[device]) return CEL_OK;
    void* ptr = nullptr;
    // the name `scratch` is a host-side stand-in; kernels get this address
    CEL_HIP(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
    g_scratch[device] = ptr;
    return CEL_OK;
}

// Device pointers of one solve's work vectors, partial sums and token.
struct WorkPtrs {
    double *r = nullptr, *p = nullptr, *v = nullptr, *part1 = nullptr, *part2 = nullptr,
           *rpart = nullptr;
    float *r32 = nullptr, *p32 = nullptr, *v32 = nullptr, *e32 = nullptr;
    Ctrl* ctrl = nullptr;
};

// One FP64 K1 of the solve on caller vectors (device pointers, n rows):
// y = A x in the CSR or DIA form, plus, with ghost, the halo products of
// the caller's ghost values (the halo records of the solve, built from
// them first), and dot = x.y summed over the K1's block partials in
// logical-block order, as the solve's K2 sums them.
int pcg_spmv_dot(Mat& A, int fmt, const double* x, const double* ghost, double* y, double* dot)
{
    CEL_TRY(ensure_init());
    const int n = A.n;
    if (fmt != CEL_FMT_CSR && fmt != CEL_FMT_DIA)
        return fail(CEL_ERR_ARG, "the fused product runs in the CSR or the DIA form");
    if (fmt == CEL_FMT_DIA && A.dia_nd == 0)
        return fail(CEL_ERR_STATE, "DIA requested but the matrix has no DIA form");
    if (A.ncols != n) return fail(CEL_ERR_ARG, "the fused product needs a square local block");
    const bool halo = ghost != nullptr;
    if (halo && A.halo.empty()) return fail(CEL_ERR_STATE, "ghost values given but no halo blocks");
    if (halo && (double)A.halo_nrows * 16.0 >= 2147483648.0 - 64.0)
        return fail(CEL_ERR_ARG, "the halo records need 31-bit offsets (< 2^27 halo rows)");
    if (fmt == CEL_FMT_DIA && (double)n * 8.0 >= 2147483648.0)
        return fail(CEL_ERR_ARG, "the DIA kernels address x with 32-bit byte offsets");
    if (dot) *dot = 0.0;
    if (n == 0) return CEL_OK;
    const int ntiles = (n + kTileRows - 1) / kTileRows;
    const int G1 = k1_cap(common_grid({(const void*)k_pcg_spmv<double, CEL_FMT_CSR, false>,
                                       (const void*)k_pcg_spmv<double, CEL_FMT_CSR, true>,
                                       (const void*)k_pcg_spmv<double, CEL_FMT_DIA, false>,
                                       (const void*)k_pcg_spmv<double, CEL_FMT_DIA, true>},
                                      ntiles),
                          false);
    Workspace& W = ws_of(A);
    CEL_TRY(W.part1.reserve((size_t)G1));
    CEL_TRY(W.ctrl.reserve(5));
    if (halo) {
        CEL_TRY(W.hrec.reserve(2 * (size_t)std::max(A.halo_nrows, 1)));
        // the CSR K1 finds the entries of rows with more than kHaloRec halo
        // entries through the per-tile compact row ranges
        CEL_TRY(build_hrow_tile(A, W, ntiles));
    }
    hipStream_t s = rt().stream;
    Ctrl c0{};
    c0.run = 1;
    CEL_HIP(hipMemcpyAsync(W.ctrl.p, &c0, sizeof(Ctrl), hipMemcpyHostToDevice, s));
    MatView V{};
    V.n = n;
    V.xadj = A.xadj.p;
    V.cols = A.cols.p;
    V.vals = A.vals.p;
    V.dvals = A.dia_vals.p;
    V.dmask = A.dia_mask.p;
    V.dmask8 = A.dia_mask8.p;
    V.nd = A.dia_nd;
    V.dzero = -1;
    for (int d = 0; d < A.dia_nd; ++d) {
        V.doff[d] = A.dia_off[d];
        if (A.dia_off[d] == 0) V.dzero = d;
    }
    if (halo) {
        V.hrow_tile = W.hrow_tile.p;
        V.hrows = A.halo_rows.p;
        V.hptr = A.halo_ptr.p;
        V.hcols = A.halo_cols.p;
        V.hvals = A.halo_vals.p;
        V.ghost = ghost;
        V.hrec = W.hrec.p;
        V.hrowq = A.halo_rowq.p;
        V.hflag = A.halo_flag.p;
        V.hwq = A.halo_wq.p;
        V.hnq = A.halo_nrows;
        if (A.halo_nrows > 0)
            hipLaunchKernelGGL(k_halo_rec, dim3((A.halo_nrows + kBlock - 1) / kBlock), dim3(kBlock), 0, s,
                               A.halo_nrows, A.halo_ptr.p, A.halo_cols.p, A.halo_vals.p, ghost, W.hrec.p,
                               W.ctrl.p);
    }
    const Prm P{};
    if (fmt == CEL_FMT_CSR && halo)
        hipLaunchKernelGGL((k_pcg_spmv<double, CEL_FMT_CSR, true>), dim3(G1), dim3(kBlock), 0, s, V, x, y,
                           W.part1.p, nullptr, 0, W.ctrl.p, W.ctrl.p + 1, P);
    else if (fmt == CEL_FMT_CSR)
        hipLaunchKernelGGL((k_pcg_spmv<double, CEL_FMT_CSR, false>), dim3(G1), dim3(kBlock), 0, s, V, x, y,
                           W.part1.p, nullptr, 0, W.ctrl.p, W.ctrl.p + 1, P);
    else if (halo)
        hipLaunchKernelGGL((k_pcg_spmv<double, CEL_FMT_DIA, true>), dim3(G1), dim3(kBlock), 0, s, V, x, y,
                           W.part1.p, nullptr, 0, W.ctrl.p, W.ctrl.p + 1, P);
    else
        hipLaunchKernelGGL((k_pcg_spmv<double, CEL_FMT_DIA, false>), dim3(G1), dim3(kBlock), 0, s, V, x, y,
                           W.part1.p, nullptr, 0, W.ctrl.p, W.ctrl.p + 1, P);
    CEL_HIP(hipGetLastError());
    std::vector<double> part((size_t)G1);
    CEL_HIP(hipMemcpyAsync(part.data(), W.part1.p, sizeof(double) * G1, hipMemcpyDeviceToHost, s));
    CEL_HIP(hipStreamSynchronize(s));
    double acc = 0.0;
    for (int b = 0; b < G1; ++b) acc = acc + part[(size_t)b];
    if (dot) *dot = acc;
    return CEL_OK;
}

int pcg_solve(Mat& A, const double* b, double* x, const cel_pcg_params& prm, cel_pcg_stats& stats,
              double* hist, bool noalloc)
{
    CEL_TRY(ensure_init());
    hipStream_t s = rt().stream;
    const int n = A.n;
    const bool mixed = prm.precision == CEL_PREC_MIXED;
    // COO (cel_fop_mv_module.f90:357-403) walks the block's row-sorted
    // entries with the CSR kernels, adding each product to v in entry order
    const bool coo = prm.format == CEL_FMT_COO;
    const int fmt = coo ? CEL_FMT_CSR : prm.format;
    const bool halo = prm.use_halo != 0;
    if (fmt != CEL_FMT_CSR && fmt != CEL_FMT_JAD && fmt != CEL_FMT_DIA)
        return fail(CEL_ERR_ARG, "format must be CSR, COO, JAD or DIA");
    if (fmt == CEL_FMT_JAD && (A.jad_mode < 0 || A.jad_ndef != A.jad_mmax))
        return fail(CEL_ERR_STATE, "JAD requested but the matrix has no complete JAD arrays");
    if (fmt == CEL_FMT_DIA && A.dia_nd == 0)
        return fail(CEL_ERR_STATE, "DIA requested but the matrix has no DIA form");
    if (halo && (mixed || fmt == CEL_FMT_JAD))
        return fail(CEL_ERR_ARG, "halo blocks are applied by the FP64 CSR and DIA forms only");
    if (halo && (double)A.halo_nrows * 16.0 >= 2147483648.0 - 64.0)
        return fail(CEL_ERR_ARG, "halo solve: the halo records need 31-bit offsets (< 2^27 halo rows)");
    if (!A.has_jac) return fail(CEL_ERR_STATE, "cel_gpu_mat_set_jacobi has not been called");
    if (A.ncols != n) return fail(CEL_ERR_ARG, "PCG needs a square local block");
    if (prm.iter_max < 0) return fail(CEL_ERR_ARG, "iter_max < 0");
    if (noalloc) {
        if (mixed || halo || fmt == CEL_FMT_JAD)
            return fail(CEL_ERR_ARG, "the allocation-free solve runs FP64 in the CSR, COO or DIA form, "
                                     "without halo blocks");
        if (prm.memspace != CEL_MEM_DEVICE)
            return fail(CEL_ERR_ARG, "the allocation-free solve takes device pointers");
        if (hist) return fail(CEL_ERR_ARG, "the allocation-free solve records no history");
        if ((size_t)n > kNoAllocMaxRows)
            return fail(CEL_ERR_ARG, "the allocation-free solve holds at most " +
                                         std::to_string(kNoAllocMaxRows) + " rows");
        if (rt().device < 0 || rt().device >= kMaxDevices || !g_scratch[rt().device])
            return fail(CEL_ERR_STATE, "no scratch address for this device (cel_gpu_init)");
    }
    if (n == 0) {  // nothing to solve (the kernels assume at least one row)
        stats = cel_pcg_stats{};
        stats.status = CEL_OK;
        return CEL_OK;
    }

    // a JAD solve runs in the permuted row order (vectors gathered on entry,
    // x scattered back on exit), with the columns renumbered into it
    const bool jadp = fmt == CEL_FMT_JAD;
    if (jadp) {
        // the JAD K1 addresses the entries with unsigned 32-bit byte offsets
        const double lim = 4294967296.0 - 64.0;
        if ((double)A.jad_nnz * (mixed ? 4.0 : 8.0) >= lim || (double)n * 8.0 >= lim)
            return fail(CEL_ERR_ARG, "JAD solve: entries beyond the 32-bit offsets of the JAD kernel");
        CEL_TRY(jad_prepare_permuted(A));
    }

    const int ntiles = std::max(1, (n + kTileRows - 1) / kTileRows);
    // grids: K1 variants that must agree (with and without halo) share one
    // grid; the residual kernels share one; K2 and K3 share one.
    int G1, Gr, Gv;
    // the DIA K1 runs on the CSR K1's grid (the same rows per block, so the
    // same partial sums); the residual kernels of a DIA solve use the CSR form
    const int rfmt = fmt == CEL_FMT_DIA ? CEL_FMT_CSR : fmt;
    if (!mixed) {
        if (fmt == CEL_FMT_DIA || fmt == CEL_FMT_CSR) {
            G1 = common_grid({(const void*)k_pcg_spmv<double, CEL_FMT_CSR, false>,
                              (const void*)k_pcg_spmv<double, CEL_FMT_CSR, true>,
                              (const void*)k_pcg_spmv<double, CEL_FMT_DIA, false>,
                              (const void*)k_pcg_spmv<double, CEL_FMT_DIA, true>}, ntiles);
            Gr = common_grid({(const void*)k_residual<double, CEL_FMT_CSR, false, false>,
                              (const void*)k_residual<double, CEL_FMT_CSR, true, false>}, ntiles);
        } else {
            G1 = common_grid({(const void*)k_pcg_spmv<double, CEL_FMT_JAD, false>}, ntiles);
            Gr = common_grid({(const void*)k_residual<double, CEL_FMT_JAD, false, false>}, ntiles);
        }
    } else {
        if (fmt == CEL_FMT_DIA || fmt == CEL_FMT_CSR) {
            G1 = common_grid({(const void*)k_pcg_spmv<float, CEL_FMT_CSR, false>,
                              (const void*)k_pcg_spmv<float, CEL_FMT_DIA, false>}, ntiles);
            Gr = common_grid({(const void*)k_residual<float, CEL_FMT_CSR, false, false>,
                              (const void*)k_residual<float, CEL_FMT_CSR, false, true>}, ntiles);
        } else {
            G1 = common_grid({(const void*)k_pcg_spmv<float, CEL_FMT_JAD, false>}, ntiles);
            Gr = common_grid({(const void*)k_residual<float, CEL_FMT_JAD, false, false>,
                              (const void*)k_residual<float, CEL_FMT_JAD, false, true>}, ntiles);
        }
    }
    G1 = k1_cap(G1, mixed);
    {
        const long long units = ((long long)n + 2 * kBlock - 1) / (2 * kBlock);
        Gv = (int)std::min<long long>((long long)rt().num_cus * kVecBlocksPerCu,
                                      std::max<long long>(1, units));
        if (Gv >= kXcds) Gv = Gv / kXcds * kXcds;
    }

    if (noalloc && ((size_t)G1 > kNoAllocMaxGrid || (size_t)Gr > kNoAllocMaxGrid ||
                    (size_t)Gv > kNoAllocMaxGrid))
        return fail(CEL_ERR_ARG, "the allocation-free solve holds at most " +
                                     std::to_string(kNoAllocMaxGrid) + " blocks per kernel");

    // matrix data for the mixed inner operator (once per matrix)
    if (mixed) {
        if (fmt == CEL_FMT_DIA) CEL_TRY(dia_prepare_f32(A));
        // the DIA K1 reads dia_vals32 and the residual kernels the FP64 CSR
        // values, so only a CSR solve needs the FP32 CSR copy
        if (fmt == CEL_FMT_CSR && !A.vals32.p) {
            CEL_TRY(A.vals32.alloc(A.nnz > 0 ? A.nnz : 1));
            CEL_TRY(launch_to_f32(A.vals.p, A.vals32.p, A.nnz));
        }
        if (fmt == CEL_FMT_JAD && !A.jad_vals32.p) {
            CEL_TRY(A.jad_vals32.alloc(A.jad_nnz > 0 ? A.jad_nnz : 1));
            CEL_TRY(launch_to_f32(A.jad_vals.p, A.jad_vals32.p, A.jad_nnz));
        }
    }

    // workspace: the matrix handle's buffers, or the scratch array
    const size_t nn = n > 0 ? n : 1;
    Workspace* Wp = noalloc ? nullptr : &ws_of(A);
    WorkPtrs w;
    if (noalloc) {
        double* base = static_cast<double*>(g_scratch[rt().device]);
        w.r = base;
        w.p = base + kNoAllocMaxRows;
        w.v = base + 2 * kNoAllocMaxRows;
        w.part1 = base + 3 * kNoAllocMaxRows;
        w.part2 = w.part1 + kNoAllocMaxGrid;
        w.rpart = w.part2 + 2 * kNoAllocMaxGrid;
        w.ctrl = reinterpret_cast<Ctrl*>(base + kNoAllocDoubles);
    } else {
        Workspace& W = *Wp;
        CEL_TRY(W.part1.reserve((size_t)G1));
        CEL_TRY(W.part2.reserve(2 * (size_t)Gv));
        CEL_TRY(W.rpart.reserve(3 * (size_t)Gr));
        CEL_TRY(W.ctrl.reserve(5));
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
        w.r = W.r.p;
        w.p = W.p.p;
        w.v = W.v.p;
        w.part1 = W.part1.p;
        w.part2 = W.part2.p;
        w.rpart = W.rpart.p;
        w.ctrl = W.ctrl.p;
        w.r32 = W.r32.p;
        w.p32 = W.p32.p;
        w.v32 = W.v32.p;
        w.e32 = W.e32.p;
    }
    const double* db = b;
    double* dx = x;
    if (prm.memspace == CEL_MEM_HOST) {
        Workspace& W = *Wp;
        CEL_TRY(W.b.alloc(nn));
        CEL_TRY(W.x.alloc(nn));
        CEL_HIP(hipMemcpyAsync(W.b.p, b, sizeof(double) * n, hipMemcpyHostToDevice, s));
        if (prm.x0_mode == CEL_X0_USER)
            CEL_HIP(hipMemcpyAsync(W.x.p, x, sizeof(double) * n, hipMemcpyHostToDevice, s));
        db = W.b.p;
        dx = W.x.p;
    }
    double* const dx_out = dx;
    const double* jac64 = A.jac.p;
    const float* jac32 = A.jac32.p;
    if (jadp) {
        Workspace& W = *Wp;
        CEL_TRY(W.bp.alloc(nn));
        CEL_TRY(W.xp.alloc(nn));
        CEL_TRY(W.jacp.alloc(nn));
        if (mixed) CEL_TRY(W.jacp32.alloc(nn));
        jac64 = W.jacp.p;
        jac32 = mixed ? W.jacp32.p : nullptr;
    }
    const int hist_cap = hist ? prm.iter_max + 1 : 0;
    if (hist_cap) CEL_TRY(Wp->hist.alloc(hist_cap));
    if (halo) {
        Workspace& W = *Wp;
        CEL_TRY(W.ghost.alloc(A.ghost_len > 0 ? A.ghost_len : 1));
        CEL_TRY(W.sendbuf.alloc(A.send_total > 0 ? A.send_total : 1));
        hipLaunchKernelGGL(k_fill, dim3(64), dim3(kBlock), 0, s, (long long)W.ghost.n, W.ghost.p, 0.0);
        CEL_TRY(W.hrec.reserve(2 * (size_t)std::max(A.halo_nrows, 1)));
        CEL_TRY(build_hrow_tile(A, W, ntiles));
    }

    Prm P{};
    P.tol = prm.tol;
    P.tol_mode = prm.tol_mode;
    P.iter_max = prm.iter_max;
    P.small = 1.1 * 2.220446049250313e-16;  // cel_real_small (cel_types_module.f90:33)
    P.ru_delta = prm.ru_delta > 0 ? prm.ru_delta : 0.1;
    P.ru_restart = prm.ru_restart;
    P.alpha_stop = prm.alpha_stop;
    P.ru_max_inner = 1000;
    P.hist = hist_cap ? Wp->hist.p : nullptr;
    P.hist_cap = hist_cap;
    Ctrl c0{};
    c0.run = 1;
    // every solve starts from a fresh token (the scratch array keeps the
    // last call's contents; everything else is written before it is read)
    CEL_HIP(hipMemcpyAsync(w.ctrl, &c0, sizeof(Ctrl), hipMemcpyHostToDevice, s));
    // token slots: 0 -> K1 -> 1 -> K2 -> 2 -> K3 -> 0 (FP64);
    // mixed: K3 -> 3 -> replacement residual -> 0 (K3 folds e into x64)
    Ctrl* C0 = w.ctrl;
    Ctrl* C1 = w.ctrl + 1;
    Ctrl* C2 = w.ctrl + 2;
    Ctrl* C3 = w.ctrl + 3;

    MatView V{};
    V.n = n;
    V.coo = coo ? 1 : 0;
    V.mmax = A.jad_mmax;
    V.jnnz = A.jad_nnz;
    V.xadj = A.xadj.p;
    V.cols = A.cols.p;
    V.vals = A.vals.p;
    V.vals32 = A.vals32.p;
    V.jbegin = A.jad_begin.p;
    V.jcols = jadp ? A.jad_pcols.p : A.jad_cols.p;
    V.jperm = A.jad_perm.p;
    V.jvals = A.jad_vals.p;
    V.jvals32 = A.jad_vals32.p;
    V.dvals = A.dia_vals.p;
    V.dvals32 = A.dia_vals32.p;
    V.dmask = A.dia_mask.p;
    V.dmask8 = A.dia_mask8.p;
    V.nd = A.dia_nd;
    V.dzero = -1;
    for (int d = 0; d < A.dia_nd; ++d) {
        V.doff[d] = A.dia_off[d];
        if (A.dia_off[d] == 0) V.dzero = d;
    }
    if (halo) {
        V.hrow_tile = Wp->hrow_tile.p;
        V.hrows = A.halo_rows.p;
        V.hrowq = A.halo_rowq.p;
        V.hptr = A.halo_ptr.p;
        V.hcols = A.halo_cols.p;
        V.hvals = A.halo_vals.p;
        V.ghost = Wp->ghost.p;
        V.hrec = Wp->hrec.p;
        V.hflag = A.halo_flag.p;
        V.hwq = A.halo_wq.p;
        V.hnq = A.halo_nrows;
    }

    // the allocation-free solve creates no events: it times with the host
    // clock around a synchronised stream
    StreamScope R(s);
    hipEvent_t e0 = nullptr, e1 = nullptr, e2 = nullptr;
    const auto t_host0 = std::chrono::steady_clock::now();
    if (!noalloc) {
        CEL_TRY(R.event(e0));
        CEL_TRY(R.event(e1));
        CEL_TRY(R.event(e2));
        CEL_HIP(hipEventRecord(e0, s));
    }
    if (jadp) {
        Workspace& W = *Wp;
        const int gp = std::max(1, std::min(Gv * 2, 4096));
        hipLaunchKernelGGL(k_perm_gather<double>, dim3(gp), dim3(kBlock), 0, s, (long long)n,
                           A.jad_perm.p, db, W.bp.p);
        hipLaunchKernelGGL(k_perm_gather<double>, dim3(gp), dim3(kBlock), 0, s, (long long)n,
                           A.jad_perm.p, A.jac.p, W.jacp.p);
        if (mixed)
            hipLaunchKernelGGL(k_perm_gather<float>, dim3(gp), dim3(kBlock), 0, s, (long long)n,
                               A.jad_perm.p, A.jac32.p, W.jacp32.p);
        if (prm.x0_mode == CEL_X0_USER)
            hipLaunchKernelGGL(k_perm_gather<double>, dim3(gp), dim3(kBlock), 0, s, (long long)n,
                               A.jad_perm.p, dx, W.xp.p);
        CEL_HIP(hipGetLastError());
        db = W.bp.p;
        dx = W.xp.p;
    }

    // x0 (cel_cgalg_module.f90:209 uses 1.0)
    if (prm.x0_mode != CEL_X0_USER)
        hipLaunchKernelGGL(k_fill, dim3(std::max(1, std::min(Gv * 2, 4096))), dim3(kBlock), 0, s,
                           (long long)n, dx, prm.x0_mode == CEL_X0_ONE ? 1.0 : 0.0);
    if (halo) CEL_TRY(launch_face_pack(A, dx, Wp->sendbuf.p));

    // r0, p0 and their partials; the first K1 folds them in.  The set-up
    // reads and writes token slot 0 (only block 0 writes, after its read).
#define RES_ARGS(OT, R, P_)                                                                \
    V, dx, db, jac64, jac32, (OT*)(R), (OT*)(P_), w.e32, w.rpart, C0, C0, P
    if (!mixed) {
        if (rfmt == CEL_FMT_CSR && halo)
            hipLaunchKernelGGL((k_residual<double, CEL_FMT_CSR, true, false>), dim3(Gr), dim3(kBlock), 0, s,
                               RES_ARGS(double, w.r, w.p));
        else if (rfmt == CEL_FMT_CSR)
            hipLaunchKernelGGL((k_residual<double, CEL_FMT_CSR, false, false>), dim3(Gr), dim3(kBlock), 0,
                               s, RES_ARGS(double, w.r, w.p));
        else
            hipLaunchKernelGGL((k_residual<double, CEL_FMT_JAD, false, false>), dim3(Gr), dim3(kBlock), 0,
                               s, RES_ARGS(double, w.r, w.p));
    } else {
        if (rfmt == CEL_FMT_CSR)
            hipLaunchKernelGGL((k_residual<float, CEL_FMT_CSR, false, false>), dim3(Gr), dim3(kBlock), 0, s,
                               RES_ARGS(float, w.r32, w.p32));
        else
            hipLaunchKernelGGL((k_residual<float, CEL_FMT_JAD, false, false>), dim3(Gr), dim3(kBlock), 0, s,
                               RES_ARGS(float, w.r32, w.p32));
    }
#undef RES_ARGS
    CEL_HIP(hipGetLastError());
    if (!noalloc) CEL_HIP(hipEventRecord(e1, s));

    // profiling events around K1
    std::vector<hipEvent_t> pev;
    int pev_used = 0;
    auto prof_event = [&]() -> hipEvent_t {
        if (pev_used == (int)pev.size()) {
            hipEvent_t ev;
            if (R.event(ev) < 0) return nullptr;
            pev.push_back(ev);
        }
        return pev[pev_used++];
    };

    const long long n_ll = n;
    auto enqueue_iter = [&](bool profile) -> int {
        hipEvent_t ea = nullptr, eb = nullptr;
        if (halo) {
            CEL_TRY(launch_face_pack(A, w.p, Wp->sendbuf.p));
            // (a multi-GPU solve exchanges the faces here, then:)
            if (A.halo_nrows > 0)
                hipLaunchKernelGGL(k_halo_rec, dim3((A.halo_nrows + kBlock - 1) / kBlock), dim3(kBlock),
                                   0, s, A.halo_nrows, A.halo_ptr.p, A.halo_cols.p,
                                   A.halo_vals.p, Wp->ghost.p, Wp->hrec.p, C0);
        }
        if (profile) {
            ea = prof_event();
            eb = prof_event();
            if (!ea || !eb) return fail(CEL_ERR_HIP, "hipEventCreate failed");
            CEL_HIP(hipEventRecord(ea, s));
        }
        if (!mixed) {
            if (halo && fmt == CEL_FMT_DIA)
                hipLaunchKernelGGL((k_pcg_spmv<double, CEL_FMT_DIA, true>), dim3(G1), dim3(kBlock), 0, s, V,
                                   w.p, w.v, w.part1, w.rpart, Gr, C0, C1, P);
            else if (halo)
                hipLaunchKernelGGL((k_pcg_spmv<double, CEL_FMT_CSR, true>), dim3(G1), dim3(kBlock), 0, s, V,
                                   w.p, w.v, w.part1, w.rpart, Gr, C0, C1, P);
            else if (fmt == CEL_FMT_CSR)
                hipLaunchKernelGGL((k_pcg_spmv<double, CEL_FMT_CSR, false>), dim3(G1), dim3(kBlock), 0, s, V,
                                   w.p, w.v, w.part1, w.rpart, Gr, C0, C1, P);
            else if (fmt == CEL_FMT_DIA)
                hipLaunchKernelGGL((k_pcg_spmv<double, CEL_FMT_DIA, false>), dim3(G1), dim3(kBlock), 0, s, V,
                                   w.p, w.v, w.part1, w.rpart, Gr, C0, C1, P);
            else
                hipLaunchKernelGGL((k_pcg_spmv<double, CEL_FMT_JAD, false>), dim3(G1), dim3(kBlock), 0, s, V,
                                   w.p, w.v, w.part1, w.rpart, Gr, C0, C1, P);
            if (profile) CEL_HIP(hipEventRecord(eb, s));
            hipLaunchKernelGGL((k_pcg_update_r<double>), dim3(Gv), dim3(kBlock), 0, s, n_ll, w.r, w.v,
                               jac64, w.part2, w.part1, G1, C1, C2);
            hipLaunchKernelGGL((k_pcg_update_xp<double, double, false>), dim3(Gv), dim3(kBlock), 0, s, n_ll,
                               dx, w.p, w.r, jac64, nullptr, w.part2, Gv, C2, C0, P);
        } else {
            if (fmt == CEL_FMT_CSR)
                hipLaunchKernelGGL((k_pcg_spmv<float, CEL_FMT_CSR, false>), dim3(G1), dim3(kBlock), 0, s, V,
                                   w.p32, w.v32, w.part1, w.rpart, Gr, C0, C1, P);
            else if (fmt == CEL_FMT_DIA)
                hipLaunchKernelGGL((k_pcg_spmv<float, CEL_FMT_DIA, false>), dim3(G1), dim3(kBlock), 0, s, V,
                                   w.p32, w.v32, w.part1, w.rpart, Gr, C0, C1, P);
            else
                hipLaunchKernelGGL((k_pcg_spmv<float, CEL_FMT_JAD, false>), dim3(G1), dim3(kBlock), 0, s, V,
                                   w.p32, w.v32, w.part1, w.rpart, Gr, C0, C1, P);
            if (profile) CEL_HIP(hipEventRecord(eb, s));
            hipLaunchKernelGGL((k_pcg_update_r<float>), dim3(Gv), dim3(kBlock), 0, s, n_ll, w.r32, w.v32,
                               jac32, w.part2, w.part1, G1, C1, C2);
            hipLaunchKernelGGL((k_pcg_update_xp<float, float, true>), dim3(Gv), dim3(kBlock), 0, s, n_ll,
                               w.e32, w.p32, w.r32, jac32, dx, w.part2, Gv, C2, C3, P);
            if (rfmt == CEL_FMT_CSR)
                hipLaunchKernelGGL((k_residual<float, CEL_FMT_CSR, false, true>), dim3(Gr), dim3(kBlock), 0, s,
                                   V, dx, db, jac64, jac32, w.r32, w.p32, w.e32, w.rpart, C3,
                                   C0, P);
            else
                hipLaunchKernelGGL((k_residual<float, CEL_FMT_JAD, false, true>), dim3(Gr), dim3(kBlock), 0, s,
                                   V, dx, db, jac64, jac32, w.r32, w.p32, w.e32, w.rpart, C3,
                                   C0, P);
        }
        CEL_HIP(hipGetLastError());
        return CEL_OK;
    };

    // iteration loop: chunks of `chunk` iterations, one chunk in flight
    // ahead of the host's check of the token.  The loop needs at most
    // iter_max iterations plus one K1 that folds in the final partials.
    const int chunk = prm.check_every > 0 ? prm.check_every : 16;
    int* hrun = nullptr;
    int hrun_sync = 1;  // the allocation-free solve reads the flag synchronously
    hipEvent_t cev[2] = {nullptr, nullptr};
    if (!noalloc) {
        CEL_TRY(R.host_alloc(hrun, 2));
        hrun[0] = hrun[1] = 1;
        CEL_TRY(R.event(cev[0], hipEventDisableTiming));
        CEL_TRY(R.event(cev[1], hipEventDisableTiming));
    }

    const bool use_graph = prm.use_graph && !prm.profile && !noalloc;
    int status = CEL_OK;
    const long long budget = (long long)prm.iter_max + 1;
    long long enqueued = 0;
    int k = 0;
    while (enqueued < budget) {
        const int c = (int)std::min<long long>(chunk, budget - enqueued);
        if (use_graph && c == chunk) {
            if (!R.exec) {
                CEL_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
                for (int i = 0; i < c; ++i) {
                    int rc = enqueue_iter(false);
                    if (rc < 0) {
                        (void)hipStreamEndCapture(s, &R.graph);
                        return rc;
                    }
                }
                CEL_HIP(hipStreamEndCapture(s, &R.graph));
                CEL_HIP(hipGraphInstantiate(&R.exec, R.graph, nullptr, nullptr, 0));
            }
            CEL_HIP(hipGraphLaunch(R.exec, s));
        } else {
            for (int i = 0; i < c; ++i) {
                status = enqueue_iter(prm.profile != 0 && !noalloc);
                if (status < 0) break;
            }
            if (status < 0) break;
        }
        enqueued += c;
        if (noalloc) {
            CEL_HIP(hipMemcpyAsync(&hrun_sync, &C0->run, sizeof(int), hipMemcpyDeviceToHost, s));
            CEL_HIP(hipStreamSynchronize(s));
            if (!hrun_sync) break;
            ++k;
            continue;
        }
        CEL_HIP(hipMemcpyAsync(&hrun[k & 1], &C0->run, sizeof(int), hipMemcpyDeviceToHost, s));
        CEL_HIP(hipEventRecord(cev[k & 1], s));
        if (k > 0) {
            CEL_HIP(hipEventSynchronize(cev[(k - 1) & 1]));
            if (!hrun[(k - 1) & 1]) break;
        }
        ++k;
    }
    if (mixed) {
        // a budget that ran out mid-way leaves a correction in e: fold it in
        hipLaunchKernelGGL(k_ru_flush, dim3(Gv), dim3(kBlock), 0, s, n_ll, dx, w.e32, C0, C3, 1);
    }
    if (jadp) {  // x back to the original row order
        hipLaunchKernelGGL(k_perm_scatter<double>, dim3(std::max(1, std::min(Gv * 2, 4096))),
                           dim3(kBlock), 0, s, n_ll, A.jad_perm.p, dx, dx_out);
        CEL_HIP(hipGetLastError());
    }
    if (!noalloc) CEL_HIP(hipEventRecord(e2, s));
    CEL_HIP(hipStreamSynchronize(s));
    if (status < 0) return status;

    Ctrl out;
    CEL_HIP(hipMemcpy(&out, C0, sizeof(Ctrl), hipMemcpyDeviceToHost));
    float ms_setup = 0, ms_all = 0;
    if (!noalloc) {
        CEL_HIP(hipEventElapsedTime(&ms_setup, e0, e1));
        CEL_HIP(hipEventElapsedTime(&ms_all, e0, e2));
    } else {
        ms_all = (float)std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() -
                                                                  t_host0)
                     .count();
    }
    double spmv_ms = 0;
    int spmv_n = 0;
    for (int i = 0; i + 1 < pev_used; i += 2) {
        float t = 0;
        if (hipEventElapsedTime(&t, pev[i], pev[i + 1]) == hipSuccess) {
            spmv_ms += t;
            ++spmv_n;
        }
    }

    if (prm.memspace == CEL_MEM_HOST)
        CEL_HIP(hipMemcpy(x, dx_out, sizeof(double) * n, hipMemcpyDeviceToHost));
    if (hist && hist_cap) {
        int cnt = std::min(out.iter + 1, hist_cap);
        CEL_HIP(hipMemcpy(hist, Wp->hist.p, sizeof(double) * cnt, hipMemcpyDeviceToHost));
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
    if (jadp) stats.spmv_bytes -= 4.0 * n;  // the permuted-order K1 reads no permutation
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
    if (!converged && out.iter < prm.iter_max && prm.alpha_stop && !(std::fabs(out.alpha) > P.small))
        stats.status = CEL_OK;  // stopped on |alpha| <= 1.1 eps (reference condition)
    return stats.status;
}

void pcg_release_workspace(Mat* A)
{
    delete A->ws;
    A->ws = nullptr;
}

}  // namespace cel
