// gmres.hip -- Jacobi-GMRES(m) in HBM, the element solver's default method.
//
// Replaces cel_gmresalg_2jacobi (src/fcgalg/cel_gmresalg_module.f90:570-851)
// with one rank.  The op sequence is the reference's:
//   x = 1 ; while iter <= iter_max and not converged:
//     iter += 1 ; r = b - A x ; beta = ||r||
//     up to m+1 cycles:
//       v1 = r * (1/beta) ; gamma1 = beta
//       for j = 1..m:  z_j = v_j * jac ; w = A z_j
//                      for i = 1..j: h_ij = v_i . w ; w = w - h_ij v_i
//                      h_j+1,j = ||w|| ; rotate column j (as the reference
//                      writes it, :722-727: both statements assign h_ij, so
//                      h_i+1,j is never rotated) ; new rotation ; gamma
//                      if |gamma_j+1| <= eps: x += sum alpha_i z_i, stop
//                      v_j+1 = w / h_j+1,j
//       if not stopped: x += sum alpha_i z_i (back substitution on m)
//       r = b - A x ; beta = ||r||
// Element-wise steps round every product before the add, as the reference's
// loops do.
//
// Everything stays on the device; the host only enqueues.  The scalars (the
// Hessenberg columns, rotations, gamma, back-substitution coefficients,
// residual norms, counters and stop flags) live in one state record in HBM.
// Each vector kernel writes per-block partial sums over fixed contiguous
// chunks; the next kernel sums them in a fixed order (every block that needs
// the value, or a one-block kernel that updates the state), so the result
// depends on n only, never on timing.  In the CSR and DIA forms the products
// w = A z_j and r = b - A x form the partials of v_1 . w and r . r
// themselves, over the same chunks and in the same order as the separate
// dot kernels the JAD form runs after its product, so all three forms give
// the same bits.  The small Hessenberg algebra
// (rotations, convergence test, back substitution) runs in one thread of a
// one-block kernel, in the reference's order.  Partials alternate between
// two buffers, so no kernel writes the partials it reads, and every state
// field a kernel writes is one that no other block of that kernel reads.
//
// One outer iteration (its residual and m+1 cycles) is captured once as a
// hipGraph and replayed up to iter_max+1 times; after a replay the host
// reads the stop flag (one replay in flight).  Once the solve has converged
// every kernel of the rest of the graph returns at once: the kernels of
// later columns and cycles test `halt` (set at convergence), the cycle's
// closing residual and everything after it test `stop` (set after that
// residual).  v_m+1 is not formed (nothing reads it).
#include <cmath>
#include <string>
#include <utility>

#include "cel_internal.hpp"
#include "device_common.hpp"

namespace cel {
namespace {

constexpr int kThreads = 256;
constexpr int kMaxRestart = 32;
constexpr int kM1 = kMaxRestart + 2;  // 1-based index range of the small arrays

struct GmState {
    double H[kM1 * kM1];  // H(i, j) = H[i * kM1 + j], 1-based, rotated in place
    double cs[kM1], sn[kM1], gam[kM1], alp[kM1];
    double residuum, local_residuum, norm0, eps;
    int iter, cycles, ncols;  // outer iterations, cycles, Arnoldi columns run
    int conv_col;             // column at which the current cycle converged (0: none)
    int nalp;                 // coefficients in alp for the pending x update
    int halt;                 // converged or stopped: later columns and cycles do nothing
    int stop;                 // nothing runs any more
    int status;               // 0 or CEL_ERR_NONFINITE
};

__device__ __forceinline__ double& Hij(GmState* st, int i, int j) { return st->H[i * kM1 + j]; }

// Per-block partial of a sum, written to part[blockIdx].
__device__ __forceinline__ void put_partial(double acc, double* __restrict__ part, double* red)
{
    const double s = block_sum(acc, red);  // valid in thread 0
    if (threadIdx.x == 0) part[blockIdx.x] = s;
}

// Sum of the G partials in a fixed order, returned to every thread.
__device__ __forceinline__ double fold(const double* __restrict__ part, int G, double* red,
                                       double* bcast)
{
    const double s = sum_partials(part, G, red);
    if (threadIdx.x == 0) *bcast = s;
    __syncthreads();
    const double v = *bcast;
    __syncthreads();
    return v;
}

#define CHUNK_LOOP(i)                                                        \
    const long long c0 = (long long)blockIdx.x * chunk;                      \
    const long long c1 = c0 + chunk < n ? c0 + chunk : n;                    \
    for (long long i = c0 + threadIdx.x; i < c1; i += kThreads)

__global__ void __launch_bounds__(kThreads)
k_fill(long long n, long long chunk, double* __restrict__ x, double v)
{
    CHUNK_LOOP(i) x[i] = v;
}

// Zeroes the state record (all threads of one block, 8 bytes each) and sets eps.
static_assert(sizeof(GmState) % 8 == 0, "GmState is cleared in 8-byte words");
__global__ void __launch_bounds__(kThreads) k_state_init(GmState* st, double eps)
{
    unsigned long long* w = reinterpret_cast<unsigned long long*>(st);
    for (int i = threadIdx.x; i < (int)(sizeof(GmState) / 8); i += kThreads) w[i] = 0ull;
    __syncthreads();
    if (threadIdx.x == 0) st->eps = eps;
}

// r = b - r ; partial r . r
__global__ void __launch_bounds__(kThreads)
k_resid(long long n, long long chunk, const double* __restrict__ b, double* __restrict__ r,
        double* __restrict__ part, const GmState* __restrict__ st)
{
    __shared__ double red[kWaves];
    if (st->stop) return;
    double acc = 0.0;
    CHUNK_LOOP(i) {
        const double ri = b[i] - r[i];
        r[i] = ri;
        const double p = ri * ri;
        acc = acc + p;
    }
    put_partial(acc, part, red);
}

// Start of an outer iteration (one block): iter += 1 ; residuum = ||r||
// (the first one is norm0) ; a non-finite residual ends the solve.
__global__ void __launch_bounds__(kThreads)
k_outer_start(const double* __restrict__ part, int G, GmState* st)
{
    __shared__ double red[kWaves];
    __shared__ double bcast;
    if (st->stop) return;
    const double s = fold(part, G, red, &bcast);
    if (threadIdx.x == 0) {
        const double res = sqrt(s);
        st->iter = st->iter + 1;
        if (st->iter == 1) st->norm0 = res;
        st->residuum = res;
        st->local_residuum = res;
        if (!isfinite(res)) {
            st->status = CEL_ERR_NONFINITE;
            st->halt = 1;
            st->stop = 1;
        }
    }
}

// End of a cycle (one block): residuum = ||r|| ; cycles += 1 ; a cycle that
// converged ends the solve here.
__global__ void __launch_bounds__(kThreads)
k_cycle_end(const double* __restrict__ part, int G, GmState* st)
{
    __shared__ double red[kWaves];
    __shared__ double bcast;
    if (st->stop) return;
    const double s = fold(part, G, red, &bcast);
    if (threadIdx.x == 0) {
        const double res = sqrt(s);
        st->residuum = res;
        st->local_residuum = res;
        st->cycles = st->cycles + 1;
        if (st->conv_col) st->stop = 1;
    }
}

// v1 = r * (1/beta) ; z1 = v1 * jac ; gamma1 = beta
__global__ void __launch_bounds__(kThreads)
k_scale_jac(long long n, long long chunk, const double* __restrict__ r,
            const double* __restrict__ jac, double* __restrict__ v, double* __restrict__ z,
            GmState* st)
{
    if (st->halt) return;
    const double tmp = 1.0 / st->local_residuum;
    CHUNK_LOOP(i) {
        const double vi = r[i] * tmp;
        v[i] = vi;
        z[i] = vi * jac[i];
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) st->gam[1] = st->residuum;
}

// partial a . b
__global__ void __launch_bounds__(kThreads)
k_dot(long long n, long long chunk, const double* __restrict__ a, const double* __restrict__ b,
      double* __restrict__ part, const GmState* __restrict__ st)
{
    __shared__ double red[kWaves];
    if (st->halt) return;
    double acc = 0.0;
    CHUNK_LOOP(i) {
        const double p = a[i] * b[i];
        acc = acc + p;
    }
    put_partial(acc, part, red);
}

// h_ii,jj = sum(prev partials) ; w = w - h vi ; partial (next ? next . w : w . w)
__global__ void __launch_bounds__(kThreads)
k_mgs(long long n, long long chunk, int ii, int jj, const double* __restrict__ prev, int G,
      const double* __restrict__ vi, double* __restrict__ w, const double* __restrict__ next,
      double* __restrict__ part, GmState* st)
{
    __shared__ double red[kWaves];
    __shared__ double bcast;
    if (st->halt) return;
    const double h = fold(prev, G, red, &bcast);
    if (blockIdx.x == 0 && threadIdx.x == 0) Hij(st, ii, jj) = h;
    double acc = 0.0;
    CHUNK_LOOP(i) {
        const double t = h * vi[i];
        const double wi = w[i] - t;
        w[i] = wi;
        const double p = (next ? next[i] : wi) * wi;
        acc = acc + p;
    }
    put_partial(acc, part, red);
}

// Back substitution on the first k columns (cel_gmresalg_module.f90:760-772
// order), one thread.
__device__ void back_substitute(GmState* st, int k)
{
    for (int ii = k; ii >= 1; --ii) {
        double a = 0.0;
        for (int kk = ii + 1; kk <= k; ++kk) {
            const double t = Hij(st, ii, kk) * st->alp[kk];
            a = a + t;
        }
        const double inv = 1.0 / Hij(st, ii, ii);
        st->alp[ii] = inv * (st->gam[ii] - a);
    }
    st->nalp = k;
}

// One block: h_jj+1,jj = sqrt(sum of the last partials) ; the column's
// rotations as the reference writes them ; new rotation ; gamma ;
// convergence test ; the back-substitution coefficients when the column
// converged or is the cycle's last.
__global__ void __launch_bounds__(kThreads)
k_givens(int jj, int m, const double* __restrict__ part, int G, GmState* st)
{
    __shared__ double red[kWaves];
    __shared__ double bcast;
    if (st->halt) return;
    const double s = fold(part, G, red, &bcast);
    if (threadIdx.x != 0) return;
    Hij(st, jj + 1, jj) = sqrt(s);
    for (int ii = 1; ii <= jj - 1; ++ii) {  // both statements assign h(ii,jj) (:722-727)
        double a = Hij(st, ii, jj) * st->cs[ii];
        double c = Hij(st, ii + 1, jj) * st->sn[ii];
        Hij(st, ii, jj) = a + c;
        a = -Hij(st, ii, jj) * st->sn[ii];
        c = Hij(st, ii + 1, jj) * st->cs[ii];
        Hij(st, ii, jj) = a + c;
    }
    {
        const double a = Hij(st, jj, jj) * Hij(st, jj, jj);
        const double c = Hij(st, jj + 1, jj) * Hij(st, jj + 1, jj);
        const double betta = sqrt(a + c);
        st->sn[jj] = Hij(st, jj + 1, jj) / betta;
        st->cs[jj] = Hij(st, jj, jj) / betta;
        Hij(st, jj, jj) = betta;
        st->gam[jj + 1] = -st->sn[jj] * st->gam[jj];
        st->gam[jj] = st->cs[jj] * st->gam[jj];
    }
    st->ncols = st->ncols + 1;
    st->nalp = 0;
    if (fabs(st->gam[jj + 1]) <= st->eps) {
        back_substitute(st, jj);
        st->conv_col = jj;
        st->halt = 1;
    } else if (jj == m) {
        back_substitute(st, m);
    }
}

// After column jj: with coefficients pending (the column converged or is the
// cycle's last) x = x + a_1 z_1 + ... + a_k z_k, added in that order; else
// v_jj+1 = w / h_jj+1,jj and z_jj+1 = v_jj+1 * jac.
__global__ void __launch_bounds__(kThreads)
k_next(long long n, long long chunk, int jj, int m, const double* __restrict__ w,
       const double* __restrict__ jac, double* __restrict__ V, double* __restrict__ Z,
       double* __restrict__ x, const GmState* __restrict__ st)
{
    if (st->stop) return;
    if (st->halt && st->conv_col != jj) return;  // converged at an earlier column
    const int k = st->nalp;
    if (k > 0) {
        CHUNK_LOOP(i) {
            double xi = x[i];
            for (int j = 0; j < k; ++j) {
                const double t = st->alp[j + 1] * Z[(long long)j * n + i];
                xi = xi + t;
            }
            x[i] = xi;
        }
        return;
    }
    if (jj == m) return;
    const double h = st->H[(jj + 1) * kM1 + jj];
    double* v = V + (long long)jj * n;  // v_jj+1
    double* z = Z + (long long)jj * n;  // z_jj+1
    CHUNK_LOOP(i) {
        const double vi = w[i] / h;
        v[i] = vi;
        z[i] = vi * jac[i];
    }
}

struct Gmres {
    const Mat& A;
    int format;
    long long n, chunk;
    int G;
    hipStream_t st;

    Gmres(const Mat& M, int fmt) : A(M), format(fmt), n(M.n), st(rt().stream)
    {
        // fixed chunks of a multiple of the workgroup size: the grouping of
        // every dot product depends on n only (2048 rows per chunk up to
        // 1024 chunks: one 8-row step per thread of the fused DIA product)
        long long g = (n + 2047) / 2048;
        G = (int)(g < 1 ? 1 : g > 1024 ? 1024 : g);
        chunk = ((n + G - 1) / G + kThreads - 1) / kThreads * kThreads;
        if (chunk < kThreads) chunk = kThreads;
    }
    // w = A z with the partials of v1 . w (skipped once halt is set)
    int spmv_dot(const double* z, double* w, const double* v1, double* part, GmState* gs)
    {
        const int* skip = &gs->halt;
        if (format == CEL_FMT_DIA)
            return launch_dia_spmv_fused(A, kFuseDot, chunk, G, z, w, v1, part, skip);
        if (format == CEL_FMT_CSR)
            return launch_csr_spmv_fused(A, kFuseDot, chunk, G, z, w, v1, part, skip);
        CEL_TRY(launch_jad_spmv(A, z, w, 0, skip));
        k_dot<<<G, kThreads, 0, st>>>(n, chunk, v1, w, part, gs);
        return CEL_OK;
    }
    // r = b - A x with the partials of r . r (skipped once stop is set)
    int spmv_resid(const double* x, double* r, const double* b, double* part, GmState* gs)
    {
        const int* skip = &gs->stop;
        if (format == CEL_FMT_DIA)
            return launch_dia_spmv_fused(A, kFuseResid, chunk, G, x, r, b, part, skip);
        if (format == CEL_FMT_CSR)
            return launch_csr_spmv_fused(A, kFuseResid, chunk, G, x, r, b, part, skip);
        CEL_TRY(launch_jad_spmv(A, x, r, 0, skip));
        k_resid<<<G, kThreads, 0, st>>>(n, chunk, b, r, part, gs);
        return CEL_OK;
    }
};

}  // namespace

int gmres_solve(Mat& A, const double* b_in, double* x_out, const cel_gmres_params& prm,
                cel_pcg_stats& stats)
{
    const int m = prm.restart;
    if (m < 1 || m > kMaxRestart) return fail(CEL_ERR_ARG, "restart must be 1..32");
    if (prm.iter_max < 0) return fail(CEL_ERR_ARG, "iter_max must be >= 0");
    if (!A.has_jac) return fail(CEL_ERR_STATE, "cel_gpu_mat_set_jacobi has not been called");
    if (prm.format == CEL_FMT_JAD && A.jad_mode < 0)
        return fail(CEL_ERR_STATE, "JAD requested before cel_gpu_mat_to_jad");
    if (prm.format == CEL_FMT_JAD && A.jad_ndef < A.jad_mmax)
        return fail(CEL_ERR_REF_UNDEFINED, "the JAD form of this matrix is undefined in the reference");
    if (prm.format == CEL_FMT_DIA && A.dia_nd == 0)
        return fail(CEL_ERR_STATE, "DIA requested before cel_gpu_mat_to_dia");
    if (A.n != A.ncols) return fail(CEL_ERR_ARG, "GMRES needs a square local block");
    const long long n = A.n;
    stats = cel_pcg_stats{};
    if (n == 0) return CEL_OK;
    Gmres S(A, prm.format);
    hipStream_t st = S.st;
    const int G = S.G;
    const long long ck = S.chunk;
    DevBuf<double> xb, bb, r, w, V, Z, partA, partB;
    DevBuf<GmState> state;
    CEL_TRY(r.alloc(n));
    CEL_TRY(w.alloc(n));
    CEL_TRY(V.alloc((size_t)n * m));
    CEL_TRY(Z.alloc((size_t)n * m));
    CEL_TRY(partA.alloc(G));
    CEL_TRY(partB.alloc(G));
    CEL_TRY(state.alloc(1));
    const double* b = b_in;
    double* x = x_out;
    if (prm.memspace == CEL_MEM_HOST) {
        CEL_TRY(bb.alloc(n));
        CEL_TRY(xb.alloc(n));
        CEL_HIP(hipMemcpyAsync(bb.p, b_in, sizeof(double) * n, hipMemcpyHostToDevice, st));
        if (prm.x0_mode == CEL_X0_USER)
            CEL_HIP(hipMemcpyAsync(xb.p, x_out, sizeof(double) * n, hipMemcpyHostToDevice, st));
        b = bb.p;
        x = xb.p;
    }
    GmState* gs = state.p;
    const int* stop = &gs->stop;
    auto vv = [&](int k) { return V.p + (size_t)(k - 1) * n; };  // v_k, 1-based
    auto zz = [&](int k) { return Z.p + (size_t)(k - 1) * n; };

    // one outer iteration: r = b - A x, then up to m+1 cycles
    auto enqueue_outer = [&]() -> int {
        CEL_TRY(S.spmv_resid(x, r.p, b, partA.p, gs));
        k_outer_start<<<1, kThreads, 0, st>>>(partA.p, G, gs);
        for (int c = 0; c <= m; ++c) {
            k_scale_jac<<<G, kThreads, 0, st>>>(n, ck, r.p, A.jac.p, vv(1), zz(1), gs);
            for (int jj = 1; jj <= m; ++jj) {
                double* cur = partA.p;
                double* nxt = partB.p;
                CEL_TRY(S.spmv_dot(zz(jj), w.p, vv(1), cur, gs));
                for (int ii = 1; ii <= jj; ++ii) {
                    const double* next = ii < jj ? vv(ii + 1) : nullptr;
                    k_mgs<<<G, kThreads, 0, st>>>(n, ck, ii, jj, cur, G, vv(ii), w.p, next, nxt, gs);
                    std::swap(cur, nxt);
                }
                k_givens<<<1, kThreads, 0, st>>>(jj, m, cur, G, gs);
                k_next<<<G, kThreads, 0, st>>>(n, ck, jj, m, w.p, A.jac.p, V.p, Z.p, x, gs);
            }
            CEL_TRY(S.spmv_resid(x, r.p, b, partA.p, gs));
            k_cycle_end<<<1, kThreads, 0, st>>>(partA.p, G, gs);
        }
        CEL_HIP(hipGetLastError());
        return CEL_OK;
    };

    StreamScope R(st);
    hipEvent_t e0, e1, cev[2];
    CEL_TRY(R.event(e0));
    CEL_TRY(R.event(e1));
    CEL_TRY(R.event(cev[0], hipEventDisableTiming));
    CEL_TRY(R.event(cev[1], hipEventDisableTiming));
    int* hstop = nullptr;
    CEL_TRY(R.host_alloc(hstop, 2));
    hstop[0] = hstop[1] = 0;

    CEL_HIP(hipEventRecord(e0, st));
    k_state_init<<<1, kThreads, 0, st>>>(gs, prm.eps);
    if (prm.x0_mode == CEL_X0_ONE || prm.x0_mode == CEL_X0_ZERO)
        k_fill<<<G, kThreads, 0, st>>>(n, ck, x, prm.x0_mode == CEL_X0_ONE ? 1.0 : 0.0);
    CEL_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
    {
        const int rc = enqueue_outer();
        const hipError_t ec = hipStreamEndCapture(st, &R.graph);
        if (rc < 0) return rc;
        CEL_HIP(ec);
    }
    CEL_HIP(hipGraphInstantiate(&R.exec, R.graph, nullptr, nullptr, 0));
    // outer iterations: the reference runs them while iter <= iter_max
    // (iter counted from 1), so at most iter_max + 1
    const long long budget = (long long)prm.iter_max + 1;
    for (long long k = 0; k < budget; ++k) {
        CEL_HIP(hipGraphLaunch(R.exec, st));
        CEL_HIP(hipMemcpyAsync(&hstop[k & 1], stop, sizeof(int), hipMemcpyDeviceToHost, st));
        CEL_HIP(hipEventRecord(cev[k & 1], st));
        if (k > 0) {
            CEL_HIP(hipEventSynchronize(cev[(k - 1) & 1]));
            if (hstop[(k - 1) & 1]) break;
        }
    }
    CEL_HIP(hipEventRecord(e1, st));
    if (prm.memspace == CEL_MEM_HOST)
        CEL_HIP(hipMemcpyAsync(x_out, x, sizeof(double) * n, hipMemcpyDeviceToHost, st));
    CEL_HIP(hipStreamSynchronize(st));
    GmState out{};
    CEL_HIP(hipMemcpy(&out, gs, sizeof(GmState), hipMemcpyDeviceToHost));
    float ms = 0.f;
    CEL_HIP(hipEventElapsedTime(&ms, e0, e1));

    stats.iters = out.iter;
    stats.outer_iters = out.cycles;
    stats.norm0 = out.norm0;
    stats.final_norm = out.residuum;
    stats.eps_abs = prm.eps;
    stats.solve_ms = ms;
    // products run: one per outer residual, per cycle residual, per column
    stats.spmv_launches = out.iter + out.cycles + out.ncols;
    if (out.status != 0) {
        stats.status = out.status;
        return fail(out.status, "the residual is not finite");
    }
    stats.status = out.conv_col ? CEL_OK : CEL_NOT_CONVERGED;
    return stats.status;
}

}  // namespace cel

using namespace cel;

extern "C" void cel_gpu_gmres_params_default(cel_gmres_params* p)
{
    if (!p) return;
    // the element solver's settings (src/felmsolver/cel_elmsolver_module.f90:140-151)
    p->eps = 0.1e-07;
    p->iter_max = 5000;
    p->restart = 4 - 1;
    p->x0_mode = CEL_X0_ONE;
    p->memspace = CEL_MEM_HOST;
    p->format = CEL_FMT_CSR;
}

extern "C" int cel_gpu_gmres_solve(cel_gpu_mat* A, const double* b, double* x,
                                   const cel_gmres_params* params, cel_pcg_stats* stats)
{
    CEL_TRY(ensure_init());
    if (!A || !b || !x) return fail(CEL_ERR_ARG, "null argument to cel_gpu_gmres_solve");
    cel_gmres_params p;
    if (params) p = *params;
    else cel_gpu_gmres_params_default(&p);
    if ((p.memspace != CEL_MEM_HOST && p.memspace != CEL_MEM_DEVICE) ||
        (p.format != CEL_FMT_CSR && p.format != CEL_FMT_JAD && p.format != CEL_FMT_DIA) ||
        (p.x0_mode != CEL_X0_ONE && p.x0_mode != CEL_X0_ZERO && p.x0_mode != CEL_X0_USER) ||
        !(p.eps >= 0.0))
        return fail(CEL_ERR_ARG, "bad cel_gmres_params");
    cel_pcg_stats s{};
    int rc = gmres_solve(A->m, b, x, p, s);
    if (stats) *stats = s;
    return rc;
}
