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
// loops do; the vector kernels are single passes over HBM with the
// Gram-Schmidt update of w fused with the next dot product (v_i+1 . w, or
// ||w|| after the last) and z_j formed where v_j is.  Dot products are
// per-workgroup partial sums over fixed contiguous chunks, added on the host
// in workgroup order: the result does not depend on timing, only on n.  The
// small Hessenberg algebra runs on the host in the reference's order.
#include "cel_internal.hpp"

namespace cel {
namespace {

constexpr int kThreads = 256;
constexpr int kMaxRestart = 32;

struct Coefs {
    double a[kMaxRestart];
};

// Per-workgroup partial of sum(a .* b) over [blk*chunk, min(n, (blk+1)*chunk)).
__device__ inline void block_partial(double acc, double* __restrict__ part)
{
    __shared__ double red[kThreads];
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int s = kThreads / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) red[threadIdx.x] = red[threadIdx.x] + red[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = red[0];
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

// r = b - r ; partial r . r
__global__ void __launch_bounds__(kThreads)
k_resid(long long n, long long chunk, const double* __restrict__ b, double* __restrict__ r,
        double* __restrict__ part)
{
    double acc = 0.0;
    CHUNK_LOOP(i) {
        const double ri = b[i] - r[i];
        r[i] = ri;
        const double p = ri * ri;
        acc = acc + p;
    }
    block_partial(acc, part);
}

// v = r * tmp ; z = v * jac
__global__ void __launch_bounds__(kThreads)
k_scale_jac(long long n, long long chunk, const double* __restrict__ r, double tmp,
            const double* __restrict__ jac, double* __restrict__ v, double* __restrict__ z)
{
    CHUNK_LOOP(i) {
        const double vi = r[i] * tmp;
        v[i] = vi;
        z[i] = vi * jac[i];
    }
}

// v = w / h ; z = v * jac
__global__ void __launch_bounds__(kThreads)
k_div_jac(long long n, long long chunk, const double* __restrict__ w, double h,
          const double* __restrict__ jac, double* __restrict__ v, double* z)
{
    CHUNK_LOOP(i) {
        const double vi = w[i] / h;
        v[i] = vi;
        if (z) z[i] = vi * jac[i];
    }
}

// partial a . b
__global__ void __launch_bounds__(kThreads)
k_dot(long long n, long long chunk, const double* __restrict__ a, const double* __restrict__ b,
      double* __restrict__ part)
{
    double acc = 0.0;
    CHUNK_LOOP(i) {
        const double p = a[i] * b[i];
        acc = acc + p;
    }
    block_partial(acc, part);
}

// w = w - h vi ; partial (next ? next . w : w . w)
__global__ void __launch_bounds__(kThreads)
k_mgs(long long n, long long chunk, double h, const double* __restrict__ vi,
      double* __restrict__ w, const double* __restrict__ next, double* __restrict__ part)
{
    double acc = 0.0;
    CHUNK_LOOP(i) {
        const double t = h * vi[i];
        const double wi = w[i] - t;
        w[i] = wi;
        const double p = (next ? next[i] : wi) * wi;
        acc = acc + p;
    }
    block_partial(acc, part);
}

// x = x + a_1 z_1 + ... + a_k z_k, added in that order
__global__ void __launch_bounds__(kThreads)
k_xupdate(long long n, long long chunk, int k, Coefs c, const double* __restrict__ Z,
          long long ld, double* __restrict__ x)
{
    CHUNK_LOOP(i) {
        double xi = x[i];
        for (int j = 0; j < k; ++j) {
            const double t = c.a[j] * Z[(long long)j * ld + i];
            xi = xi + t;
        }
        x[i] = xi;
    }
}

struct Gmres {
    const Mat& A;
    int format;
    long long n, chunk;
    int G;
    hipStream_t st;
    DevBuf<double> part;
    double* hpart = nullptr;
    int launches = 0;

    Gmres(const Mat& M, int fmt) : A(M), format(fmt), n(M.n), st(rt().stream)
    {
        // fixed chunks of a multiple of the workgroup size: the grouping of
        // every dot product depends on n only
        long long g = (n + 4095) / 4096;
        G = (int)(g < 1 ? 1 : g > 1024 ? 1024 : g);
        chunk = ((n + G - 1) / G + kThreads - 1) / kThreads * kThreads;
        if (chunk < kThreads) chunk = kThreads;
    }
    ~Gmres()
    {
        if (hpart) (void)hipHostFree(hpart);
    }
    int init()
    {
        CEL_TRY(part.alloc(G));
        CEL_HIP(hipHostMalloc(reinterpret_cast<void**>(&hpart), sizeof(double) * G, hipHostMallocDefault));
        return CEL_OK;
    }
    // sum of the workgroup partials in workgroup order
    int reduce(double& out)
    {
        CEL_HIP(hipGetLastError());
        CEL_HIP(hipMemcpyAsync(hpart, part.p, sizeof(double) * G, hipMemcpyDeviceToHost, st));
        CEL_HIP(hipStreamSynchronize(st));
        double s = 0.0;
        for (int i = 0; i < G; ++i) s = s + hpart[i];
        out = s;
        return CEL_OK;
    }
    int spmv(const double* x, double* y)
    {
        ++launches;
        if (format == CEL_FMT_JAD) return launch_jad_spmv(A, x, y, 0);
        if (format == CEL_FMT_DIA) return launch_dia_spmv(A, x, y, 0);
        return launch_csr_spmv(A, x, y, 0);
    }
    // r = b - A x, returns ||r||
    int residual(const double* b, const double* x, double* r, double& norm)
    {
        CEL_TRY(spmv(x, r));
        k_resid<<<G, kThreads, 0, st>>>(n, chunk, b, r, part.p);
        double s;
        CEL_TRY(reduce(s));
        norm = std::sqrt(s);
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
    CEL_TRY(S.init());
    hipStream_t st = S.st;
    const int G = S.G;
    const long long ck = S.chunk;
    DevBuf<double> xb, bb, r, w, V, Z;
    CEL_TRY(r.alloc(n));
    CEL_TRY(w.alloc(n));
    CEL_TRY(V.alloc((size_t)n * (m + 1)));
    CEL_TRY(Z.alloc((size_t)n * m));
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
    auto vv = [&](int k) { return V.p + (size_t)(k - 1) * n; };  // v_k, 1-based
    auto zz = [&](int k) { return Z.p + (size_t)(k - 1) * n; };
    hipEvent_t e0, e1;
    CEL_HIP(hipEventCreate(&e0));
    CEL_HIP(hipEventCreate(&e1));
    CEL_HIP(hipEventRecord(e0, st));
    if (prm.x0_mode == CEL_X0_ONE || prm.x0_mode == CEL_X0_ZERO)
        k_fill<<<G, kThreads, 0, st>>>(n, ck, x, prm.x0_mode == CEL_X0_ONE ? 1.0 : 0.0);

    // 1-based small arrays as the reference's
    const int M1 = m + 2;
    std::vector<double> hh((size_t)M1 * M1, 0.0), cc(M1, 0.0), ss(M1, 0.0), gam(M1, 0.0), alp(M1, 0.0);
    auto H = [&](int i, int j) -> double& { return hh[(size_t)i * M1 + j]; };
    int conv = 0, iter = 0, cycles = 0;
    double residuum = 0.0;
    bool first = true;
    while (iter <= prm.iter_max && conv == 0) {
        iter += 1;
        CEL_TRY(S.residual(b, x, r.p, residuum));
        if (first) stats.norm0 = residuum;
        first = false;
        if (!std::isfinite(residuum)) {
            stats.final_norm = residuum;
            return fail(CEL_ERR_NONFINITE, "the residual is not finite");
        }
        gam[1] = residuum;
        double local_residuum = residuum;
        int restart_count = 0;
        while (restart_count <= m && conv == 0) {
            const double tmp = 1.0 / local_residuum;
            k_scale_jac<<<G, kThreads, 0, st>>>(n, ck, r.p, tmp, A.jac.p, vv(1), zz(1));
            gam[1] = residuum;
            for (int jj = 1; jj <= m; ++jj) {
                CEL_TRY(S.spmv(zz(jj), w.p));
                k_dot<<<G, kThreads, 0, st>>>(n, ck, vv(1), w.p, S.part.p);
                CEL_TRY(S.reduce(H(1, jj)));
                for (int ii = 1; ii <= jj; ++ii) {
                    const double* next = ii < jj ? vv(ii + 1) : nullptr;
                    k_mgs<<<G, kThreads, 0, st>>>(n, ck, H(ii, jj), vv(ii), w.p, next, S.part.p);
                    double s;
                    CEL_TRY(S.reduce(s));
                    if (ii < jj) H(ii + 1, jj) = s;
                    else H(jj + 1, jj) = std::sqrt(s);
                }
                for (int ii = 1; ii <= jj - 1; ++ii) {  // as the reference writes it
                    double a = H(ii, jj) * cc[ii];
                    double c = H(ii + 1, jj) * ss[ii];
                    H(ii, jj) = a + c;
                    a = -H(ii, jj) * ss[ii];
                    c = H(ii + 1, jj) * cc[ii];
                    H(ii, jj) = a + c;
                }
                {
                    const double a = H(jj, jj) * H(jj, jj);
                    const double c = H(jj + 1, jj) * H(jj + 1, jj);
                    const double betta = std::sqrt(a + c);
                    ss[jj] = H(jj + 1, jj) / betta;
                    cc[jj] = H(jj, jj) / betta;
                    H(jj, jj) = betta;
                    gam[jj + 1] = -ss[jj] * gam[jj];
                    gam[jj] = cc[jj] * gam[jj];
                }
                if (std::fabs(gam[jj + 1]) <= prm.eps) {
                    Coefs c{};
                    for (int ii = jj; ii >= 1; --ii) {
                        alp[ii] = 0.0;
                        for (int kk = ii + 1; kk <= jj; ++kk) {
                            const double t = H(ii, kk) * alp[kk];
                            alp[ii] = alp[ii] + t;
                        }
                        const double inv = 1.0 / H(ii, ii);
                        alp[ii] = inv * (gam[ii] - alp[ii]);
                    }
                    for (int ii = 1; ii <= jj; ++ii) c.a[ii - 1] = alp[ii];
                    k_xupdate<<<G, kThreads, 0, st>>>(n, ck, jj, c, Z.p, n, x);
                    conv = 2;
                    break;
                }
                // z_m+1 is never used
                k_div_jac<<<G, kThreads, 0, st>>>(n, ck, w.p, H(jj + 1, jj), A.jac.p, vv(jj + 1),
                                                  jj < m ? zz(jj + 1) : nullptr);
            }
            if (conv == 0) {
                Coefs c{};
                for (int ii = m; ii >= 1; --ii) {
                    alp[ii] = 0.0;
                    for (int kk = ii + 1; kk <= m; ++kk) {
                        const double t = H(ii, kk) * alp[kk];
                        alp[ii] = alp[ii] + t;
                    }
                    const double inv = 1.0 / H(ii, ii);
                    alp[ii] = inv * (gam[ii] - alp[ii]);
                }
                for (int ii = 1; ii <= m; ++ii) c.a[ii - 1] = alp[ii];
                k_xupdate<<<G, kThreads, 0, st>>>(n, ck, m, c, Z.p, n, x);
            }
            CEL_HIP(hipGetLastError());
            CEL_TRY(S.residual(b, x, r.p, residuum));
            local_residuum = residuum;
            cycles += 1;
            restart_count += 1;
        }
    }
    CEL_HIP(hipEventRecord(e1, st));
    if (prm.memspace == CEL_MEM_HOST)
        CEL_HIP(hipMemcpyAsync(x_out, x, sizeof(double) * n, hipMemcpyDeviceToHost, st));
    CEL_HIP(hipStreamSynchronize(st));
    float ms = 0.f;
    CEL_HIP(hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    stats.iters = iter;
    stats.outer_iters = cycles;
    stats.final_norm = residuum;
    stats.eps_abs = prm.eps;
    stats.solve_ms = ms;
    stats.spmv_launches = S.launches;
    stats.status = conv == 2 ? CEL_OK : CEL_NOT_CONVERGED;
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
