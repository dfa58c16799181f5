/*
 * cel_oracle.c -- CPU restatement of the CEL Jacobi-PCG hot path.
 *
 * TEST INFRASTRUCTURE ONLY (results to check against, and the CPU baseline); the product
 * library never calls into this file.  See cel_oracle.h for conventions.
 * Build: oracle/Makefile (gcc -O2 -fopenmp -ffp-contract=off).
 */
#include "cel_oracle.h"

#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

double orc_real_small(void) { return 1.1 * DBL_EPSILON; }

/* cel_omp_shared_work_module.f90:183-230.  Chunks are multiples of
 * AW_SMALLEST_OMP_TASK elements; the last worker also takes the remainder
 * of the 4-aligned blocks and the unaligned tail. */
void orc_work_distr(int length, int num_workers, int worker, int* start, int* end)
{
    int aligned_len = (length / ORC_AW_SMALLEST_OMP_TASK) * ORC_AW_SMALLEST_OMP_TASK;
    int tail = length - aligned_len;
    int nblocks = aligned_len / ORC_AW_SMALLEST_OMP_TASK;
    int per_worker = nblocks / num_workers;
    int spare_blocks = nblocks - per_worker * num_workers;
    int s = per_worker * ORC_AW_SMALLEST_OMP_TASK * (worker - 1) + 1;
    int e = s + per_worker * ORC_AW_SMALLEST_OMP_TASK - 1;
    /* worker/num_workers is 1 only for the last worker (integer division) */
    e += (worker / num_workers) * (tail + spare_blocks * ORC_AW_SMALLEST_OMP_TASK);
    *start = s > 0 ? s : 0;
    *end = e > -1 ? e : -1;
}

/* ------------------------------------------------------------------ */
/* generators                                                          */

long long orc_poisson7_nnz(int nx, int ny, int nz)
{
    long long n = (long long)nx * ny * nz;
    /* every cell has itself plus one entry per interior face */
    return n + 2LL * ((long long)(nx - 1) * ny * nz + (long long)nx * (ny - 1) * nz +
                      (long long)nx * ny * (nz - 1));
}

/* Shared stencil walk: column order z-, y-, x-, self, x+, y+, z+, which is
 * ascending for the row-major cell numbering. */
static void gen7(int nx, int ny, int nz, const double* kappa, int* xadj, int* cols,
                 double* vals)
{
    long long sxy = (long long)nx * ny;
    long long pos = 0;
    xadj[0] = 1;
    for (int z = 0; z < nz; ++z)
        for (int y = 0; y < ny; ++y)
            for (int x = 0; x < nx; ++x) {
                long long i = x + nx * (long long)y + sxy * z;
                int has[6] = {z > 0, y > 0, x > 0, x < nx - 1, y < ny - 1, z < nz - 1};
                long long nb[6] = {i - sxy, i - nx, i - 1, i + 1, i + nx, i + sxy};
                double kf[6];
                double ki = kappa ? kappa[i] : 1.0;
                for (int f = 0; f < 6; ++f) {
                    if (!kappa) {
                        kf[f] = 1.0;
                    } else if (has[f]) {
                        double kj = kappa[nb[f]];
                        kf[f] = (2.0 * ki * kj) / (ki + kj);
                    } else {
                        kf[f] = ki;
                    }
                }
                double diag = 0.0;
                for (int f = 0; f < 6; ++f) diag = diag + kf[f];
                if (!kappa) diag = 6.0;
                for (int f = 0; f < 3; ++f)
                    if (has[f]) { cols[pos] = (int)(nb[f] + 1); vals[pos] = -kf[f]; ++pos; }
                cols[pos] = (int)(i + 1); vals[pos] = diag; ++pos;
                for (int f = 3; f < 6; ++f)
                    if (has[f]) { cols[pos] = (int)(nb[f] + 1); vals[pos] = -kf[f]; ++pos; }
                xadj[i + 1] = (int)(pos + 1);
            }
}

void orc_gen_poisson7(int nx, int ny, int nz, int* xadj, int* cols, double* vals)
{
    gen7(nx, ny, nz, NULL, xadj, cols, vals);
}

void orc_gen_diffusion7(int nx, int ny, int nz, const double* kappa, int* xadj, int* cols,
                        double* vals)
{
    gen7(nx, ny, nz, kappa, xadj, cols, vals);
}

/* ------------------------------------------------------------------ */
/* formats                                                             */

void orc_coo_to_csr(int nrows, int nnz, const int* rows, int row_offset, int* xadj)
{
    int* cnt = (int*)calloc((size_t)nrows + 1, sizeof(int));
    for (int k = 0; k < nnz; ++k) cnt[rows[k] - row_offset - 1] += 1;
    xadj[0] = 1;
    for (int i = 1; i <= nrows; ++i) xadj[i] = xadj[i - 1] + cnt[i - 1];
    free(cnt);
}

void orc_stable_sort3(int* key, int* a, double* v, int n)
{
    if (n < 2) return;
    /* bottom-up merge sort of an index permutation; ties take the left run */
    int* idx = (int*)malloc((size_t)n * sizeof(int));
    int* tmp = (int*)malloc((size_t)n * sizeof(int));
    for (int i = 0; i < n; ++i) idx[i] = i;
    for (int w = 1; w < n; w *= 2) {
        for (int lo = 0; lo < n; lo += 2 * w) {
            int mid = lo + w < n ? lo + w : n;
            int hi = lo + 2 * w < n ? lo + 2 * w : n;
            int i = lo, j = mid, k = lo;
            while (i < mid && j < hi) tmp[k++] = key[idx[i]] <= key[idx[j]] ? idx[i++] : idx[j++];
            while (i < mid) tmp[k++] = idx[i++];
            while (j < hi) tmp[k++] = idx[j++];
        }
        int* s = idx; idx = tmp; tmp = s;
    }
    int* ki = (int*)malloc((size_t)n * sizeof(int));
    int* ai = (int*)malloc((size_t)n * sizeof(int));
    double* vi = (double*)malloc((size_t)n * sizeof(double));
    for (int i = 0; i < n; ++i) { ki[i] = key[idx[i]]; ai[i] = a[idx[i]]; vi[i] = v[idx[i]]; }
    memcpy(key, ki, (size_t)n * sizeof(int));
    memcpy(a, ai, (size_t)n * sizeof(int));
    memcpy(v, vi, (size_t)n * sizeof(double));
    free(ki); free(ai); free(vi); free(idx); free(tmp);
}

int orc_coo_sort_cel(int nrows, int nnz, int* rows, int* cols, double* vals)
{
    orc_stable_sort3(rows, cols, vals, nnz);                 /* :532-535 */
    /* the row-sorted state, restored when the loop runs off the end */
    size_t ni = (size_t)(nnz > 0 ? nnz : 1);
    int* r0 = (int*)malloc(ni * sizeof(int));
    int* c0 = (int*)malloc(ni * sizeof(int));
    double* v0 = (double*)malloc(ni * sizeof(double));
    memcpy(r0, rows, (size_t)nnz * sizeof(int));
    memcpy(c0, cols, (size_t)nnz * sizeof(int));
    memcpy(v0, vals, (size_t)nnz * sizeof(double));
    int status = 0;
    /* :544-575 with the reference's 1-based indices */
    int cur_index = 1, last_row = 1, row_counter = 0, start_index = 1, end_index = 1;
    while (row_counter < nrows - 1) {
        if (cur_index > nnz) { status = -1; break; }         /* rows(nnz+1) */
        int cur_row = rows[cur_index - 1];
        if (last_row != cur_row || cur_index == nnz) {
            if (last_row > cur_row) { status = -2; break; }  /* :555-558 */
            if (cur_index == nnz) end_index = cur_index;
            if (start_index < end_index)                     /* :562-567 */
                orc_stable_sort3(cols + start_index - 1, rows + start_index - 1,
                                 vals + start_index - 1, end_index - start_index + 1);
            start_index = end_index + 1;
            row_counter += 1;
            last_row = cur_row;
        }
        end_index = cur_index;
        cur_index += 1;
    }
    if (status != 0) {
        memcpy(rows, r0, (size_t)nnz * sizeof(int));
        memcpy(cols, c0, (size_t)nnz * sizeof(int));
        memcpy(vals, v0, (size_t)nnz * sizeof(double));
    }
    free(r0); free(c0); free(v0);
    return status;
}

void orc_coo_sort_rowcol(int nnz, int* rows, int* cols, double* vals)
{
    /* two stable passes: by column, then by row */
    orc_stable_sort3(cols, rows, vals, nnz);
    orc_stable_sort3(rows, cols, vals, nnz);
}

void orc_csr_diagonal(int n, const int* xadj, const int* cols, const double* vals,
                      int row_offset, double* diag)
{
    for (int i = 0; i < n; ++i) {
        int grow = i + 1 + row_offset;
        for (int j = xadj[i]; j < xadj[i + 1]; ++j)
            if (cols[j - 1] == grow) diag[i] = vals[j - 1];
    }
}

/* ------------------------------------------------------------------ */
/* SpMV                                                                */

static void csr_rows(int r0, int r1, const int* xadj, const int* cols, const double* vals,
                     int col_offset, const double* x, double* y)
{
    for (int i = r0; i <= r1; ++i) {
        double tmp = 0.0;
        for (int j = xadj[i - 1]; j < xadj[i]; ++j) {
            double prod = vals[j - 1] * x[cols[j - 1] - col_offset - 1];
            tmp = tmp + prod;
        }
        y[i - 1] = y[i - 1] + tmp;
    }
}

void orc_spmv_csr(int n, const int* xadj, const int* cols, const double* vals, int col_offset,
                  const double* x, double* y, int workers)
{
    if (workers <= 1) {
        csr_rows(1, n, xadj, cols, vals, col_offset, x, y);
        return;
    }
#pragma omp parallel for num_threads(workers) schedule(static, 1)
    for (int w = 1; w <= workers; ++w) {
        int s, e;
        orc_work_distr(n, workers, w, &s, &e);
        csr_rows(s, e, xadj, cols, vals, col_offset, x, y);
    }
}

void orc_spmv_csr_f32(int n, const int* xadj, const int* cols, const float* vals,
                      int col_offset, const float* x, float* y)
{
    for (int i = 0; i < n; ++i) {
        float tmp = 0.0f;
        for (int j = xadj[i]; j < xadj[i + 1]; ++j) {
            float prod = vals[j - 1] * x[cols[j - 1] - col_offset - 1];
            tmp = tmp + prod;
        }
        y[i] = y[i] + tmp;
    }
}

void orc_spmv_jad(int n, int mmax, const int* begin, const int* jcols, const double* jvals,
                  const int* perm, const double* x, double* y)
{
    double* tmp = (double*)calloc((size_t)(n > 2 * ORC_VRL ? n : 2 * ORC_VRL), sizeof(double));
    for (int d = 1; d <= mmax; ++d) {
        int k1 = begin[d - 1] - 1;
        int len = begin[d] - k1 - 1;
        for (int k = 1; k <= len; ++k) {
            double prod = jvals[k1 + k - 1] * x[jcols[k1 + k - 1] - 1];
            tmp[k - 1] = tmp[k - 1] + prod;
        }
    }
    for (int k = 1; k <= n; ++k) y[perm[k - 1] - 1] = tmp[k - 1];
    free(tmp);
}

void orc_spmv_coo(int nnz, const int* rows, const int* cols, const double* vals,
                  int row_offset, int col_offset, const double* x, double* y)
{
    for (int k = 0; k < nnz; ++k) {
        double prod = vals[k] * x[cols[k] - col_offset - 1];
        y[rows[k] - row_offset - 1] = y[rows[k] - row_offset - 1] + prod;
    }
}

/* ------------------------------------------------------------------ */
/* CSR -> JAD                                                          */

void orc_jad_dcsort(const int* ival, int nn, int* icnt, int* index, int ilo, int ihi)
{
    int nb = ihi - ilo + 1;
    for (int v = 0; v < nb; ++v) icnt[v] = 0;
    for (int i = 0; i < nn; ++i) icnt[ival[i] - ilo] += 1;
    for (int v = nb - 2; v >= 0; --v) icnt[v] += icnt[v + 1];
    for (int j = nn; j >= 1; --j) {
        int v = ival[j - 1] - ilo;
        index[icnt[v] - 1] = j;
        icnt[v] -= 1;
    }
}

int orc_jad_convert(int n, const int* xadj, const int* cols, const double* vals, int col_offset,
                    int mode, int* perm, int* begin, int* jcols, double* jvals, int* mmax,
                    int* ilo_out, long long* nnz_written)
{
    int* len = (int*)malloc(sizeof(int) * (size_t)(n > 0 ? n : 1));
    int ilo = n, ihi = 0;
    for (int i = 0; i < n; ++i) {
        len[i] = xadj[i + 1] - xadj[i];
        if (len[i] < ilo) ilo = len[i];
        if (len[i] > ihi) ihi = len[i];
    }
    *mmax = ihi;
    *ilo_out = ilo;
    /* the reference aliases dcsort's icnt(ilo:ihi) onto jad_begin(1:ihi-ilo+1) */
    int* icnt = (int*)malloc(sizeof(int) * (size_t)(ihi - ilo + 1 > 0 ? ihi - ilo + 1 : 1));
    orc_jad_dcsort(len, n, icnt, perm, ilo, ihi);

    int ndef = ihi;
    int* count = (int*)malloc(sizeof(int) * (size_t)(ihi + 1));
    if (mode == ORC_JAD_SPARSKIT || ilo == 0) {
        /* #{rows with len >= d}; with ilo == 0 the dcsort leftover
         * icnt[d-1] = #{len > d-1} is the same number */
        for (int d = 1; d <= ihi; ++d) count[d] = 0;
        for (int i = 0; i < n; ++i)
            for (int d = 1; d <= len[i]; ++d) count[d] += 1;
    } else {
        ndef = ihi - ilo + 1;
        if (ndef > ihi) ndef = ihi;
        for (int d = 1; d <= ndef; ++d) count[d] = icnt[d - 1];
    }

    long long k1 = 1, k0 = 1;
    for (int d = 1; d <= ndef; ++d) {
        for (int k = 1; k <= count[d]; ++k) {
            int src = xadj[perm[k - 1] - 1] + d - 1;
            jvals[k1 - 1] = vals[src - 1];
            jcols[k1 - 1] = cols[src - 1] - col_offset;
            ++k1;
        }
        begin[d - 1] = (int)k0;
        k0 = k1;
    }
    /* the start of diagonal ndef+1 is written before its (undefined) count
     * is read, so begin(ndef+1) is defined; later starts are not */
    begin[ndef] = (int)k1;
    for (int d = ndef + 2; d <= ihi + 1; ++d) begin[d - 1] = -1;
    *nnz_written = k1 - 1;
    free(count);
    free(icnt);
    free(len);
    return ndef;
}

/* ------------------------------------------------------------------ */
/* Jacobi vector and scaling                                           */

void orc_set_jacobi(int n, const double* diag, int scaled, double* jac)
{
    for (int i = 0; i < n; ++i) {
        if (scaled) {
            double sd = sqrt(fabs(diag[i]));
            jac[i] = sd / diag[i];
        } else {
            jac[i] = 1.0 / diag[i];
        }
    }
}

void orc_scale_csr(int n, const int* xadj, double* vals, const double* diag)
{
    for (int i = 0; i < n; ++i) {
        double s = sqrt(fabs(diag[i]));
        for (int j = xadj[i]; j < xadj[i + 1]; ++j) vals[j - 1] = vals[j - 1] / s;
    }
}

void orc_scale_coo(int nnz, const int* rows, int row_offset, double* vals, const double* diag)
{
    for (int k = 0; k < nnz; ++k) {
        double s = sqrt(fabs(diag[rows[k] - row_offset - 1]));
        vals[k] = vals[k] / s;
    }
}

void orc_scale_vec(int n, double* b, const double* diag)
{
    for (int i = 0; i < n; ++i) b[i] = b[i] / sqrt(fabs(diag[i]));
}

/* ------------------------------------------------------------------ */
/* Jacobi-PCG, reference op order (cel_cgalg_module.f90:206-304)         */

/* Worker chunk sum, combined by the master in worker order after a barrier
 * (cel_fop_vec_distr_module.f90:460-472, 724-775). */
#define DOT_PARTIAL(expr)                                   \
    do {                                                    \
        if (w > 0) {                                        \
            double acc = 0.0;                               \
            for (int i = s - 1; i < e; ++i) {               \
                double prod = (expr);                       \
                acc = acc + prod;                           \
            }                                               \
            partial[w] = acc;                               \
        }                                                   \
    } while (0)

static double combine(const double* partial, int workers)
{
    double sum = 0.0;
    for (int w = 1; w <= workers; ++w) sum = sum + partial[w];
    return sum;
}

int orc_pcg_jacobi(int n, const int* xadj, const int* cols, const double* vals, int col_offset,
                   const double* jac, const double* b, double* x, int x0_mode, double eps,
                   int iter_max, int workers, double* hist, double* alpha_hist,
                   double* sigma_hist, double* final_norm)
{
    if (workers < 1) workers = 1;
    double* r = (double*)malloc(sizeof(double) * (size_t)n);
    double* z = (double*)malloc(sizeof(double) * (size_t)n);
    double* p = (double*)malloc(sizeof(double) * (size_t)n);
    double* v = (double*)malloc(sizeof(double) * (size_t)n);
    double* partial = (double*)calloc((size_t)workers + 1, sizeof(double));
    double alpha = 0.0, sigma = 0.0, norm2 = 0.0;
    int iters = 0;
    const double small = orc_real_small();

#pragma omp parallel num_threads(workers + 1)
    {
#ifdef _OPENMP
        int w = omp_get_thread_num(); /* 0 = master (computes nothing), 1..W workers */
        if (omp_get_num_threads() != workers + 1) w = -1;
#else
        int w = 1;
#endif
        int s = 1, e = 0;
        if (w > 0) orc_work_distr(n, workers, w, &s, &e);
#ifndef _OPENMP
        if (workers > 1) { s = 1; e = n; workers = 1; }
#endif
        /* r = 0 ; x = 1.0 */
        if (w > 0)
            for (int i = s - 1; i < e; ++i) {
                r[i] = 0.0;
                if (x0_mode == ORC_X0_ONE) x[i] = 1.0;
                else if (x0_mode == ORC_X0_ZERO) x[i] = 0.0;
            }
#pragma omp barrier
        /* r = r + A x */
        if (w > 0) csr_rows(s, e, xadj, cols, vals, col_offset, x, r);
#pragma omp barrier
        /* r = b - r */
        if (w > 0)
            for (int i = s - 1; i < e; ++i) r[i] = b[i] - r[i];
#pragma omp barrier
        /* p = jacobi * r */
        if (w > 0)
            for (int i = s - 1; i < e; ++i) p[i] = r[i] * jac[i];
#pragma omp barrier
        DOT_PARTIAL(r[i] * p[i]);
#pragma omp barrier
#pragma omp master
        alpha = combine(partial, workers);
#pragma omp barrier
        DOT_PARTIAL(r[i] * r[i]);
#pragma omp barrier
#pragma omp master
        {
            norm2 = sqrt(combine(partial, workers));
            if (hist) hist[0] = norm2;
            if (alpha_hist) alpha_hist[0] = alpha;
            if (sigma_hist) sigma_hist[0] = 0.0;
        }
#pragma omp barrier
        int ii = 0;
        while (ii < iter_max && fabs(alpha) > small && norm2 > eps) {
            /* v = 0 ; v = v + A p */
            if (w > 0)
                for (int i = s - 1; i < e; ++i) v[i] = 0.0;
#pragma omp barrier
            if (w > 0) csr_rows(s, e, xadj, cols, vals, col_offset, p, v);
#pragma omp barrier
            ii = ii + 1;
            /* sigma = alpha / (v . p) */
            DOT_PARTIAL(v[i] * p[i]);
#pragma omp barrier
#pragma omp master
            sigma = alpha / combine(partial, workers);
#pragma omp barrier
            /* x = x + sigma p ; r = r + (-sigma) v */
            if (w > 0) {
                double ms = -sigma;
                for (int i = s - 1; i < e; ++i) {
                    double t = sigma * p[i];
                    x[i] = x[i] + t;
                }
                for (int i = s - 1; i < e; ++i) {
                    double t = ms * v[i];
                    r[i] = r[i] + t;
                }
                for (int i = s - 1; i < e; ++i) z[i] = r[i] * jac[i];
            }
#pragma omp barrier
            /* beta = (r . z) / alpha ; alpha = r . z */
            DOT_PARTIAL(r[i] * z[i]);
#pragma omp barrier
#pragma omp master
            {
                double rz = combine(partial, workers);
                sigma = rz / alpha;
                alpha = rz;
            }
#pragma omp barrier
            /* p = beta p + z */
            if (w > 0)
                for (int i = s - 1; i < e; ++i) {
                    double t = sigma * p[i];
                    p[i] = t + z[i];
                }
#pragma omp barrier
            DOT_PARTIAL(r[i] * r[i]);
#pragma omp barrier
#pragma omp master
            {
                norm2 = sqrt(combine(partial, workers));
                if (hist) hist[ii] = norm2;
                if (alpha_hist) alpha_hist[ii] = alpha;
                if (sigma_hist) sigma_hist[ii] = sigma;
                iters = ii;
            }
#pragma omp barrier
        }
    }
    /* sigma_hist[k] holds beta_k for k >= 1 (the reference reuses sigma_new) */
    *final_norm = norm2;
    free(r); free(z); free(p); free(v); free(partial);
    return iters;
}

/* ------------------------------------------------------------------ */
/* Jacobi-GMRES, reference op order (cel_gmresalg_module.f90:570-851)   */

/* Dot product as the reference's workers form it (:724-775 of
 * cel_fop_vec_distr_module.f90): each worker sums its chunk left to right
 * from 0.0, the partials are added in worker order from 0.0. */
static double dot_chunks(int n, int workers, const double* a, const double* b)
{
    double sum = 0.0;
    for (int w = 1; w <= workers; ++w) {
        int s, e;
        orc_work_distr(n, workers, w, &s, &e);
        double acc = 0.0;
        for (int i = s - 1; i < e; ++i) {
            double prod = a[i] * b[i];
            acc = acc + prod;
        }
        sum = sum + acc;
    }
    return sum;
}

/* r = 0 ; r = r + A x ; r = b - r ; returns sqrt(r . r) */
static double gm_residual(int n, const int* xadj, const int* cols, const double* vals,
                          int col_offset, const double* b, const double* x, double* r, int workers)
{
    for (int i = 0; i < n; ++i) r[i] = 0.0;
    csr_rows(1, n, xadj, cols, vals, col_offset, x, r);
    for (int i = 0; i < n; ++i) r[i] = b[i] - r[i];
    return sqrt(dot_chunks(n, workers, r, r));
}

int orc_gmres_jacobi(int n, const int* xadj, const int* cols, const double* vals, int col_offset,
                     const double* jac, const double* b, double* x, int x0_mode, double eps,
                     int iter_max, int m, int workers, double* final_norm, int* reason,
                     int* cycles_out, double* hist, int hist_cap)
{
    if (workers < 1) workers = 1;
    const size_t nn = (size_t)(n > 0 ? n : 1);
    double* r = (double*)malloc(sizeof(double) * nn);
    double* V = (double*)malloc(sizeof(double) * nn * (size_t)(m + 1));  /* v_1 .. v_{m+1} */
    double* Z = (double*)malloc(sizeof(double) * nn * (size_t)m);        /* z_1 .. z_m */
    double* W = (double*)malloc(sizeof(double) * nn * (size_t)m);        /* w_1 .. w_m */
    /* 1-based small arrays as the reference's: hh(m+1, m+1), cc, ss, gamma, alpha(m+1) */
    const int M1 = m + 2;
    double* hh = (double*)calloc((size_t)M1 * M1, sizeof(double));
    double* cc = (double*)calloc((size_t)M1, sizeof(double));
    double* ss = (double*)calloc((size_t)M1, sizeof(double));
    double* gam = (double*)calloc((size_t)M1, sizeof(double));
    double* alp = (double*)calloc((size_t)M1, sizeof(double));
#define H(i, j) hh[(size_t)(i) * M1 + (j)]
#define VV(k) (V + (size_t)((k) - 1) * nn)
#define ZZ(k) (Z + (size_t)((k) - 1) * nn)
#define WW(k) (W + (size_t)((k) - 1) * nn)
    const double small = orc_real_small();
    int conv = 0, conv_field = 0, iter = 0, cycles = 0;
    double residuum_norm = 0.0;
    /* :637 x = 1.0 */
    for (int i = 0; i < n; ++i) {
        if (x0_mode == ORC_X0_ONE) x[i] = 1.0;
        else if (x0_mode == ORC_X0_ZERO) x[i] = 0.0;
    }
    while (iter <= iter_max && conv == 0) {                     /* :640 */
        iter = iter + 1;
        residuum_norm = gm_residual(n, xadj, cols, vals, col_offset, b, x, r, workers);
        gam[1] = residuum_norm;                                 /* :656-660 */
        conv_field = fabs(residuum_norm) <= small ? 1 : 0;
        double local_residuum = residuum_norm;
        int restart_count = 0;
        while (restart_count <= m && conv == 0) {               /* :665 */
            double tmp = 1.0 / local_residuum;                  /* scale_inv */
            for (int i = 0; i < n; ++i) VV(1)[i] = r[i] * tmp;
            gam[1] = residuum_norm;
            for (int jj = 1; jj <= m; ++jj) {
                double* w = WW(jj);
                double* z = ZZ(jj);
                for (int i = 0; i < n; ++i) w[i] = 0.0;
                for (int i = 0; i < n; ++i) z[i] = VV(jj)[i] * jac[i];
                csr_rows(1, n, xadj, cols, vals, col_offset, z, w);
                for (int ii = 1; ii <= jj; ++ii) {              /* :697-711 */
                    H(ii, jj) = dot_chunks(n, workers, VV(ii), w);
                    const double h = H(ii, jj);
                    for (int i = 0; i < n; ++i) {
                        double t = h * VV(ii)[i];
                        w[i] = w[i] - t;
                    }
                }
                H(jj + 1, jj) = sqrt(dot_chunks(n, workers, w, w));   /* :713-717 */
                for (int ii = 1; ii <= jj - 1; ++ii) {          /* :722-727, as written */
                    double a = H(ii, jj) * cc[ii];
                    double c = H(ii + 1, jj) * ss[ii];
                    H(ii, jj) = a + c;
                    a = -H(ii, jj) * ss[ii];
                    c = H(ii + 1, jj) * cc[ii];
                    H(ii, jj) = a + c;
                }
                {
                    double a = H(jj, jj) * H(jj, jj);           /* :728 */
                    double c = H(jj + 1, jj) * H(jj + 1, jj);
                    double betta = sqrt(a + c);
                    ss[jj] = H(jj + 1, jj) / betta;
                    cc[jj] = H(jj, jj) / betta;
                    H(jj, jj) = betta;
                    gam[jj + 1] = -ss[jj] * gam[jj];
                    gam[jj] = cc[jj] * gam[jj];
                }
                if (fabs(gam[jj + 1]) <= eps) {                 /* :737-766 */
                    for (int ii = jj; ii >= 1; --ii) {
                        alp[ii] = 0.0;
                        for (int kk = ii + 1; kk <= jj; ++kk) {
                            double t = H(ii, kk) * alp[kk];
                            alp[ii] = alp[ii] + t;
                        }
                        double inv = 1.0 / H(ii, ii);
                        alp[ii] = inv * (gam[ii] - alp[ii]);
                    }
                    for (int ii = 1; ii <= jj; ++ii) {
                        const double a = alp[ii];
                        for (int i = 0; i < n; ++i) {
                            double t = a * ZZ(ii)[i];
                            x[i] = x[i] + t;
                        }
                    }
                    conv_field = 2;
                    conv = 2;
                    break;
                }
                const double hd = H(jj + 1, jj);                /* :767-778 */
                for (int i = 0; i < n; ++i) VV(jj + 1)[i] = w[i] / hd;
            }
            if (conv == 0) {                                    /* :780-800 */
                for (int ii = m; ii >= 1; --ii) {
                    alp[ii] = 0.0;
                    for (int kk = ii + 1; kk <= m; ++kk) {
                        double t = H(ii, kk) * alp[kk];
                        alp[ii] = alp[ii] + t;
                    }
                    double inv = 1.0 / H(ii, ii);
                    alp[ii] = inv * (gam[ii] - alp[ii]);
                }
                for (int ii = 1; ii <= m; ++ii) {
                    const double a = alp[ii];
                    for (int i = 0; i < n; ++i) {
                        double t = a * ZZ(ii)[i];
                        x[i] = x[i] + t;
                    }
                }
            }
            residuum_norm = gm_residual(n, xadj, cols, vals, col_offset, b, x, r, workers);
            local_residuum = residuum_norm;                     /* :802-815 */
            if (hist && cycles < hist_cap) hist[cycles] = residuum_norm;
            cycles += 1;
            restart_count += 1;
        }
    }
#undef H
#undef VV
#undef ZZ
#undef WW
    *final_norm = residuum_norm;
    if (reason) *reason = conv_field;
    if (cycles_out) *cycles_out = cycles;
    free(r); free(V); free(Z); free(W); free(hh); free(cc); free(ss); free(gam); free(alp);
    return iter;
}

/* ------------------------------------------------------------------ */
/* distributed block split                                             */

int orc_owner_of(int nranks, const int* vtxdist, int col)
{
    for (int p = 0; p < nranks; ++p)
        if (vtxdist[p] <= col && col < vtxdist[p + 1]) return p;
    return -1;
}

int orc_distr_split(int nranks, const int* vtxdist, int my_rank, int nnz, const int* rows,
                    const int* cols, const double* vals, int* procs, int* block_nnz,
                    int* out_rows, int* out_cols, double* out_vals, int* row_offset,
                    int* col_offset, int* xadj, int* comm_len, int* comm_perm, double* diagonal)
{
    int n = vtxdist[my_rank + 1] - vtxdist[my_rank];
    int my_off = vtxdist[my_rank] - 1;
    int* per_rank = (int*)calloc((size_t)nranks, sizeof(int));
    int* owner = (int*)malloc(sizeof(int) * (size_t)(nnz > 0 ? nnz : 1));
    for (int k = 0; k < nnz; ++k) {
        owner[k] = orc_owner_of(nranks, vtxdist, cols[k]);
        if (owner[k] < 0) { free(per_rank); free(owner); return -1; }
        per_rank[owner[k]] += 1;
    }
    /* ascending ranks with entries, then my rank moved to the front */
    int nb = 0;
    for (int p = 0; p < nranks; ++p)
        if (per_rank[p] > 0) procs[nb++] = p;
    if (nb > 1 && procs[0] != my_rank) {
        int idx = -1;
        for (int b = 0; b < nb; ++b)
            if (procs[b] == my_rank) idx = b;
        if (idx < 0) { free(per_rank); free(owner); return -1; }
        for (int b = idx; b >= 1; --b) procs[b] = procs[b - 1];
        procs[0] = my_rank;
    }
    int* block_of = (int*)malloc(sizeof(int) * (size_t)nranks);
    for (int p = 0; p < nranks; ++p) block_of[p] = -1;
    for (int b = 0; b < nb; ++b) block_of[procs[b]] = b;
    long long* start = (long long*)malloc(sizeof(long long) * (size_t)(nb + 1));
    start[0] = 0;
    for (int b = 0; b < nb; ++b) {
        block_nnz[b] = per_rank[procs[b]];
        start[b + 1] = start[b] + block_nnz[b];
    }
    long long* fill = (long long*)malloc(sizeof(long long) * (size_t)(nb > 0 ? nb : 1));
    for (int b = 0; b < nb; ++b) fill[b] = start[b];
    for (int i = 0; i < n; ++i) diagonal[i] = 0.0;
    for (int k = 0; k < nnz; ++k) {
        int b = block_of[owner[k]];
        long long q = fill[b]++;
        out_rows[q] = rows[k];
        out_cols[q] = cols[k];
        out_vals[q] = vals[k];
        if (cols[k] == rows[k]) diagonal[rows[k] - my_off - 1] = vals[k];
    }
    *row_offset = my_off;
    long long perm_pos = 0;
    char* mark = NULL;
    for (int b = 0; b < nb; ++b) {
        int p = procs[b];
        int plen = vtxdist[p + 1] - vtxdist[p];
        col_offset[b] = vtxdist[p] - 1;
        orc_coo_to_csr(n, block_nnz[b], out_rows + start[b], my_off, xadj + (size_t)b * (n + 1));
        mark = (char*)realloc(mark, (size_t)(plen > 0 ? plen : 1));
        memset(mark, 0, (size_t)(plen > 0 ? plen : 1));
        int cnt = 0;
        for (long long q = start[b]; q < start[b + 1]; ++q) {
            int lc = out_cols[q] - col_offset[b];
            if (!mark[lc - 1]) { mark[lc - 1] = 1; ++cnt; }
        }
        comm_len[b] = cnt;
        if (b >= 1)
            for (int j = 1; j <= plen; ++j)
                if (mark[j - 1]) comm_perm[perm_pos++] = j;
    }
    free(mark); free(fill); free(start); free(block_of); free(owner); free(per_rank);
    return nb;
}

void orc_face_pack(int len, const int* perm, const double* x, double* buf)
{
    for (int j = 0; j < len; ++j) buf[j] = x[perm[j] - 1];
}

/* ------------------------------------------------------------------ */
/* build-only extensions                                               */

void orc_sddmm_csr(int n, const int* xadj, const int* cols, int col_offset,
                   const double* lambda, const double* x, double* g)
{
    for (int i = 0; i < n; ++i)
        for (int j = xadj[i]; j < xadj[i + 1]; ++j) {
            double t = lambda[i] * x[cols[j - 1] - col_offset - 1];
            g[j - 1] = -t;
        }
}

void orc_diffusion7_kappa_grad(int nx, int ny, int nz, const double* kappa, const double* x,
                               const double* lambda, double* grad)
{
    long long sxy = (long long)nx * ny;
    for (int z = 0; z < nz; ++z)
        for (int y = 0; y < ny; ++y)
            for (int xx = 0; xx < nx; ++xx) {
                long long i = xx + nx * (long long)y + sxy * z;
                int has[6] = {z > 0, y > 0, xx > 0, xx < nx - 1, y < ny - 1, z < nz - 1};
                long long nb[6] = {i - sxy, i - nx, i - 1, i + 1, i + nx, i + sxy};
                double ki = kappa[i];
                double g = 0.0;
                for (int f = 0; f < 6; ++f) {
                    double term;
                    if (has[f]) {
                        long long j = nb[f];
                        double kj = kappa[j];
                        double dLdk = -((lambda[i] - lambda[j]) * (x[i] - x[j]));
                        double s = ki + kj;
                        double dk = (2.0 * kj * kj) / (s * s);
                        term = dLdk * dk;
                    } else {
                        term = -(lambda[i] * x[i]);
                    }
                    g = g + term;
                }
                grad[i] = g;
            }
}
