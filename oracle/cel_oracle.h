/*
 * cel_oracle.h -- CPU restatement of the CEL Jacobi-PCG hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is what the GPU library's results
 * are compared against, and the CPU baseline timed beside it.  Only
 * tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load
 * it.  The product (cel_v3_amd/libcel_gpu.so) never links or calls it.
 *
 * Conventions follow the reference Fortran exactly:
 *   - every index array is 1-based (xadj(1) = 1, cols in 1..ncols),
 *   - INDEX = int32, REAL = double (src/include/base/cel_types.h:12-13,
 *     src/fbase/cel_types_module.f90:13-14),
 *   - element-wise arithmetic is compiled with -ffp-contract=off so every
 *     product is rounded before it is added (no fused multiply-add).
 *
 * Each function cites the reference file:line it restates.  How each one is
 * pinned (golden vectors, closed forms, brute force, the compiled reference
 * generator) is stated in its comment and tested under tests/.
 */
#ifndef CEL_ORACLE_H
#define CEL_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* ---- constants (src/fbase/cel_types_module.f90:33,57; cel_omp_shared_work_module.f90) */
#define ORC_AW_SMALLEST_OMP_TASK 4
#define ORC_VRL 64

/* cel_real_small = 1.1 * EPSILON(1.0_rk)  (cel_types_module.f90:33) */
double orc_real_small(void);

/* Work partition of the reference OpenMP master/worker model
 * (cel_omp_shared_work_module.f90:183-230, cel_omp_module.f90:71-83).
 * worker in 1..num_workers.  Writes the 1-based inclusive [start, end]. */
void orc_work_distr(int length, int num_workers, int worker, int* start, int* end);

/* ---- generators (closed forms; no reference counterpart for the diag-6 matrix) */

/* nnz of the 7-point Dirichlet-outside matrix on an nx*ny*nz grid. */
long long orc_poisson7_nnz(int nx, int ny, int nz);

/* 7-point Poisson, diag 6, off-diagonal -1, Dirichlet outside.
 * Cell (x,y,z) -> row x + nx*(y + ny*z) + 1.  Columns ascending.  1-based CSR. */
void orc_gen_poisson7(int nx, int ny, int nz, int* xadj, int* cols, double* vals);

/* Variable-coefficient 7-point diffusion.  Face coefficient between cells
 * i and j is the harmonic mean k = (2*ki*kj)/(ki+kj); a face on the domain
 * boundary uses k = ki.  a_ij = -k, a_ii = sum of the six face k in the
 * order z-, y-, x-, x+, y+, z+.  With kappa == 1 this is orc_gen_poisson7. */
void orc_gen_diffusion7(int nx, int ny, int nz, const double* kappa,
                        int* xadj, int* cols, double* vals);

/* ---- sparse formats */

/* cel_sp_mat_to_csr (src/fspmt/cel_sp_mat_module.f90:430-456): xadj from row
 * counts of a row-sorted COO; entries are not moved; xadj(1) = 1. */
void orc_coo_to_csr(int nrows, int nnz, const int* rows, int row_offset, int* xadj);

/* Stable ascending sort of n entries by key[], carrying a[] and v[] (ties
 * keep their input order).  The output of a stable sort is fixed by the
 * keys and the input order alone, so this gives the result of the
 * reference's cel_sp_mat_merge_sort (src/fspmt/cel_sp_mat_module.f90:620-701:
 * merge on '<=' takes the left run, :706-749 (:727); bubble-sort leaves swap
 * on '>' only, :588-618), whatever its recursion and leaf sizes. */
void orc_stable_sort3(int* key, int* a, double* v, int n);

/* cel_sp_mat_format_sparse_31_sort (src/fspmt/cel_sp_mat_module.f90:518-585),
 * 1-based COO in place: a stable sort by row (:532-535), then the column
 * loop (:544-575) restated statement by statement.  Its quirks follow from
 * the loop: the segment of the last transition it handles is the last one
 * column-sorted, so the last row keeps its row-sorted (file) order; when
 * rows(1) != 1 the first entry is excluded from its row's column sort; a
 * last row of one entry is sorted together with the row before it; with
 * nrows <= 1 nothing is column-sorted.
 * Returns 0, or -1 where the loop would read rows(nnz+1) (fewer row
 * transitions than nrows-1: empty rows, nnz == 0): the reference reads past
 * its arrays there and its result is undefined; the arrays are then left
 * in the row-sorted state (column sorts already done by the loop are
 * undone).  -2 where the reference stops with "rows not sorted" (a row
 * index below 1), arrays likewise row-sorted. */
int orc_coo_sort_cel(int nrows, int nnz, int* rows, int* cols, double* vals);

/* Stable sort by (row, col): what the column loop above intends. */
void orc_coo_sort_rowcol(int nnz, int* rows, int* cols, double* vals);

/* Extract the diagonal of a CSR block whose global row of local row i is
 * i + row_offset and whose column indices are global (the "diagonal" array
 * captured in cel_sp_mat_distr_module.f90:197-207; last write wins). */
void orc_csr_diagonal(int n, const int* xadj, const int* cols, const double* vals,
                      int row_offset, double* diag);

/* ---- SpMV (src/fop/cel_fop_mv_module.f90) */

/* CSR, :215-229.  y(i) += sum_j a_ij * x(cols(j) - col_offset), each row
 * summed left to right from 0.0, then added to y(i).  workers >= 1 splits
 * rows by orc_work_distr (values are independent of the split). */
void orc_spmv_csr(int n, const int* xadj, const int* cols, const double* vals,
                  int col_offset, const double* x, double* y, int workers);

/* Same in FP32: FP32 values, FP32 x, FP32 row accumulation, y(i) += tmp. */
void orc_spmv_csr_f32(int n, const int* xadj, const int* cols, const float* vals,
                      int col_offset, const float* x, float* y);

/* JAD, :250-352.  Race-free restatement of the three phases:
 * tmp = 0; tmp(k) += jad_values * x(jad_cols) per diagonal in order;
 * y(perm(k)) = tmp(k)  (overwrite, as the reference does). */
void orc_spmv_jad(int n, int mmax, const int* begin, const int* jcols,
                  const double* jvals, const int* perm, const double* x, double* y);

/* COO, :357-403.  y(row - row_offset) += a * x(col - col_offset), entries in
 * storage order (the reference uses omp atomic; sequential order here). */
void orc_spmv_coo(int nnz, const int* rows, const int* cols, const double* vals,
                  int row_offset, int col_offset, const double* x, double* y);

/* ---- CSR -> JAD (src/fspmtconverter/cel_sp_mat_jad_converter_module.f90) */

#define ORC_JAD_SPARSKIT 0   /* correct diagonal lengths #{len >= d} */
#define ORC_JAD_CEL_COMPAT 1 /* the reference's lengths from dcsort leftovers */

/* dcsort, :57-88.  Stable descending counting sort of rows by length.
 * icnt has ihi-ilo+1 entries; on return icnt[v-ilo] = #{len > v}. */
void orc_jad_dcsort(const int* ival, int nn, int* icnt, int* index, int ilo, int ihi);

/* csrjad, :93-175.  Outputs are 1-based.  begin needs mmax+1 entries,
 * jcols/jvals need nnz entries.  Returns the number of diagonals whose
 * begin/cols/values are defined: mmax for SPARSKIT and for CEL_COMPAT with
 * ilo == 0; ihi-ilo+1 for CEL_COMPAT with ilo >= 1, where the reference
 * reads uninitialised counts for the remaining diagonals.  begin(1..ndef+1)
 * is defined (the start of diagonal ndef+1 is written before its count is
 * read); later entries are set to -1 and nothing is fabricated. */
int orc_jad_convert(int n, const int* xadj, const int* cols, const double* vals,
                    int col_offset, int mode, int* perm, int* begin, int* jcols,
                    double* jvals, int* mmax, int* ilo_out, long long* nnz_written);

/* ---- Jacobi vector and row scaling (src/fspmtdistr/cel_sp_mat_distr_format_module.f90) */

/* :149-203.  jac = 1/d, or sqrt|d|/d when scaled != 0. */
void orc_set_jacobi(int n, const double* diag, int scaled, double* jac);
/* :276-315.  a_ij = a_ij / sqrt|d_i| for every entry of row i (CSR). */
void orc_scale_csr(int n, const int* xadj, double* vals, const double* diag);
/* :318-356.  COO variant, row index via row_offset. */
void orc_scale_coo(int nnz, const int* rows, int row_offset, double* vals, const double* diag);
/* :207-239.  b_i = b_i / sqrt|d_i|. */
void orc_scale_vec(int n, double* b, const double* diag);

/* ---- Jacobi-PCG in the reference op order (src/fcgalg/cel_cgalg_module.f90:178-305) */

#define ORC_X0_ONE 0   /* the reference: x = 1.0 (:209) */
#define ORC_X0_ZERO 1
#define ORC_X0_USER 2

/* Runs the loop exactly as the reference with one rank.  The matrix is the
 * local block only.  Dot products follow the reference grouping: each
 * worker sums its orc_work_distr chunk left to right from 0.0; the partials
 * are combined in worker order 1..workers (the reference combines them with
 * omp atomic in arrival order).  threads = workers + 1, thread 0 computes
 * nothing, as in the reference.
 *
 * hist (may be NULL) receives norm2 after setup (index 0) and after every
 * iteration (index k); alpha_hist/sigma_hist likewise (sigma_hist[0] = 0).
 * Returns the number of iterations done; *final_norm receives norm2. */
int orc_pcg_jacobi(int n, const int* xadj, const int* cols, const double* vals,
                   int col_offset, const double* jac, const double* b, double* x,
                   int x0_mode, double eps, int iter_max, int workers,
                   double* hist, double* alpha_hist, double* sigma_hist,
                   double* final_norm);

/* ---- Jacobi-GMRES in the reference op order (src/fcgalg/cel_gmresalg_module.f90:570-851) */

/* cel_gmresalg_2jacobi with one rank: x = 1 (or per x0_mode); outer
 * iterations (at most iter_max+1) each recompute r = b - A x and run up to
 * m+1 restart cycles of m Arnoldi steps with right Jacobi preconditioning
 * (z_j = jac * v_j, w_j = A z_j, modified Gram-Schmidt), stopping when the
 * Givens estimate |gamma(j+1)| <= eps (absolute).  The rotation of column
 * j by the earlier rotations is restated as the reference writes it
 * (:722-727): both statements assign hh(i,j), so hh(i+1,j) is never
 * rotated.  Dot products as dot_chunks (worker chunks, in worker order).
 * Returns the outer iteration count (the reference's iter_done);
 * *final_norm = the last true residual norm (residuum_norm), *reason = 2
 * when the estimate met eps, 1 when the first residual was below
 * cel_real_small, else 0; hist[k] = true residual norm after cycle k. */
int orc_gmres_jacobi(int n, const int* xadj, const int* cols, const double* vals, int col_offset,
                     const double* jac, const double* b, double* x, int x0_mode, double eps,
                     int iter_max, int m, int workers, double* final_norm, int* reason,
                     int* cycles_out, double* hist, int hist_cap);

/* ---- distributed block split (src/fspmtdistr/cel_sp_mat_distr_module.f90:34-301) */

/* Owning rank of global column col (1-based) for vtxdist (1-based starts,
 * nranks+1 entries).  Returns -1 when col is out of range. */
int orc_owner_of(int nranks, const int* vtxdist, int col);

/* Splits this rank's row-sorted COO into block 1 (columns owned by
 * my_rank) followed by one block per other owning rank in ascending rank
 * order, entries kept in input order.  Outputs (caller sizes them for the
 * worst case: nranks blocks, nnz entries):
 *   procs[b]          owning rank of block b
 *   block_nnz[b]      entries in block b; out_rows/cols/vals concatenated
 *   row_offset        vtxdist(my_rank+1)-1; col_offset[b] = vtxdist(procs[b]+1)-1
 *   xadj              (n+1) per block, concatenated (cel_sp_mat_to_csr)
 *   comm_len[b]       number of distinct columns referenced in block b
 *   comm_perm         ascending list of referenced local column indices of
 *                     every block b >= 2, concatenated (com_buffer_perm)
 *   diagonal          local diagonal (n)
 * Returns the number of blocks, or -1 on a column out of range. */
int orc_distr_split(int nranks, const int* vtxdist, int my_rank, int nnz,
                    const int* rows, const int* cols, const double* vals,
                    int* procs, int* block_nnz, int* out_rows, int* out_cols,
                    double* out_vals, int* row_offset, int* col_offset, int* xadj,
                    int* comm_len, int* comm_perm, double* diagonal);

/* Face packing, cel_sp_mat_distr_vec_module.f90:482-486:
 * buf(j) = x(perm(j)), j = 1..len. */
void orc_face_pack(int len, const int* perm, const double* x, double* buf);

/* ---- build-only extensions (no reference counterpart) */

/* FP64 reference gradient for L = 0.5*||x - x_obs||^2 with A(kappa) x = b
 * (orc_gen_diffusion7 pattern).  Given x and the adjoint lambda
 * (A^T lambda = x - x_obs), writes dL/dkappa per cell:
 *   interior face f=(i,j): dL/dk_f = -(l_i - l_j)(x_i - x_j),
 *                          dk_f/dki = 2 kj^2 / (ki + kj)^2
 *   boundary face of i:    dL/dk_f = -l_i x_i,  dk_f/dki = 1. */
void orc_diffusion7_kappa_grad(int nx, int ny, int nz, const double* kappa,
                               const double* x, const double* lambda, double* grad);

/* Sampled dense product on the CSR pattern: g(j) = -lambda(i) * x(cols(j)). */
void orc_sddmm_csr(int n, const int* xadj, const int* cols, int col_offset,
                   const double* lambda, const double* x, double* g);

#ifdef __cplusplus
}
#endif
#endif
