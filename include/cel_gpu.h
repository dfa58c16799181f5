/*
 * cel_gpu.h -- C ABI of the MI355X (gfx950) Jacobi-PCG library.
 *
 * Drop-in boundary for the hot path of the CEL linear solver.  The
 * reference exposes this path only as Fortran module procedures; every
 * entry point below names the reference interface it replaces
 * (paths relative to the reference tree).  Fortran reaches these
 * functions through ISO_C_BINDING (cel_v3_amd/fortran/cel_gpu_mod.f90).
 *
 * Conventions
 *   - INDEX = int (32-bit), REAL = double, as src/include/base/cel_types.h:12-13.
 *   - Index arrays take an explicit index_base (0 or 1).  Arrays handed back
 *     for parity checks are 1-based, exactly as the reference stores them.
 *   - memspace says whether a vector pointer is host (CEL_MEM_HOST) or
 *     device memory (CEL_MEM_DEVICE, from cel_gpu_malloc).
 *   - Every function returns a status code; nothing aborts the process
 *     (the reference STOPs in cel_error, src/fbase/cel_error_module.f90:17-32).
 *     cel_gpu_last_error() returns the message of the last failure.
 *   - One host thread per matrix handle.  All work is ordered on the
 *     library's stream; calls return after their results are complete.
 */
#ifndef CEL_GPU_H
#define CEL_GPU_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ------------------------------------------------------ */
#define CEL_OK 0
#define CEL_NOT_CONVERGED 1     /* iter_max reached; results and stats valid */
#define CEL_ERR_ARG (-1)
#define CEL_ERR_HIP (-2)
#define CEL_ERR_NO_DEVICE (-3)
#define CEL_ERR_BAD_MATRIX (-4) /* index out of range, unsorted rows, bad xadj */
#define CEL_ERR_BAD_DIAG (-5)   /* zero or missing diagonal entry */
#define CEL_ERR_BREAKDOWN (-6)  /* p.Ap <= 0 or not finite */
#define CEL_ERR_NONFINITE (-7)  /* NaN/Inf in the right-hand side */
#define CEL_ERR_OOM (-8)
#define CEL_ERR_REF_UNDEFINED (-9) /* the reference's result is undefined here */
#define CEL_ERR_STATE (-10)     /* call order (e.g. JAD requested before conversion) */
#define CEL_ERR_FILE (-11)      /* file not readable, or not a Matrix Market file the
                                   reference's reader accepts */

/* ---- enumerations ------------------------------------------------------ */
/* storage formats; values equal MV_CSR/MV_JAD/MV_COO of
 * src/fbase/cel_types_module.f90:57-59 */
#define CEL_FMT_CSR 1
#define CEL_FMT_JAD 2
#define CEL_FMT_COO 3
#define CEL_FMT_DIA 4  /* diagonal storage of a CSR block whose entries lie on at most
                          32 diagonals (no reference counterpart; see cel_gpu_mat_to_dia) */

#define CEL_MEM_HOST 0
#define CEL_MEM_DEVICE 1

/* initial guess: the reference hard-codes x0 = 1 (src/fcgalg/cel_cgalg_module.f90:209) */
#define CEL_X0_ONE 0
#define CEL_X0_ZERO 1
#define CEL_X0_USER 2

/* stopping quantity: the reference stops on the absolute 2-norm,
 * norm2 > eps (cel_cgalg_module.f90:247-249) */
#define CEL_TOL_ABS 0
#define CEL_TOL_REL_B 1   /* eps = tol * ||b|| */
#define CEL_TOL_REL_R0 2  /* eps = tol * ||b - A x0|| */

#define CEL_PREC_FP64 0
#define CEL_PREC_MIXED 1  /* FP32 inner PCG, FP64 residual/solution (no reference counterpart) */

/* CSR -> JAD diagonal lengths (src/fspmtconverter/cel_sp_mat_jad_converter_module.f90:134-143) */
#define CEL_JAD_SPARSKIT 0
#define CEL_JAD_CEL_COMPAT 1

typedef struct cel_gpu_mat cel_gpu_mat;

/* Solver parameters: the fields of cel_cgalg_parameter_type
 * (src/fcgalg/cel_alg_module.f90:26-40) plus the choices the reference
 * hard-codes.  cel_gpu_pcg_params_default() fills the reference settings
 * (x0 = 1, absolute tolerance, CSR, FP64). */
typedef struct {
    double tol;        /* eps (absolute) or relative factor, see tol_mode */
    int tol_mode;      /* CEL_TOL_* */
    int iter_max;      /* cel_cgalg_parameter_type%iter_max */
    int x0_mode;       /* CEL_X0_* */
    int precision;     /* CEL_PREC_* */
    int format;        /* CEL_FMT_CSR or CEL_FMT_JAD for the local block */
    int memspace;      /* where b and x live */
    int use_halo;      /* 1: pack the send faces and add halo blocks every SpMV */
    double ru_delta;   /* mixed: replace the FP32 residual by the FP64 one when it
                          has dropped by this factor (default 0.1) */
    int ru_restart;    /* mixed: 1 also restarts p (classic iterative refinement) */
    int check_every;   /* iterations per host poll of the device state (default 16) */
    int use_graph;     /* 1: replay one captured chunk of iterations as a hipGraph */
    int profile;       /* 1: time every SpMV launch with HIP events (stats) */
    int alpha_stop;    /* 1 (reference): also stop when |alpha| <= 1.1 eps_mach, an absolute
                          test (cel_cgalg_module.f90:248); 0: only the norm test */
} cel_pcg_params;

/* Per-solve results.  The reference returns iter_done and norm2
 * (cel_cgalg_module.f90:302-304) and prints the history at E12.6. */
typedef struct {
    int status;          /* CEL_OK / CEL_NOT_CONVERGED / error code */
    int iters;           /* iter_done (inner iterations for mixed) */
    int outer_iters;     /* mixed: number of FP64 residual replacements */
    double norm0;        /* ||b - A x0|| */
    double final_norm;   /* norm2 at exit (FP64 true residual for mixed) */
    double eps_abs;      /* absolute threshold used */
    double alpha;        /* final alpha (r.z) */
    double solve_ms;     /* device time of the iteration loop (HIP events) */
    double setup_ms;     /* device time of r0/p0 setup */
    double spmv_ms;      /* profile=1: summed duration of the SpMV launches */
    int spmv_launches;   /* profile=1: number of timed SpMV launches */
    double spmv_bytes;   /* algorithmic bytes of one SpMV of the iteration */
    double iter_bytes;   /* algorithmic bytes of one full iteration */
} cel_pcg_stats;

/* ---- runtime ----------------------------------------------------------- */
/* Replaces MPI_Init / cel_omp_init set-up (src/fbase/cel_omp_module.f90:27-91):
 * selects the GPU and creates the library stream. */
int cel_gpu_init(int device);
int cel_gpu_finalize(void);
const char* cel_gpu_last_error(void);
int cel_gpu_device_info(char* name, int name_len, int* num_cus);

/* device memory for callers that keep inputs resident in HBM */
int cel_gpu_malloc(void** ptr, size_t bytes);
int cel_gpu_free(void* ptr);
int cel_gpu_memcpy_h2d(void* dst, const void* src, size_t bytes);
int cel_gpu_memcpy_d2h(void* dst, const void* src, size_t bytes);
int cel_gpu_memcpy_d2d(void* dst, const void* src, size_t bytes);
int cel_gpu_synchronize(void);

/* ---- matrices ---------------------------------------------------------- */
/* CSR local block.  Replaces cel_sp_mat_type (src/fspmt/cel_sp_mat_module.f90:38-71)
 * holding xadj/cols/values with col_offset subtracted on use
 * (src/fop/cel_fop_mv_module.f90:219-224).  Arrays are copied; the caller
 * keeps ownership.  Validates ranges and row order. */
int cel_gpu_mat_create_csr(cel_gpu_mat** A, int nrows, int ncols, int nnz, const int* xadj,
                           const int* cols, const double* vals, int index_base, int row_offset,
                           int col_offset, int memspace);

/* Row-sorted COO.  Replaces cel_sp_mat_to_csr (src/fspmt/cel_sp_mat_module.f90:430-456):
 * xadj from row counts, entries not moved. */
int cel_gpu_mat_create_coo(cel_gpu_mat** A, int nrows, int ncols, int nnz, const int* rows,
                           const int* cols, const double* vals, int index_base, int row_offset,
                           int col_offset);

/* Device generators of the closed-form test matrices (no reference
 * counterpart for this matrix: the reference generator is
 * src/base/cel_discretisation_poisson_3d.c, which scales by 1/h^2 and keeps
 * Dirichlet identity rows).  poisson7: diag 6, off-diagonal -1, Dirichlet
 * outside, cell (x,y,z) -> row x + nx*(y + ny*z).  diffusion7: harmonic
 * mean face coefficients of kappa, boundary faces use kappa_i. */
int cel_gpu_mat_gen_poisson7(cel_gpu_mat** A, int nx, int ny, int nz);
int cel_gpu_mat_gen_diffusion7(cel_gpu_mat** A, int nx, int ny, int nz, const double* kappa,
                               int memspace);

int cel_gpu_mat_destroy(cel_gpu_mat* A);
int cel_gpu_mat_info(const cel_gpu_mat* A, int* nrows, int* ncols, int* nnz, int* jad_mmax,
                     int* nblocks);
/* 1-based host copies of the local block (xadj n+1, cols/vals nnz). */
int cel_gpu_mat_get_csr(const cel_gpu_mat* A, int* xadj, int* cols, double* vals);

/* ---- Matrix Market files, COO ordering, COO -> CSR --------------------- */
#define CEL_MM_COORDINATE 0
#define CEL_MM_ARRAY 1
#define CEL_MM_REAL 0
#define CEL_MM_COMPLEX 1
#define CEL_MM_PATTERN 2
#define CEL_MM_GENERAL 0
#define CEL_MM_SYMMETRIC 1
#define CEL_MM_HERMITIAN 2
#define CEL_MM_SKEW_SYMMETRIC 3
#define CEL_MM_AS_STORED 0         /* the reference: entries exactly as in the file */
#define CEL_MM_EXPAND_SYMMETRIC 1  /* also append (j,i) for every stored (i,j), i != j
                                      (value negated for skew-symmetric); no
                                      reference counterpart */

/* mminfo (src/fbase/cel_mmio_module.f90:342-466): header and size line.
 * Host only (no GPU call).  nnz is the stored entry count (array files:
 * the reference's count for the symmetry).  Any output may be NULL. */
int cel_gpu_mm_info(const char* path, int* rep, int* field, int* symm, int* nrows, int* ncols,
                    int* nnz);
/* cel_sp_mat_read_matrix_market_file + mmread
 * (src/fspmt/cel_sp_mat_generate_module.f90:219-273, src/fbase/cel_mmio_module.f90:17-339):
 * the coordinate entries in file order, indices as written (1-based, not
 * range-checked, as the reference), real values as read.  Host only.
 * Pattern files leave vals untouched (the reference leaves its values
 * unset).  Array files and complex files of more than one entry return
 * CEL_ERR_REF_UNDEFINED (the reference sets no indices for the first and
 * overruns a one-element buffer for the second); 'integer' fields are
 * rejected with CEL_ERR_FILE as the reference rejects them.  cap is the
 * capacity of rows/cols/vals; *nnz_out receives the entries written. */
int cel_gpu_mm_read(const char* path, int flags, int cap, int* rows, int* cols, double* vals,
                    int* nnz_out);

#define CEL_SORT_ROW_COL 0     /* stable by (row, col) */
#define CEL_SORT_CEL_COMPAT 1  /* cel_sp_mat_format_sparse_31_sort, quirks included */

/* cel_sp_mat_format_sparse_31_sort (src/fspmt/cel_sp_mat_module.f90:518-585),
 * in place on host or device arrays, sorted in HBM.  Row i of the block is
 * rows = i + index_base + row_offset (i < nrows), columns likewise with
 * col_offset and ncols; anything else is CEL_ERR_BAD_MATRIX.
 * CEL_SORT_CEL_COMPAT gives the reference's order entry for entry: a stable
 * sort by row, then stable by column within each segment its loop closes,
 * so the last row keeps its file order, a first entry whose row value is
 * not 1 stays first, and a last row of one entry is column-sorted together
 * with the row before it.  Where its loop reads past the arrays (fewer row
 * changes than nrows-1, e.g. an empty row) the arrays are left sorted by
 * row and CEL_ERR_REF_UNDEFINED is returned.  The sort keeps its device
 * workspace (about 50 bytes per entry; up to 90 with host arrays or column
 * groups too long for the in-LDS sort; shared with cel_gpu_coo_row_offsets
 * and cel_gpu_mat_create_coo_sorted) from call to call until
 * cel_gpu_finalize; calls that overlap from several host threads race. */
int cel_gpu_coo_sort(int nrows, int ncols, int nnz, int* rows, int* cols, double* vals,
                     int index_base, int row_offset, int col_offset, int mode, int memspace);
/* cel_sp_mat_to_csr (src/fspmt/cel_sp_mat_module.f90:430-456): xadj from
 * row counts (entries are not moved), xadj[0] = index_base; nrows+1
 * entries.  A row outside the block is CEL_ERR_BAD_MATRIX. */
int cel_gpu_coo_row_offsets(int nrows, int nnz, const int* rows, int index_base, int row_offset,
                            int* xadj, int memspace);
/* Sort (as cel_gpu_coo_sort), count rows, and build the matrix, all in
 * HBM: what the reference does between reading a COO and its CSR SpMV
 * (src/felmsolver/cel_elmsolver_module.f90:372, cel_sp_mat_to_csr).  Rows
 * that decrease after the sort (a CEL_SORT_CEL_COMPAT quirk on a last row
 * of one entry) are CEL_ERR_BAD_MATRIX. */
int cel_gpu_mat_create_coo_sorted(cel_gpu_mat** A, int nrows, int ncols, int nnz, const int* rows,
                                  const int* cols, const double* vals, int index_base,
                                  int row_offset, int col_offset, int mode, int memspace);

/* ---- CSR -> JAD on the device ------------------------------------------ */
/* Replaces cel_sp_mat_jad_from_sp_mat_local + cel_sp_mat_jad_dcsort
 * (src/fspmtconverter/cel_sp_mat_jad_converter_module.f90:57-175).
 * mode CEL_JAD_SPARSKIT: diagonal d holds #{rows with len >= d} entries.
 * mode CEL_JAD_CEL_COMPAT: the reference's counts (dcsort leftovers); when
 * the shortest row is non-empty the reference reads uninitialised counts,
 * so only the defined leading diagonals are built and CEL_ERR_REF_UNDEFINED
 * is returned (the JAD SpMV then refuses the matrix). */
int cel_gpu_mat_to_jad(cel_gpu_mat* A, int mode);
/* 1-based copies: perm (n), begin (mmax+1, -1 where undefined), cols/vals
 * (entries written).  *ndefined = diagonals whose data is defined. */
int cel_gpu_mat_get_jad(const cel_gpu_mat* A, int* perm, int* begin, int* cols, double* vals,
                        int* mmax, int* ndefined, long long* nnz_written);

/* ---- CSR -> DIA on the device (no reference counterpart) --------------- */
/* For a block whose entries lie on at most 32 diagonals (offsets col - row,
 * e.g. 7/19/27-point stencils) and whose rows hold strictly ascending
 * columns: values stored per diagonal (diagonal-major, n entries each,
 * absent entries marked in a per-row mask), offsets ascending.  A DIA row
 * sum visits the row's entries in the CSR column order and skips absent
 * ones, so y = A x is bit-identical to the CSR product, with no column
 * indices read and x read contiguously per diagonal.  Fails with
 * CEL_ERR_BAD_MATRIX (and builds nothing) for more than 32 diagonals,
 * unsorted or repeated columns, more than twice nnz stored slots, or more
 * than 2^28 rows. */
int cel_gpu_mat_to_dia(cel_gpu_mat* A);
/* Number of diagonals and their offsets (ascending; up to 32). */
int cel_gpu_mat_dia_info(const cel_gpu_mat* A, int* ndiag, int* offsets);

/* ---- diagonal, Jacobi vector, row scaling ------------------------------ */
/* cel_sp_mat_distr_set_jacobi (src/fspmtdistr/cel_sp_mat_distr_format_module.f90:149-203):
 * jacobi = 1/d, or sqrt|d|/d when row_scaled.  d is the diagonal captured
 * before any scaling (cel_sp_mat_distr_module.f90:197-207). */
int cel_gpu_mat_set_jacobi(cel_gpu_mat* A, int row_scaled);
/* cel_sp_mat_distr_scale (cel_sp_mat_distr_format_module.f90:91-145,276-356):
 * every local and halo entry of row i divided by sqrt|d_i|. */
int cel_gpu_mat_scale_diagonal(cel_gpu_mat* A);
/* cel_sp_mat_distr_boundary_diag_scale (:207-239): b_i = b_i / sqrt|d_i|. */
int cel_gpu_vec_scale_diagonal(const cel_gpu_mat* A, double* b, int memspace);
int cel_gpu_mat_get_jacobi(const cel_gpu_mat* A, double* jac);
int cel_gpu_mat_get_diagonal(const cel_gpu_mat* A, double* diag);

/* ---- SpMV -------------------------------------------------------------- */
/* cel_fop_mv_axy_distr_worker_local_{csr,jad} (src/fop/cel_fop_mv_module.f90:189-352):
 * y = A x (accumulate = 0) or y += A x (accumulate = 1), each row summed in
 * column order from 0.0 and then added to y.  The reference CSR kernel
 * accumulates; its JAD kernel overwrites (:323-326).  x has ncols entries. */
int cel_gpu_spmv(cel_gpu_mat* A, int format, const double* x, double* y, int accumulate,
                 int memspace);
/* Same in FP32 (FP32 values, x, y and accumulation): the mixed-precision
 * inner operator.  No reference counterpart. */
int cel_gpu_spmv_f32(cel_gpu_mat* A, int format, const float* x, float* y, int accumulate,
                     int memspace);
/* Times `reps` back-to-back launches of the device SpMV with HIP events on
 * the library stream (x, y device pointers).  Returns the mean launch time. */
int cel_gpu_spmv_time(cel_gpu_mat* A, int format, int precision, const void* x, void* y,
                      int reps, double* mean_ms, double* bytes);

/* ---- distributed block split, face packing, halo SpMV ------------------ */
/* cel_sp_mat_distr_from_sp_mat_local (src/fspmtdistr/cel_sp_mat_distr_module.f90:34-301):
 * this rank's row-sorted COO (global ids) is split into the local block
 * (columns owned by my_rank, becomes CSR) and one halo block per other
 * owning rank, in ascending rank order, entries kept in input order.
 * Each halo block's receive list is the ascending list of its referenced
 * columns (com_buffer_perm, :270-280).  vtxdist has nranks+1 entries with
 * the reference's 1-based row starts. */
int cel_gpu_mat_distr_split(cel_gpu_mat** A, int nranks, const int* vtxdist, int my_rank,
                            int nnz, const int* rows, const int* cols, const double* vals,
                            int index_base);
/* Block b (0 = local) of the split, reference layout (global 1-based ids):
 * proc, entries, rows/cols/vals (NULL to skip), xadj (n+1, NULL to skip),
 * recv_perm (receive list, blocks >= 1, NULL to skip). */
int cel_gpu_mat_get_block(const cel_gpu_mat* A, int b, int* proc, int* nnz, int* rows, int* cols,
                          double* vals, int* xadj, int* recv_len, int* recv_perm);
/* Appends one halo block to a local block built by any constructor (e.g. a
 * generated subdomain): the same per-rank block the split would create for
 * owner `proc` (cel_sp_mat_distr_module.f90:117-144, :249).  rows are this
 * rank's local rows, cols the owner's local rows (both index_base, row-
 * sorted); proc_first_row is vtxdist(proc+1) (1-based), proc_len the owner's
 * row count.  Blocks must be added in ascending rank order. */
int cel_gpu_mat_add_halo_block(cel_gpu_mat* A, int proc, int proc_first_row, int proc_len,
                               int nnz, const int* rows, const int* cols, const double* vals,
                               int index_base);
/* Send lists: the owner packs, for each peer, the entries that peer listed
 * (cel_comm_get_comm_buffer_connections, src/fcomm/cel_comm_module.f90:312-357).
 * perm is 1-based local indices. */
int cel_gpu_mat_set_send_list(cel_gpu_mat* A, int peer, int len, const int* perm);
/* Face packing (cel_sp_mat_distr_vec_module.f90:482-486): for every send
 * list in peer order, sendbuf(j) = x(perm(j)); lists are concatenated. */
int cel_gpu_face_pack(cel_gpu_mat* A, const double* x, double* sendbuf, int memspace);
/* Distributed SpMV (cel_fop_mv_axy_distr, src/fop/cel_fop_mv_module.f90:28-87):
 * y = A_loc x, then for every halo block in order y += A_b ghost_b, where
 * ghost holds the received values of all halo blocks concatenated in block
 * order (receive-list order).  The sum order of each row matches the
 * reference: local row sum first, then halo entries block by block. */
int cel_gpu_spmv_distr(cel_gpu_mat* A, int format, const double* x, const double* ghost,
                       double* y, int memspace);
/* The Jacobi-PCG's fused product on caller vectors: y = A_loc x in the CSR
 * or DIA form plus, when ghost is not NULL, the halo blocks' products of the
 * ghost values (as cel_gpu_spmv_distr: local row sum first, then the halo
 * entries in block order), and *dot = x . y over the local rows.  It runs
 * the kernel the FP64 solve runs once per iteration (the reference's
 * v = A p and the v.p dot of cel_fop_vec_distr_alpha_allreduce in
 * cel_cgalg_jacobi, src/fcgalg/cel_cgalg_module.f90:250-261), so y is
 * bitwise the product of cel_gpu_spmv_distr and dot is summed in the
 * solve's fixed order.  FP64, square local block. */
int cel_gpu_spmv_dot(cel_gpu_mat* A, int format, const double* x, const double* ghost, double* y,
                     double* dot, int memspace);
int cel_gpu_mat_ghost_len(const cel_gpu_mat* A, int* ghost_len, int* send_total);

/* Mean time of a streaming device copy of `bytes` (read + write), for the
 * measured bandwidth ceiling next to the 8 TB/s nominal peak.  No reference
 * counterpart. */
int cel_gpu_stream_copy_time(size_t bytes, int reps, double* mean_ms);
/* Mean time of reading `bytes` from HBM as `streams` (1-8) equal arrays read
 * together (16 bytes per lane and array per step, blocks_per_cu blocks of
 * 256 threads on every CU), for the read-bandwidth ceiling of the SpMV
 * kernels, which read several arrays at once.  No reference counterpart. */
int cel_gpu_stream_read_time(size_t bytes, int streams, int blocks_per_cu, int reps,
                             double* mean_ms);

/* ---- Jacobi-PCG -------------------------------------------------------- */
void cel_gpu_pcg_params_default(cel_pcg_params* p);
/* cel_cgalg_jacobi (src/fcgalg/cel_cgalg_module.f90:178-305), same op
 * order, with x0, tolerance mode and precision selectable.  hist (may be
 * NULL) receives norm2 after setup and after every iteration
 * (iter_max+1 doubles); for the mixed form it receives the inner estimate.
 * Requires cel_gpu_mat_set_jacobi.  With use_halo the halo blocks are
 * applied with a zero ghost region (one subdomain, no neighbours present)
 * and the send faces are packed every SpMV; FP64 only, in the CSR or the
 * DIA form, with fewer than 2^27 rows that have halo entries. */
int cel_gpu_pcg_solve(cel_gpu_mat* A, const double* b, double* x, const cel_pcg_params* params,
                      cel_pcg_stats* stats, double* hist);

/* No counterpart in the reference: cel_gpu_pcg_solve for callers that
 * cannot allocate (code under graph capture, tight serving loops, harnesses
 * that forbid hipMalloc).  It calls no hipMalloc, hipFree, hipHostMalloc,
 * event or graph routine: its work vectors, partial sums and token live in
 * a fixed device array of the library (about 24.4 MiB per GPU, held until
 * the process ends), resolved once per device by cel_gpu_init.  Limits:
 * FP64, CSR, COO or DIA form, no halo blocks, no history, b and x device
 * pointers (params->memspace must be CEL_MEM_DEVICE), at most 1,048,576
 * rows (CEL_ERR_ARG beyond) and 8,192 blocks per kernel.  The same bits as
 * cel_gpu_pcg_solve with use_graph = 0; solve_ms is host time around the
 * synchronised stream and setup_ms is 0.  Every call on one GPU shares that
 * GPU's array: calls that overlap on one GPU (several host threads, or a
 * stream other than the library's) race. */
int cel_gpu_pcg_solve_noalloc(cel_gpu_mat* A, const double* b, double* x,
                              const cel_pcg_params* params, cel_pcg_stats* stats);

/* ---- adjoint (no reference counterpart) -------------------------------- */
/* For A x = b with L(x): solves A^T lambda = dLdx (A is symmetric here, so
 * the same PCG is used), returns dLdb = lambda and, when dLda is not NULL,
 * dL/da_ij = -lambda_i x_j on the CSR pattern of the local block.  x is the
 * forward solution.  Vectors in memspace. */
int cel_gpu_cg_adjoint(cel_gpu_mat* A, const double* x, const double* dLdx, double* dLdb,
                       double* dLda, const cel_pcg_params* params, cel_pcg_stats* stats);
/* dL/dkappa of the diffusion7 matrix from x and lambda (chain rule through
 * the harmonic-mean face coefficients). */
int cel_gpu_diffusion7_kappa_grad(int nx, int ny, int nz, const double* kappa, const double* x,
                                  const double* lambda, double* grad, int memspace);

/* ---- batched Jacobi-PCG (no reference counterpart) --------------------- */
/* Many independent SPD systems -- uneven sizes, and blocks of right-hand
 * sides that share one matrix -- solved together by one ring of kernels
 * replayed as a hipGraph.  Each right-hand side is a lane with its own
 * scalars and the reference's op sequence (cel_cgalg_jacobi,
 * src/fcgalg/cel_cgalg_module.f90:178-305, called once per system there).
 * FP64, CSR, no halo blocks.  The matrices (set_jacobi done) must outlive
 * the batch.  nrhs[s] in 1..64 (NULL: all 1). */
typedef struct cel_gpu_batch cel_gpu_batch;
int cel_gpu_batch_create(cel_gpu_batch** B, int nsys, cel_gpu_mat* const* mats, const int* nrhs);
int cel_gpu_batch_destroy(cel_gpu_batch* B);
/* Times candidate launch configurations (the grid of the SpMV, the grid of
 * the two vector kernels, iterations per graph replay, block apportionment)
 * on `iters` iterations of the batch with the right-hand sides b and keeps
 * the fastest for later solves.  The choice changes how dot products are
 * grouped, so results agree with an untuned solve to rounding, not bit for
 * bit. */
int cel_gpu_batch_tune(cel_gpu_batch* B, const double* const* b, int memspace, int iters);
/* The configuration in use and the time per iteration the tuner measured
 * (0 before tuning).  weight 0: blocks dealt by entries x lanes, 1: by rows x lanes. */
int cel_gpu_batch_info(const cel_gpu_batch* B, int* lanes, long long* elements, int* grid,
                       int* chunk, int* weight, int* rebalance, double* ms_per_iter);
/* No reference counterpart.  The launch configuration of later solves:
 * grid = blocks of the SpMV and set-up kernels, grid23 = blocks of the two
 * vector kernels (their own block table), chunk = iterations per graph
 * replay, weight as above.  get returns the one in use (after
 * cel_gpu_batch_tune, the tuner's); set fixes one by hand, so that solves
 * on different runs or GPUs group their dot products alike (the
 * configuration changes the grouping, and so the rounding).  set accepts
 * grids that are multiples of 8, at least the number of systems and at
 * most the resident grid (grid23 = 0: grid), chunk 1..1024, weight 0 or 1;
 * anything else is CEL_ERR_ARG. */
int cel_gpu_batch_get_config(const cel_gpu_batch* B, int* grid, int* grid23, int* chunk,
                             int* weight);
int cel_gpu_batch_set_config(cel_gpu_batch* B, int grid, int grid23, int chunk, int weight);
/* b[s], x[s]: n_s x nrhs[s] column-major (right-hand side j at offset j*n_s),
 * host or device per params->memspace.  lane_stats (may be NULL) receives one
 * entry per right-hand side in system order; total: status (most severe),
 * iters (maximum), outer_iters (graph replays), solve_ms, iter_bytes. */
int cel_gpu_batch_solve(cel_gpu_batch* B, const double* const* b, double* const* x,
                        const cel_pcg_params* params, cel_pcg_stats* lane_stats,
                        cel_pcg_stats* total);

/* ---- Jacobi-GMRES ------------------------------------------------------ */
/* cel_gmresalg_parameter_type settings used by cel_gmresalg_2jacobi
 * (src/fcgalg/cel_alg_module.f90:139-271). */
typedef struct {
    double eps;     /* absolute: stop when the Givens estimate |gamma(j+1)| <= eps */
    int iter_max;   /* outer iterations; the loop runs while iter <= iter_max */
    int restart;    /* m: Arnoldi steps per cycle (1..32); m+1 cycles per outer iteration */
    int x0_mode;    /* CEL_X0_ONE is the reference's x = 1 */
    int memspace;   /* b and x on the host or the device */
    int format;     /* CEL_FMT_CSR, CEL_FMT_JAD or CEL_FMT_DIA */
} cel_gmres_params;

/* The element solver's settings (src/felmsolver/cel_elmsolver_module.f90:140-151):
 * eps 1e-8, iter_max 5000, restart 3, x = 1, host arrays, CSR. */
void cel_gpu_gmres_params_default(cel_gmres_params* p);

/* cel_gmresalg_2jacobi (src/fcgalg/cel_gmresalg_module.f90:570-851), one
 * rank, local block only: right Jacobi preconditioning (z = jac * v),
 * modified Gram-Schmidt, Givens rotations applied as the reference writes
 * them (:722-727: the second statement assigns hh(i,j) again, so hh(i+1,j)
 * is never rotated; with restart >= 2 the estimate can meet eps while the
 * true residual is larger, exactly as in the reference).  Returns CEL_OK
 * when the estimate met eps, CEL_NOT_CONVERGED after iter_max+1 outer
 * iterations.  stats: iters = outer iterations (the reference's
 * iter_done), outer_iters = restart cycles, norm0 = first residual,
 * final_norm = the true residual norm ||b - A x|| after the last cycle
 * (the reference's residuum_norm, returned as ftol by cel_elmSolverDoIt). */
int cel_gpu_gmres_solve(cel_gpu_mat* A, const double* b, double* x, const cel_gmres_params* params,
                        cel_pcg_stats* stats);

/* ---- element solver API (src/felmsolver/cel_elmsolver_module.f90) ------ */
/* One handle per system; on one GPU it holds the whole system (the
 * reference gathers a reduce group's rows on its master rank).  The Fortran
 * module cel_gpu_elmsolver keeps one handle and offers the reference's
 * cel_elmSolver* names over these calls. */
typedef struct cel_gpu_elm cel_gpu_elm;
#define CEL_ELM_GMRES 0  /* Jacobi-GMRES(3), the method the reference forces (:146) */
#define CEL_ELM_PCG 1    /* Jacobi-PCG (symmetric positive definite systems) */

/* cel_elmSolverInitialization (:103-153): itot global rows, ibs local rows,
 * ibl first local row, nnzero[0..1] entries per row (a capacity hint; the
 * arrays grow past it where the reference's would be overrun).  One GPU:
 * ibs == itot and ibl == 1, else CEL_ERR_ARG. */
int cel_gpu_elm_create(cel_gpu_elm** E, int itot, int ibs, int ibl, const int* nnzero);
/* cel_elmSolverDatarowStart (:155-159): rows 1..ibs. */
int cel_gpu_elm_row_start(cel_gpu_elm* E, int row);
/* cel_elmSolverDataInsertion (:161-172): appends (row, col, value); an
 * entry with col == row is the row's diagonal (the last one counts). */
int cel_gpu_elm_insert(cel_gpu_elm* E, int col, double value);
/* cel_elmSolverDataReinsertion (:174-187): replaces the value of the
 * current row's diagonal entry; other columns are CEL_ERR_ARG (the
 * reference stops), a row without a diagonal CEL_ERR_REF_UNDEFINED. */
int cel_gpu_elm_reinsert(cel_gpu_elm* E, int col, double value);
/* cel_elmSolverDatarowCompleted (:189-191): nothing to do. */
int cel_gpu_elm_row_completed(cel_gpu_elm* E);
/* cel_elmSolverMatrixCompleted (:193-439): the reference's COO sort
 * (cel_gpu_coo_sort order, :372), COO->CSR, diagonal captured, rows scaled
 * by 1/sqrt|d_i| and the Jacobi vector sqrt|d|/d (:386-419), all in HBM.
 * A matrix whose entries lie on at most 32 diagonals is also kept in the
 * DIA form, which the solves then use (bit-identical products).
 * Insertion after this is CEL_ERR_STATE (the reference frees its arrays). */
int cel_gpu_elm_matrix_completed(cel_gpu_elm* E);
/* Solver of the next solves (no reference counterpart; default GMRES). */
int cel_gpu_elm_set_solver(cel_gpu_elm* E, int solver);
/* cel_elmSolverDoIt (:441-651): b (ibs, host) scaled by 1/sqrt|d_i|, then
 * the solver from x = 1 with eps = 1e-8 (absolute), iter_max = 5000,
 * GMRES restart 3.  iierr = 0 (or the error code), niter = the solver's
 * iteration count, ftol = its final residual norm (GMRES: the true
 * residual of the scaled system).  Returns the solver's status. */
int cel_gpu_elm_solve(cel_gpu_elm* E, const double* b, int* iierr, int* niter, double* ftol);
/* cel_elmSolverReDoIt (:653-788): the same with a new b that is NOT scaled
 * by the diagonal -- the reference's ReDoIt omits that step. */
int cel_gpu_elm_resolve(cel_gpu_elm* E, const double* b, int* iierr, int* niter, double* ftol);
/* cel_elmSolverCollectSolution (:790-815): x (ibs, host). */
int cel_gpu_elm_collect_solution(const cel_gpu_elm* E, double* x);
/* cel_elmSolverPerformance (:879-900): the last solve's statistics. */
int cel_gpu_elm_stats(const cel_gpu_elm* E, cel_pcg_stats* s);
/* The assembled matrix (owned by the handle), for inspection. */
int cel_gpu_elm_matrix(const cel_gpu_elm* E, cel_gpu_mat** A);
/* cel_elmSolverDestroy (:817-877). */
int cel_gpu_elm_destroy(cel_gpu_elm* E);

#ifdef __cplusplus
}
#endif
#endif
