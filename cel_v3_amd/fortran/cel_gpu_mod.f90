! cel_gpu_mod.f90 -- Fortran access to the MI355X Jacobi-PCG library.
!
! module cel_gpu          : ISO_C_BINDING interfaces of include/cel_gpu.h
! module cel_gpu_elmsolver: the assembly/solve entry points of the reference's
!                           cel_elmsolver_module (src/felmsolver/
!                           cel_elmsolver_module.f90:85-89) on one GPU, so a
!                           caller written against that API links unchanged.
!
! Types follow the reference: INDEX = integer(4), REAL = real(8), 1-based rows
! and columns (src/fbase/cel_types_module.f90:13-14).
module cel_gpu
  use, intrinsic :: iso_c_binding
  implicit none

  integer(c_int), parameter :: CEL_OK = 0, CEL_NOT_CONVERGED = 1
  integer(c_int), parameter :: CEL_FMT_CSR = 1, CEL_FMT_JAD = 2
  integer(c_int), parameter :: CEL_MEM_HOST = 0, CEL_MEM_DEVICE = 1
  integer(c_int), parameter :: CEL_X0_ONE = 0, CEL_X0_ZERO = 1, CEL_X0_USER = 2
  integer(c_int), parameter :: CEL_TOL_ABS = 0, CEL_TOL_REL_B = 1, CEL_TOL_REL_R0 = 2
  integer(c_int), parameter :: CEL_PREC_FP64 = 0, CEL_PREC_MIXED = 1
  integer(c_int), parameter :: CEL_JAD_SPARSKIT = 0, CEL_JAD_CEL_COMPAT = 1
  integer(c_int), parameter :: CEL_SORT_ROW_COL = 0, CEL_SORT_CEL_COMPAT = 1
  integer(c_int), parameter :: CEL_ELM_GMRES = 0, CEL_ELM_PCG = 1

  type, bind(c) :: cel_pcg_params
    real(c_double) :: tol
    integer(c_int) :: tol_mode, iter_max, x0_mode, precision, format, memspace, use_halo
    real(c_double) :: ru_delta
    integer(c_int) :: ru_restart, check_every, use_graph, profile, alpha_stop
  end type cel_pcg_params

  type, bind(c) :: cel_pcg_stats
    integer(c_int) :: status, iters, outer_iters
    real(c_double) :: norm0, final_norm, eps_abs, alpha, solve_ms, setup_ms, spmv_ms
    integer(c_int) :: spmv_launches
    real(c_double) :: spmv_bytes, iter_bytes
  end type cel_pcg_stats

  type, bind(c) :: cel_gmres_params
    real(c_double) :: eps
    integer(c_int) :: iter_max, restart, x0_mode, memspace, format
  end type cel_gmres_params

  interface
    integer(c_int) function cel_gpu_init(device) bind(c, name="cel_gpu_init")
      import :: c_int
      integer(c_int), value :: device
    end function
    integer(c_int) function cel_gpu_finalize() bind(c, name="cel_gpu_finalize")
      import :: c_int
    end function
    type(c_ptr) function cel_gpu_last_error() bind(c, name="cel_gpu_last_error")
      import :: c_ptr
    end function
    integer(c_int) function cel_gpu_mat_create_csr(A, nrows, ncols, nnz, xadj, cols, vals, &
        index_base, row_offset, col_offset, memspace) bind(c, name="cel_gpu_mat_create_csr")
      import :: c_int, c_ptr, c_double
      type(c_ptr), intent(out) :: A
      integer(c_int), value :: nrows, ncols, nnz, index_base, row_offset, col_offset, memspace
      integer(c_int), intent(in) :: xadj(*), cols(*)
      real(c_double), intent(in) :: vals(*)
    end function
    integer(c_int) function cel_gpu_mat_create_coo(A, nrows, ncols, nnz, rows, cols, vals, &
        index_base, row_offset, col_offset) bind(c, name="cel_gpu_mat_create_coo")
      import :: c_int, c_ptr, c_double
      type(c_ptr), intent(out) :: A
      integer(c_int), value :: nrows, ncols, nnz, index_base, row_offset, col_offset
      integer(c_int), intent(in) :: rows(*), cols(*)
      real(c_double), intent(in) :: vals(*)
    end function
    integer(c_int) function cel_gpu_mat_gen_poisson7(A, nx, ny, nz) &
        bind(c, name="cel_gpu_mat_gen_poisson7")
      import :: c_int, c_ptr
      type(c_ptr), intent(out) :: A
      integer(c_int), value :: nx, ny, nz
    end function
    integer(c_int) function cel_gpu_mat_destroy(A) bind(c, name="cel_gpu_mat_destroy")
      import :: c_int, c_ptr
      type(c_ptr), value :: A
    end function
    integer(c_int) function cel_gpu_mat_to_jad(A, mode) bind(c, name="cel_gpu_mat_to_jad")
      import :: c_int, c_ptr
      type(c_ptr), value :: A
      integer(c_int), value :: mode
    end function
    integer(c_int) function cel_gpu_mat_set_jacobi(A, row_scaled) &
        bind(c, name="cel_gpu_mat_set_jacobi")
      import :: c_int, c_ptr
      type(c_ptr), value :: A
      integer(c_int), value :: row_scaled
    end function
    integer(c_int) function cel_gpu_mat_scale_diagonal(A) bind(c, name="cel_gpu_mat_scale_diagonal")
      import :: c_int, c_ptr
      type(c_ptr), value :: A
    end function
    integer(c_int) function cel_gpu_vec_scale_diagonal(A, b, memspace) &
        bind(c, name="cel_gpu_vec_scale_diagonal")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: A
      real(c_double), intent(inout) :: b(*)
      integer(c_int), value :: memspace
    end function
    integer(c_int) function cel_gpu_spmv(A, format, x, y, accumulate, memspace) &
        bind(c, name="cel_gpu_spmv")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: A
      integer(c_int), value :: format, accumulate, memspace
      real(c_double), intent(in) :: x(*)
      real(c_double), intent(inout) :: y(*)
    end function
    subroutine cel_gpu_pcg_params_default(p) bind(c, name="cel_gpu_pcg_params_default")
      import :: cel_pcg_params
      type(cel_pcg_params), intent(out) :: p
    end subroutine
    integer(c_int) function cel_gpu_pcg_solve(A, b, x, params, stats, hist) &
        bind(c, name="cel_gpu_pcg_solve")
      import :: c_int, c_ptr, c_double, cel_pcg_params, cel_pcg_stats
      type(c_ptr), value :: A
      real(c_double), intent(in) :: b(*)
      real(c_double), intent(inout) :: x(*)
      type(cel_pcg_params), intent(in) :: params
      type(cel_pcg_stats), intent(out) :: stats
      type(c_ptr), value :: hist
    end function
    integer(c_int) function cel_gpu_mat_gen_diffusion7(A, nx, ny, nz, kappa, memspace) &
        bind(c, name="cel_gpu_mat_gen_diffusion7")
      import :: c_int, c_ptr, c_double
      type(c_ptr), intent(out) :: A
      integer(c_int), value :: nx, ny, nz, memspace
      real(c_double), intent(in) :: kappa(*)
    end function
    subroutine cel_gpu_gmres_params_default(p) bind(c, name="cel_gpu_gmres_params_default")
      import :: cel_gmres_params
      type(cel_gmres_params), intent(out) :: p
    end subroutine
    integer(c_int) function cel_gpu_gmres_solve(A, b, x, params, stats) &
        bind(c, name="cel_gpu_gmres_solve")
      import :: c_int, c_ptr, c_double, cel_gmres_params, cel_pcg_stats
      type(c_ptr), value :: A
      real(c_double), intent(in) :: b(*)
      real(c_double), intent(inout) :: x(*)
      type(cel_gmres_params), intent(in) :: params
      type(cel_pcg_stats), intent(out) :: stats
    end function
    integer(c_int) function cel_gpu_mat_create_coo_sorted(A, nrows, ncols, nnz, rows, cols, vals, &
        index_base, row_offset, col_offset, mode, memspace) bind(c, name="cel_gpu_mat_create_coo_sorted")
      import :: c_int, c_ptr, c_double
      type(c_ptr), intent(out) :: A
      integer(c_int), value :: nrows, ncols, nnz, index_base, row_offset, col_offset, mode, memspace
      integer(c_int), intent(in) :: rows(*), cols(*)
      real(c_double), intent(in) :: vals(*)
    end function
    integer(c_int) function cel_gpu_elm_create(E, itot, ibs, ibl, nnzero) bind(c, name="cel_gpu_elm_create")
      import :: c_int, c_ptr
      type(c_ptr), intent(out) :: E
      integer(c_int), value :: itot, ibs, ibl
      integer(c_int), intent(in) :: nnzero(2)
    end function
    integer(c_int) function cel_gpu_elm_row_start(E, row) bind(c, name="cel_gpu_elm_row_start")
      import :: c_int, c_ptr
      type(c_ptr), value :: E
      integer(c_int), value :: row
    end function
    integer(c_int) function cel_gpu_elm_insert(E, col, val) bind(c, name="cel_gpu_elm_insert")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: E
      integer(c_int), value :: col
      real(c_double), value :: val
    end function
    integer(c_int) function cel_gpu_elm_reinsert(E, col, val) bind(c, name="cel_gpu_elm_reinsert")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: E
      integer(c_int), value :: col
      real(c_double), value :: val
    end function
    integer(c_int) function cel_gpu_elm_row_completed(E) bind(c, name="cel_gpu_elm_row_completed")
      import :: c_int, c_ptr
      type(c_ptr), value :: E
    end function
    integer(c_int) function cel_gpu_elm_matrix_completed(E) bind(c, name="cel_gpu_elm_matrix_completed")
      import :: c_int, c_ptr
      type(c_ptr), value :: E
    end function
    integer(c_int) function cel_gpu_elm_set_solver(E, solver) bind(c, name="cel_gpu_elm_set_solver")
      import :: c_int, c_ptr
      type(c_ptr), value :: E
      integer(c_int), value :: solver
    end function
    integer(c_int) function cel_gpu_elm_solve(E, b, iierr, niter, ftol) bind(c, name="cel_gpu_elm_solve")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: E
      real(c_double), intent(in) :: b(*)
      integer(c_int), intent(out) :: iierr, niter
      real(c_double), intent(out) :: ftol
    end function
    integer(c_int) function cel_gpu_elm_resolve(E, b, iierr, niter, ftol) bind(c, name="cel_gpu_elm_resolve")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: E
      real(c_double), intent(in) :: b(*)
      integer(c_int), intent(out) :: iierr, niter
      real(c_double), intent(out) :: ftol
    end function
    integer(c_int) function cel_gpu_elm_collect_solution(E, x) bind(c, name="cel_gpu_elm_collect_solution")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: E
      real(c_double), intent(out) :: x(*)
    end function
    integer(c_int) function cel_gpu_elm_stats(E, s) bind(c, name="cel_gpu_elm_stats")
      import :: c_int, c_ptr, cel_pcg_stats
      type(c_ptr), value :: E
      type(cel_pcg_stats), intent(out) :: s
    end function
    integer(c_int) function cel_gpu_elm_destroy(E) bind(c, name="cel_gpu_elm_destroy")
      import :: c_int, c_ptr
      type(c_ptr), value :: E
    end function
    integer(c_int) function cel_gpu_cg_adjoint(A, x, dLdx, dLdb, dLda, params, stats) &
        bind(c, name="cel_gpu_cg_adjoint")
      import :: c_int, c_ptr, c_double, cel_pcg_params, cel_pcg_stats
      type(c_ptr), value :: A
      real(c_double), intent(in) :: x(*), dLdx(*)
      real(c_double), intent(out) :: dLdb(*)
      type(c_ptr), value :: dLda
      type(cel_pcg_params), intent(in) :: params
      type(cel_pcg_stats), intent(out) :: stats
    end function
    integer(c_int) function cel_gpu_diffusion7_kappa_grad(nx, ny, nz, kappa, x, lambda, grad, &
        memspace) bind(c, name="cel_gpu_diffusion7_kappa_grad")
      import :: c_int, c_double
      integer(c_int), value :: nx, ny, nz, memspace
      real(c_double), intent(in) :: kappa(*), x(*), lambda(*)
      real(c_double), intent(out) :: grad(*)
    end function
    ! batched solves: mats, b and x are arrays of C pointers (c_loc of each
    ! system's matrix handle / n x nrhs column-major array)
    integer(c_int) function cel_gpu_batch_create(B, nsys, mats, nrhs) &
        bind(c, name="cel_gpu_batch_create")
      import :: c_int, c_ptr
      type(c_ptr), intent(out) :: B
      integer(c_int), value :: nsys
      type(c_ptr), intent(in) :: mats(*)
      integer(c_int), intent(in) :: nrhs(*)
    end function
    integer(c_int) function cel_gpu_batch_destroy(B) bind(c, name="cel_gpu_batch_destroy")
      import :: c_int, c_ptr
      type(c_ptr), value :: B
    end function
    integer(c_int) function cel_gpu_batch_tune(B, bptrs, memspace, iters) &
        bind(c, name="cel_gpu_batch_tune")
      import :: c_int, c_ptr
      type(c_ptr), value :: B
      type(c_ptr), intent(in) :: bptrs(*)
      integer(c_int), value :: memspace, iters
    end function
    integer(c_int) function cel_gpu_batch_solve(B, bptrs, xptrs, params, lane_stats, total) &
        bind(c, name="cel_gpu_batch_solve")
      import :: c_int, c_ptr, cel_pcg_params, cel_pcg_stats
      type(c_ptr), value :: B
      type(c_ptr), intent(in) :: bptrs(*), xptrs(*)
      type(cel_pcg_params), intent(in) :: params
      type(cel_pcg_stats), intent(out) :: lane_stats(*)
      type(cel_pcg_stats), intent(out) :: total
    end function
    ! the launch configuration of later batch solves (grids of the SpMV and
    ! of the vector kernels, iterations per graph replay, weighting)
    integer(c_int) function cel_gpu_batch_get_config(B, grid, grid23, chunk, weight) &
        bind(c, name="cel_gpu_batch_get_config")
      import :: c_int, c_ptr
      type(c_ptr), value :: B
      integer(c_int), intent(out) :: grid, grid23, chunk, weight
    end function
    integer(c_int) function cel_gpu_batch_set_config(B, grid, grid23, chunk, weight) &
        bind(c, name="cel_gpu_batch_set_config")
      import :: c_int, c_ptr
      type(c_ptr), value :: B
      integer(c_int), value :: grid, grid23, chunk, weight
    end function
  end interface

contains

  ! message of the last failing call as a Fortran string
  function cel_gpu_error_message() result(msg)
    character(len=:), allocatable :: msg
    type(c_ptr) :: p
    character(kind=c_char), pointer :: s(:)
    integer :: n
    p = cel_gpu_last_error()
    call c_f_pointer(p, s, [4096])
    n = 0
    do while (n < 4096)
      if (s(n + 1) == c_null_char) exit
      n = n + 1
    end do
    allocate(character(len=n) :: msg)
    msg = transfer(s(1:n), msg)
  end function

end module cel_gpu


module cel_gpu_elmsolver
  ! The reference's element solver API (cel_elmsolver_module.f90:85-89,
  ! 103-900) with the same names and arguments, over one cel_gpu_elm handle
  ! of the C ABI: rows are inserted one at a time, MatrixCompleted sorts them
  ! in the reference's order and builds the diagonally scaled matrix and its
  ! Jacobi vector in HBM, DoIt/ReDoIt solve with the reference's
  ! Jacobi-GMRES(3) (x = 1, eps = 1e-8 absolute, iter_max = 5000; DoIt
  ! scales b by the diagonal, ReDoIt does not, as in the reference) and
  ! CollectSolution returns x.  One GPU holds the whole system, so the MPI
  ! communicator is accepted and not used.  cel_elmSolverSetSolver (no
  ! reference counterpart) switches to Jacobi-PCG for SPD systems.
  use, intrinsic :: iso_c_binding
  use cel_gpu
  implicit none
  private
  public :: cel_elmSolverInitialization, cel_elmSolverDatarowStart, cel_elmSolverDataInsertion, &
            cel_elmSolverDataReinsertion, cel_elmSolverDatarowCompleted, &
            cel_elmSolverMatrixCompleted, cel_elmSolverDoIt, cel_elmSolverReDoIt, &
            cel_elmSolverCollectSolution, cel_elmSolverDestroy, cel_elmSolverPerformance, &
            cel_elmSolverSetSolver

  integer, parameter :: ik = 4, rk = 8
  type(c_ptr), save :: elm = c_null_ptr
  integer(ik), save :: elm_n = 0

contains

  subroutine report(rc, where)
    integer(c_int), intent(in) :: rc
    character(len=*), intent(in) :: where
    if (rc < 0) write(0, '(A,A,A)') where, ': ', cel_gpu_error_message()
  end subroutine

  subroutine cel_elmSolverInitialization(itotSize, ibs, ibl, nnzero, comm)
    integer(ik), intent(in) :: itotSize, ibs, ibl, nnzero(2)
    integer(4), intent(in) :: comm
    integer(c_int) :: rc
    if (comm < -huge(comm)) return  ! the communicator is not used on one GPU
    rc = cel_gpu_init(0_c_int)
    call report(rc, 'cel_elmSolverInitialization')
    rc = cel_gpu_elm_create(elm, itotSize, ibs, ibl, nnzero)
    call report(rc, 'cel_elmSolverInitialization')
    elm_n = ibs
  end subroutine

  subroutine cel_elmSolverDatarowStart(row)
    integer(ik), intent(in) :: row
    call report(cel_gpu_elm_row_start(elm, row), 'cel_elmSolverDatarowStart')
  end subroutine

  subroutine cel_elmSolverDataInsertion(col, value)
    integer, intent(in) :: col
    real(rk), intent(in) :: value
    call report(cel_gpu_elm_insert(elm, col, value), 'cel_elmSolverDataInsertion')
  end subroutine

  subroutine cel_elmSolverDataReinsertion(col, value)
    integer(ik), intent(in) :: col
    real(rk), intent(in) :: value
    call report(cel_gpu_elm_reinsert(elm, col, value), 'cel_elmSolverDataReinsertion')
  end subroutine

  subroutine cel_elmSolverDatarowCompleted()
    call report(cel_gpu_elm_row_completed(elm), 'cel_elmSolverDatarowCompleted')
  end subroutine

  subroutine cel_elmSolverMatrixCompleted()
    call report(cel_gpu_elm_matrix_completed(elm), 'cel_elmSolverMatrixCompleted')
  end subroutine

  subroutine cel_elmSolverSetSolver(solver)
    integer(ik), intent(in) :: solver
    call report(cel_gpu_elm_set_solver(elm, solver), 'cel_elmSolverSetSolver')
  end subroutine

  subroutine cel_elmSolverDoIt(bmat, iierr, niter, ftol)
    real(rk), intent(in) :: bmat(:)
    integer(ik), intent(out) :: iierr, niter
    real(rk), intent(out) :: ftol
    call report(cel_gpu_elm_solve(elm, bmat, iierr, niter, ftol), 'cel_elmSolverDoIt')
  end subroutine

  subroutine cel_elmSolverReDoIt(bmat, iierr, niter, ftol)
    real(rk), intent(in) :: bmat(:)
    integer(ik), intent(out) :: iierr, niter
    real(rk), intent(out) :: ftol
    call report(cel_gpu_elm_resolve(elm, bmat, iierr, niter, ftol), 'cel_elmSolverReDoIt')
  end subroutine

  subroutine cel_elmSolverCollectSolution(bmat)
    real(rk), allocatable, intent(inout) :: bmat(:)
    if (allocated(bmat)) deallocate(bmat)
    allocate(bmat(elm_n))
    call report(cel_gpu_elm_collect_solution(elm, bmat), 'cel_elmSolverCollectSolution')
  end subroutine

  subroutine cel_elmSolverDestroy()
    integer(c_int) :: rc
    if (c_associated(elm)) rc = cel_gpu_elm_destroy(elm)
    elm = c_null_ptr
  end subroutine

  subroutine cel_elmSolverPerformance()
    type(cel_pcg_stats) :: s
    integer(c_int) :: rc
    rc = cel_gpu_elm_stats(elm, s)
    write(*, '(A,I0,A,I0,A,ES13.6,A,F12.3,A)') 'iterations: ', s%iters, '; cycles: ', &
      s%outer_iters, '; |r|: ', s%final_norm, '; solve: ', s%solve_ms, ' ms'
  end subroutine

end module cel_gpu_elmsolver
