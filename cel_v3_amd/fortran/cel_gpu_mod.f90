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
  ! One-GPU counterpart of the reference's element solver API: rows are
  ! inserted one at a time (cel_elmsolver_module.f90:155-189), the matrix is
  ! assembled once (:193-439) and solved any number of times (:441, :653).
  ! The reference forces Jacobi-GMRES with diagonal row scaling (:140-151);
  ! this layer keeps the scaling and the absolute eps = 1e-8 / iter_max = 5000
  ! settings and runs Jacobi-PCG (GMRES is not part of this library), so it
  ! serves symmetric positive definite systems.
  use, intrinsic :: iso_c_binding
  use cel_gpu
  implicit none
  private
  public :: cel_elmSolverInitialization, cel_elmSolverDatarowStart, cel_elmSolverDataInsertion, &
            cel_elmSolverDataReinsertion, cel_elmSolverDatarowCompleted, &
            cel_elmSolverMatrixCompleted, cel_elmSolverDoIt, cel_elmSolverReDoIt, &
            cel_elmSolverCollectSolution, cel_elmSolverDestroy, cel_elmSolverPerformance

  integer, parameter :: ik = 4, rk = 8
  integer(ik) :: elm_n = 0, elm_low = 1, elm_row = 0, elm_nnz = 0
  integer(ik), allocatable :: elm_rows(:), elm_cols(:), elm_diag(:)
  real(rk), allocatable :: elm_vals(:), elm_x(:)
  type(c_ptr) :: elm_mat = c_null_ptr
  type(cel_pcg_stats), save :: elm_stats

contains

  subroutine cel_elmSolverInitialization(itotSize, ibs, ibl, nnzero, comm)
    integer(ik), intent(in) :: itotSize, ibs, ibl, nnzero(2)
    integer(4), intent(in) :: comm
    integer(c_int) :: rc
    rc = cel_gpu_init(0_c_int)
    elm_n = ibs
    elm_low = ibl
    elm_nnz = 0
    allocate(elm_rows((nnzero(1) + nnzero(2)) * ibs), elm_cols((nnzero(1) + nnzero(2)) * ibs))
    allocate(elm_vals((nnzero(1) + nnzero(2)) * ibs), elm_diag(ibs))
    elm_diag = 0
    if (itotSize /= ibs .or. comm < -huge(comm)) then
      ! one rank owns all rows here
      write(0, '(A)') 'cel_elmSolverInitialization: one GPU solves the whole system'
    end if
  end subroutine

  subroutine cel_elmSolverDatarowStart(row)
    integer(ik), intent(in) :: row
    elm_row = row
  end subroutine

  subroutine cel_elmSolverDataInsertion(col, value)
    integer, intent(in) :: col
    real(rk), intent(in) :: value
    elm_nnz = elm_nnz + 1
    elm_rows(elm_nnz) = elm_row
    elm_cols(elm_nnz) = col
    elm_vals(elm_nnz) = value
    if (col == elm_row) elm_diag(elm_row - elm_low + 1) = elm_nnz
  end subroutine

  subroutine cel_elmSolverDataReinsertion(col, value)
    integer(ik), intent(in) :: col
    real(rk), intent(in) :: value
    if (col == elm_row) then
      elm_vals(elm_diag(elm_row - elm_low + 1)) = value
    else
      write(0, '(A)') 'cel_elmSolverDataReinsertion: only the diagonal can be reinserted'
    end if
  end subroutine

  subroutine cel_elmSolverDatarowCompleted()
  end subroutine

  ! Sorts the inserted entries by (row, column), builds the device matrix,
  ! captures the diagonal, scales rows by 1/sqrt|d| and sets the Jacobi
  ! vector sqrt|d|/d (cel_elmsolver_module.f90:372-419).
  subroutine cel_elmSolverMatrixCompleted()
    integer(ik), allocatable :: order(:)
    integer(c_int) :: rc
    integer(ik) :: k
    allocate(order(elm_nnz))
    call sort_row_col(order)
    elm_rows(1:elm_nnz) = elm_rows(order)
    elm_cols(1:elm_nnz) = elm_cols(order)
    elm_vals(1:elm_nnz) = elm_vals(order)
    do k = 1, elm_nnz
      elm_rows(k) = elm_rows(k) - elm_low + 1
      elm_cols(k) = elm_cols(k) - elm_low + 1
    end do
    rc = cel_gpu_mat_create_coo(elm_mat, elm_n, elm_n, elm_nnz, elm_rows, elm_cols, elm_vals, &
                                1_c_int, 0_c_int, 0_c_int)
    if (rc < 0) write(0, '(A,A)') 'cel_elmSolverMatrixCompleted: ', cel_gpu_error_message()
    rc = cel_gpu_mat_set_jacobi(elm_mat, 1_c_int)
    if (rc < 0) write(0, '(A,A)') 'cel_elmSolverMatrixCompleted: ', cel_gpu_error_message()
    rc = cel_gpu_mat_scale_diagonal(elm_mat)
    allocate(elm_x(elm_n))
  end subroutine

  subroutine cel_elmSolverDoIt(bmat, iierr, niter, ftol)
    real(rk), intent(inout) :: bmat(:)
    integer(ik), intent(out) :: iierr, niter
    real(rk), intent(out) :: ftol
    type(cel_pcg_params) :: p
    integer(c_int) :: rc
    real(rk), allocatable :: b(:)
    call cel_gpu_pcg_params_default(p)
    p%tol = 0.1e-07_rk            ! the API's eps (cel_elmsolver_module.f90:141)
    p%iter_max = 5000             ! :140
    p%x0_mode = CEL_X0_ONE        ! cel_cgalg_module.f90:209
    allocate(b(elm_n))
    b = bmat(1:elm_n)
    rc = cel_gpu_vec_scale_diagonal(elm_mat, b, CEL_MEM_HOST)
    rc = cel_gpu_pcg_solve(elm_mat, b, elm_x, p, elm_stats, c_null_ptr)
    iierr = merge(0_ik, int(rc, ik), rc == CEL_OK)
    niter = elm_stats%iters
    ftol = elm_stats%final_norm
    bmat(1:elm_n) = elm_x
  end subroutine

  subroutine cel_elmSolverReDoIt(bmat, iierr, niter, ftol)
    real(rk), intent(inout) :: bmat(:)
    integer(ik), intent(out) :: iierr, niter
    real(rk), intent(out) :: ftol
    call cel_elmSolverDoIt(bmat, iierr, niter, ftol)
  end subroutine

  subroutine cel_elmSolverCollectSolution(bmat)
    real(rk), allocatable, intent(inout) :: bmat(:)
    if (.not. allocated(bmat)) allocate(bmat(elm_n))
    bmat(1:elm_n) = elm_x
  end subroutine

  subroutine cel_elmSolverDestroy()
    integer(c_int) :: rc
    if (c_associated(elm_mat)) rc = cel_gpu_mat_destroy(elm_mat)
    elm_mat = c_null_ptr
    if (allocated(elm_rows)) deallocate(elm_rows, elm_cols, elm_vals, elm_diag)
    if (allocated(elm_x)) deallocate(elm_x)
  end subroutine

  subroutine cel_elmSolverPerformance()
    write(*, '(A,I0,A,ES13.6,A,F12.3,A)') 'iterations: ', elm_stats%iters, '; |r|: ', &
      elm_stats%final_norm, '; solve: ', elm_stats%solve_ms, ' ms'
  end subroutine

  ! stable order of the inserted entries by (row, column)
  subroutine sort_row_col(order)
    integer(ik), intent(out) :: order(:)
    integer(ik) :: i, j, t
    do i = 1, elm_nnz
      order(i) = i
    end do
    ! insertion sort per row run; rows arrive mostly in order
    do i = 2, elm_nnz
      t = order(i)
      j = i - 1
      do while (j >= 1)
        if (.not. before(t, order(j))) exit
        order(j + 1) = order(j)
        j = j - 1
      end do
      order(j + 1) = t
    end do
  end subroutine

  logical function before(a, b)
    integer(ik), intent(in) :: a, b
    before = elm_rows(a) < elm_rows(b) .or. &
             (elm_rows(a) == elm_rows(b) .and. elm_cols(a) < elm_cols(b))
  end function

end module cel_gpu_elmsolver
