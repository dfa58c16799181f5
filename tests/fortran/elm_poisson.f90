! Assembles the 7-point Poisson matrix (diag 6, off -1) of an n^3 grid row by
! row through the element-solver API, solves A x = b for b = A * x_true with
! x_true(i) = mod(i, 7) / 7, and prints the iteration count, the final norm
! and the largest error.  Mirrors the flow of the reference's
! tests/cel_cg/src/cel_elm.f90:149-234 on one GPU.
program elm_poisson
  use cel_gpu_elmsolver
  implicit none
  integer, parameter :: n = 12
  integer :: i, j, k, row, col, ierr, niter, nnz(2), nrow
  real(8), allocatable :: b(:), xt(:)
  real(8) :: ftol
  nrow = n * n * n
  nnz = [7, 0]
  call cel_elmSolverInitialization(nrow, nrow, 1, nnz, 0)
  allocate(b(nrow), xt(nrow))
  do row = 1, nrow
    xt(row) = real(mod(row, 7), 8) / 7.0d0
  end do
  b = 0.0d0
  do k = 0, n - 1
    do j = 0, n - 1
      do i = 0, n - 1
        row = 1 + i + n * (j + n * k)
        call cel_elmSolverDatarowStart(row)
        if (k > 0) call insert(row - n * n, -1.0d0)
        if (j > 0) call insert(row - n, -1.0d0)
        if (i > 0) call insert(row - 1, -1.0d0)
        call insert(row, 6.0d0)
        if (i < n - 1) call insert(row + 1, -1.0d0)
        if (j < n - 1) call insert(row + n, -1.0d0)
        if (k < n - 1) call insert(row + n * n, -1.0d0)
        call cel_elmSolverDatarowCompleted()
      end do
    end do
  end do
  call cel_elmSolverMatrixCompleted()
  call cel_elmSolverDoIt(b, ierr, niter, ftol)
  write(*, '(A,I0,A,I0,A,ES24.16,A,ES12.4)') 'status ', ierr, ' iterations ', niter, &
    ' norm ', ftol, ' maxerr ', maxval(abs(b - xt))
  call cel_elmSolverPerformance()
  call cel_elmSolverDestroy()
contains
  subroutine insert(c, v)
    integer, intent(in) :: c
    real(8), intent(in) :: v
    call cel_elmSolverDataInsertion(c, v)
    b(row) = b(row) + v * xt(c)
  end subroutine
end program
