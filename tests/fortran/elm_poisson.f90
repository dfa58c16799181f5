! Assembles a 7-point operator on an n^3 grid row by row through the
! element-solver API -- diagonal 6, off-diagonal -1, plus a first-order
! upwind convection term in x that makes it nonsymmetric -- solves A x = b
! for b = A * x_true with x_true(i) = mod(i, 7) / 7 with the reference's
! default method (Jacobi-GMRES), then again with ReDoIt, and once with
! Jacobi-PCG on the symmetric part.  Prints one line per solve: status,
! iterations, the final norm and the largest error of the collected x.
! Mirrors the flow of the reference's tests/cel_cg/src/cel_cg/cel_elm.f90:149-234
! on one GPU.
program elm_poisson
  use cel_gpu_elmsolver
  implicit none
  integer, parameter :: n = 12
  integer :: ierr, niter, nnz(2), nrow, pass
  real(8), allocatable :: b(:), xt(:), x(:)
  real(8) :: ftol, conv
  nrow = n * n * n
  nnz = [7, 0]
  allocate(b(nrow), xt(nrow))
  do pass = 1, 2
    conv = merge(0.5d0, 0.0d0, pass == 1)
    call assemble(conv)
    if (pass == 2) call cel_elmSolverSetSolver(1)
    call cel_elmSolverDoIt(b, ierr, niter, ftol)
    call cel_elmSolverCollectSolution(x)
    write(*, '(A,I0,A,I0,A,ES24.16,A,ES12.4)') 'status ', ierr, ' iterations ', niter, &
      ' norm ', ftol, ' maxerr ', maxval(abs(x - xt))
    if (pass == 1) then
      call cel_elmSolverReDoIt(b, ierr, niter, ftol)
      call cel_elmSolverCollectSolution(x)
      write(*, '(A,I0,A,I0,A,ES24.16,A,ES12.4)') 'status ', ierr, ' iterations ', niter, &
        ' norm ', ftol, ' maxerr ', maxval(abs(x - xt))
    end if
    call cel_elmSolverPerformance()
    call cel_elmSolverDestroy()
  end do
contains
  subroutine assemble(c)
    real(8), intent(in) :: c
    integer :: i, j, k, row
    call cel_elmSolverInitialization(nrow, nrow, 1, nnz, 0)
    do row = 1, nrow
      xt(row) = real(mod(row, 7), 8) / 7.0d0
    end do
    b = 0.0d0
    do k = 0, n - 1
      do j = 0, n - 1
        do i = 0, n - 1
          row = 1 + i + n * (j + n * k)
          call cel_elmSolverDatarowStart(row)
          if (k > 0) call insert(row, row - n * n, -1.0d0)
          if (j > 0) call insert(row, row - n, -1.0d0)
          if (i > 0) call insert(row, row - 1, -1.0d0 - c)
          call insert(row, row, 6.0d0 + c)
          if (i < n - 1) call insert(row, row + 1, -1.0d0)
          if (j < n - 1) call insert(row, row + n, -1.0d0)
          if (k < n - 1) call insert(row, row + n * n, -1.0d0)
          call cel_elmSolverDatarowCompleted()
        end do
      end do
    end do
    call cel_elmSolverMatrixCompleted()
  end subroutine
  subroutine insert(row, c, v)
    integer, intent(in) :: row, c
    real(8), intent(in) :: v
    call cel_elmSolverDataInsertion(c, v)
    b(row) = b(row) + v * xt(c)
  end subroutine
end program
