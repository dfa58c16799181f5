! batch_poisson.f90 -- batched Jacobi-PCG through the Fortran bindings:
! a 10^3 and a 6x7x8 Poisson system, the second with 3 right-hand sides,
! solved together on a launch configuration fixed by hand (the vector
! kernels on a grid of their own); prints the configuration read back,
! per-lane status, iterations and the true relative residual computed with
! the library's SpMV.
program batch_poisson
  use, intrinsic :: iso_c_binding
  use cel_gpu
  implicit none
  type(c_ptr) :: A(2), B
  integer(c_int) :: nrhs(2), ierr, i, j, n(2), lane
  integer(c_int) :: cg, cg23, cch, cw
  type(c_ptr) :: bp(2), xp(2)
  real(c_double), allocatable, target :: b1(:), x1(:), b2(:), x2(:)
  real(c_double), allocatable :: ax(:)
  type(cel_pcg_params) :: p
  type(cel_pcg_stats) :: st(4), tot
  real(c_double) :: rel, worst

  ierr = cel_gpu_init(0_c_int)
  if (ierr /= CEL_OK) stop 1
  ierr = cel_gpu_mat_gen_poisson7(A(1), 10_c_int, 10_c_int, 10_c_int)
  ierr = ior(ierr, cel_gpu_mat_gen_poisson7(A(2), 6_c_int, 7_c_int, 8_c_int))
  ierr = ior(ierr, cel_gpu_mat_set_jacobi(A(1), 0_c_int))
  ierr = ior(ierr, cel_gpu_mat_set_jacobi(A(2), 0_c_int))
  if (ierr /= CEL_OK) stop 2
  n = [1000, 336]
  nrhs = [1, 3]
  allocate(b1(n(1)), x1(n(1)), b2(n(2) * 3), x2(n(2) * 3), ax(n(1)))
  do i = 1, n(1)
    b1(i) = 1.0d0 + mod(i, 7) * 0.25d0
  end do
  do j = 0, 2
    do i = 1, n(2)
      b2(j * n(2) + i) = dble(mod(i * (j + 3), 11)) - 5.0d0
    end do
  end do
  ierr = cel_gpu_batch_create(B, 2_c_int, A, nrhs)
  if (ierr /= CEL_OK) stop 3
  ierr = cel_gpu_batch_set_config(B, 16_c_int, 8_c_int, 4_c_int, 1_c_int)
  if (ierr /= CEL_OK) stop 5
  ierr = cel_gpu_batch_get_config(B, cg, cg23, cch, cw)
  if (ierr /= CEL_OK) stop 6
  write (*, '(a,4i6)') 'config ', cg, cg23, cch, cw
  call cel_gpu_pcg_params_default(p)
  p%tol = 1.0d-10
  p%tol_mode = CEL_TOL_REL_B
  p%x0_mode = CEL_X0_ZERO
  p%alpha_stop = 0
  p%iter_max = 2000
  p%memspace = CEL_MEM_HOST
  bp = [c_loc(b1), c_loc(b2)]
  xp = [c_loc(x1), c_loc(x2)]
  ierr = cel_gpu_batch_solve(B, bp, xp, p, st, tot)
  if (ierr < 0) stop 4
  worst = 0.0d0
  ierr = cel_gpu_spmv(A(1), CEL_FMT_CSR, x1, ax, 0_c_int, CEL_MEM_HOST)
  rel = sqrt(sum((b1 - ax(1:n(1)))**2)) / sqrt(sum(b1**2))
  worst = max(worst, rel)
  write (*, '(a,i2,a,i3,a,i5,a,es10.3)') 'lane ', 1, ' status ', st(1)%status, ' iterations ', &
      st(1)%iters, ' relres ', rel
  do j = 0, 2
    lane = 2 + j
    ierr = cel_gpu_spmv(A(2), CEL_FMT_CSR, x2(j * n(2) + 1:), ax, 0_c_int, CEL_MEM_HOST)
    rel = sqrt(sum((b2(j * n(2) + 1:(j + 1) * n(2)) - ax(1:n(2)))**2)) / &
          sqrt(sum(b2(j * n(2) + 1:(j + 1) * n(2))**2))
    worst = max(worst, rel)
    write (*, '(a,i2,a,i3,a,i5,a,es10.3)') 'lane ', lane, ' status ', st(lane)%status, &
        ' iterations ', st(lane)%iters, ' relres ', rel
  end do
  write (*, '(a,i3,a,es10.3)') 'batch status ', tot%status, ' worst relres ', worst
  ierr = cel_gpu_batch_destroy(B)
  ierr = cel_gpu_mat_destroy(A(1))
  ierr = cel_gpu_mat_destroy(A(2))
  ierr = cel_gpu_finalize()
end program batch_poisson
