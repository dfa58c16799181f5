! ref_mm_bind.f90 -- C entry points into the reference's own Matrix Market
! reader and COO sort, compiled from the reference sources into oracle/_ref.
!
! TEST INFRASTRUCTURE ONLY: the results that tests/ compare the oracle's
! restatement and the GPU library against.  Nothing here is the product.
!
! ref_mm_info / ref_mm_read make the calls that
! cel_sp_mat_read_matrix_market_file makes
! (src/fspmt/cel_sp_mat_generate_module.f90:219-273): mminfo, new, then
! mmread into the matrix's rows/cols/values with one-element integer and
! complex buffers.  ref_coo_sort runs cel_sp_mat_format_sparse_31_sort
! (src/fspmt/cel_sp_mat_module.f90:518-585) on a matrix built from the
! caller's arrays.  The reference routines STOP on a malformed file; callers
! run such files in a child process.
module ref_mm_bind
  use, intrinsic :: iso_c_binding
  use cel_types_module
  use cel_mmio_module
  use cel_sp_mat_module
  implicit none
contains

  subroutine open_path(path, plen, iunit)
    character(kind=c_char), dimension(*), intent(in) :: path
    integer(c_int), value :: plen
    integer, intent(out) :: iunit
    character(len=plen) :: name
    integer :: i
    do i = 1, plen
      name(i:i) = path(i)
    end do
    open(newunit=iunit, file=name, access='sequential', action='read')
  end subroutine

  integer(c_int) function ref_mm_info(path, plen, nrows, ncols, nnz) bind(c, name="ref_mm_info")
    character(kind=c_char), dimension(*), intent(in) :: path
    integer(c_int), value :: plen
    integer(c_int), intent(out) :: nrows, ncols, nnz
    character(len=7) :: field
    character(len=19) :: symm
    character(len=10) :: rep
    integer :: iunit
    integer(kind=ik) :: u, r, c, z
    call open_path(path, plen, iunit)
    u = iunit
    call mminfo(u, rep, field, symm, r, c, z)
    close(iunit)
    nrows = r
    ncols = c
    nnz = z
    ref_mm_info = 0
  end function

  integer(c_int) function ref_mm_read(path, plen, cap, rows, cols, vals) bind(c, name="ref_mm_read")
    character(kind=c_char), dimension(*), intent(in) :: path
    integer(c_int), value :: plen, cap
    integer(c_int), intent(inout) :: rows(*), cols(*)
    real(c_double), intent(inout) :: vals(*)
    type(cel_sp_mat_type) :: sp_mat
    character(len=7) :: field
    character(len=19) :: symm
    character(len=10) :: rep
    integer :: iunit
    integer(kind=ik) :: u, nrows, ncols, nnz, ival(1)
    complex :: cval(1)
    call open_path(path, plen, iunit)
    u = iunit
    call mminfo(u, rep, field, symm, nrows, ncols, nnz)
    if (nnz > cap) then
      close(iunit)
      ref_mm_read = -1
      return
    end if
    call new(sp_mat, nrows, ncols, 1_ik, 1_ik, nnz)
    sp_mat%values = 0.0_rk
    call mmread(u, rep, field, symm, nrows, ncols, nnz, huge(1_ik), &
                sp_mat%rows, sp_mat%cols, ival, sp_mat%values, cval)
    close(iunit)
    rows(1:nnz) = sp_mat%rows(1:nnz)
    cols(1:nnz) = sp_mat%cols(1:nnz)
    vals(1:nnz) = sp_mat%values(1:nnz)
    ref_mm_read = nnz
  end function

  integer(c_int) function ref_coo_sort(nrows, nnz, rows, cols, vals) bind(c, name="ref_coo_sort")
    integer(c_int), value :: nrows, nnz
    integer(c_int), intent(inout) :: rows(*), cols(*)
    real(c_double), intent(inout) :: vals(*)
    type(cel_sp_mat_type) :: sp_mat
    integer(kind=ik) :: err
    call new(sp_mat, nrows, nrows, 1_ik, 1_ik, nnz)
    sp_mat%rows(1:nnz) = rows(1:nnz)
    sp_mat%cols(1:nnz) = cols(1:nnz)
    sp_mat%values(1:nnz) = vals(1:nnz)
    call cel_sp_mat_format_sparse_31_sort(sp_mat, .false._lk, err)
    rows(1:nnz) = sp_mat%rows(1:nnz)
    cols(1:nnz) = sp_mat%cols(1:nnz)
    vals(1:nnz) = sp_mat%values(1:nnz)
    ref_coo_sort = err
  end function

end module ref_mm_bind
