! Drop-in replacement of flowlab's solv_m (flowlab/src/solv_m.f90; the same
! source is cfdpack's and pcns's): lin_eqns_solver keeps its interface
! (solv_m.f90:10-30) and runs cg_solver / bicg_solver with the reference's
! preconditioner and direct_solve on the GPU through the fvk C ABI
! (fvk_fl_lin_eqns_solver).  The cell matrix pattern goes to the device on
! the first call and again only when a call brings another one; each call
! sends the matrix values, phi and the right-hand side and returns phi, the
! initial residual in rhs and the final L1 norm, as the reference does.
module solv_m
  use constants_m
  use iso_c_binding
  implicit none
  private
  public :: lin_eqns_solver

  type(c_ptr), save :: fl_ctx = c_null_ptr

  interface
    integer(c_int) function fvk_fl_csr_create(n, ia, ja, device, out) bind(C, name='fvk_fl_csr_create')
      import :: c_int, c_ptr
      integer(c_int), value :: n, device
      integer(c_int), intent(in) :: ia(*), ja(*)
      type(c_ptr) :: out
    end function
    integer(c_int) function fvk_fl_csr_matches(ctx, n, ia, ja) bind(C, name='fvk_fl_csr_matches')
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: n
      integer(c_int), intent(in) :: ia(*), ja(*)
    end function
    subroutine fvk_fl_ctx_destroy(ctx) bind(C, name='fvk_fl_ctx_destroy')
      import :: c_ptr
      type(c_ptr), value :: ctx
    end subroutine
    integer(c_int) function fvk_fl_lin_eqns_solver(ctx, a, phi, rhs, norm, n_iter, reduc, isol) &
        bind(C, name='fvk_fl_lin_eqns_solver')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(in) :: a(*)
      real(c_double), intent(inout) :: phi(*), rhs(*)
      real(c_double), intent(out) :: norm
      integer(c_int), value :: n_iter, isol
      real(c_double), value :: reduc
    end function
    function fvk_last_error() bind(C, name='fvk_last_error')
      import :: c_ptr
      type(c_ptr) :: fvk_last_error
    end function
  end interface

contains

subroutine lin_eqns_solver(a_ij, phi, rhs, norm, ja, ia, n_iter, reduc, isol, id)
  real(rp), dimension(:), intent(in) :: a_ij
  real(rp), dimension(:), intent(inout) :: phi
  real(rp), dimension(:), intent(inout) :: rhs
  real(rp), intent(inout) :: norm
  real(rp), intent(in) :: reduc
  integer, dimension(:), intent(in) :: ja, ia
  integer, intent(in) :: n_iter
  integer, intent(in) :: isol, id

  integer(c_int) :: n, rc
  real(c_double) :: nrm
  real(c_double), allocatable :: p(:), r(:)

  n = size(ia) - 1
  if (c_associated(fl_ctx)) then
    if (fvk_fl_csr_matches(fl_ctx, n, ia, ja) == 0) then
      call fvk_fl_ctx_destroy(fl_ctx)
      fl_ctx = c_null_ptr
    end if
  end if
  if (.not. c_associated(fl_ctx)) then
    rc = fvk_fl_csr_create(n, ia, ja, 0_c_int, fl_ctx)
    if (rc /= 0) call fail(rc)
  end if
  ! phi and rhs may be rows of a (n_cmp, n) array: contiguous copies
  p = phi
  r = rhs
  rc = fvk_fl_lin_eqns_solver(fl_ctx, a_ij, p, r, nrm, int(n_iter, c_int), real(reduc, c_double), int(isol, c_int))
  if (rc /= 0) call fail(rc)
  phi = p
  rhs = r
  norm = nrm
end subroutine

subroutine fail(rc)
  integer(c_int), intent(in) :: rc
  character(kind=c_char), pointer :: msg(:)
  integer :: k
  call c_f_pointer(fvk_last_error(), msg, [512])
  k = 0
  do while (k < 512)
    if (msg(k + 1) == c_null_char) exit
    k = k + 1
  end do
  print "(a,i0,a,512a)", "fvk: lin_eqns_solver failed (", rc, "): ", msg(1:k)
  stop 1
end subroutine

end module
