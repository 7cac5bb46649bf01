! Drop-in replacements of the reference's linear-solver modules.  Module and
! procedure names, dummy arguments and their meaning are those of
!   solv_res_m     spray-simulation/solv_res.f90:1-31
!   solv_precond_m spray-simulation/solv_precond.f90:1-31
!   solv_direct_m  spray-simulation/solv_direct.f90:1-44
!   solv_bicg_m    spray-simulation/solv_bicg.f90:1-122
!   solv_cg_m      spray-simulation/solv_cg.f90:1-88
! and they read the same module data (solv_data_m: pl, ng, v_lst).  The work
! runs on the GPU through the fvk C ABI on the context grd_solv created; the
! current matrix is uploaded on every call, since the caller may have changed
! pl or ng in between.

module fvk_solv_util_m
  use solv_data_m
  use fvk_c
  implicit none
contains

  subroutine fvk_need_ctx(where)
    character(len=*),intent(in)::where
    if(.not.c_associated(fvk_ctx))then
      write(*,*)'fvk: ',where,' called before grd_solv created the device context'
      stop 1
    end if
  end subroutine

  ! pl, ng%p, ng%n of solv_data_m -> device
  subroutine fvk_upload_matrix(where)
    character(len=*),intent(in)::where
    real(c_double),allocatable::gp(:),gn(:)
    integer::ifce
    call fvk_need_ctx(where)
    allocate(gp(size(ng)),gn(size(ng)))
    do ifce=1,size(ng)
      gp(ifce)=ng(ifce)%p
      gn(ifce)=ng(ifce)%n
    end do
    call fvk_check(fvk_set_matrix(fvk_ctx,pl,gp,gn),where)
  end subroutine

  ! phi(ip)%vc(1:n_vc), ip = 1..size(pl)  <->  contiguous x(n_vc, n)
  subroutine fvk_pack_phi(phi,n_vc,x)
    type(phi_t),dimension(:),intent(in)::phi
    integer,intent(in)::n_vc
    real(c_double),intent(out)::x(n_vc,size(pl))
    integer::ip
    do ip=1,size(pl)
      x(:,ip)=phi(ip)%vc(1:n_vc)
    end do
  end subroutine

  subroutine fvk_unpack_phi(x,n_vc,phi)
    integer,intent(in)::n_vc
    real(c_double),intent(in)::x(n_vc,size(pl))
    type(phi_t),dimension(:),intent(inout)::phi
    integer::ip
    do ip=1,size(pl)
      phi(ip)%vc(1:n_vc)=x(:,ip)
    end do
  end subroutine
end module


module solv_res_m
  use fvk_solv_util_m
  implicit none
contains
  subroutine solv_res(phi,res0,res0_n,not_r0)
    type(phi_t),dimension(:),intent(inout)::phi
    real(wp),dimension(:,:),intent(inout)::res0
    real(wp),dimension(:),intent(inout)::res0_n
    logical,intent(in)::not_r0
    real(c_double),allocatable::x(:,:),r(:,:)
    integer::n_vc
    n_vc=size(res0_n)
    call fvk_upload_matrix('solv_res')
    allocate(x(n_vc,size(pl)),r(size(pl),n_vc))
    if(.not.not_r0)then
      call fvk_pack_phi(phi,n_vc,x)
    else
      x=0
    end if
    r=res0(1:size(pl),1:n_vc)
    call fvk_check(fvk_solv_res(fvk_ctx,n_vc,x,r,res0_n,merge(1,0,not_r0)),'solv_res')
    res0(1:size(pl),1:n_vc)=r
  end subroutine
end module


module solv_precond_m
  use fvk_solv_util_m
  implicit none
contains
  subroutine solv_precond(imat,c,c_r)
    integer,intent(in)::imat
    real(wp),dimension(:),intent(inout)::c,c_r
    real(c_double),allocatable::cc(:),cr(:)
    call fvk_upload_matrix('solv_precond')
    allocate(cc(size(pl)),cr(size(pl)))
    call fvk_check(fvk_solv_precond(fvk_ctx,imat,cc,cr),'solv_precond')
    c(1:size(pl))=cc
    c_r(1:size(pl))=cr
  end subroutine
end module


module solv_direct_m
  use fvk_solv_util_m
  implicit none
contains
  subroutine solv_direct(imat,c,c_r,x,b)
    integer,intent(in)::imat
    real(wp),dimension(:),intent(in)::c,c_r,b
    real(wp),dimension(:),intent(inout)::x
    real(c_double),allocatable::cc(:),cr(:),bb(:),xx(:)
    integer::n
    n=size(pl)
    call fvk_upload_matrix('solv_direct')
    allocate(cc(n),cr(n),bb(n),xx(n))
    cc=c(1:n); cr=c_r(1:n); bb=b(1:n)
    call fvk_check(fvk_solv_direct(fvk_ctx,imat,cc,cr,bb,xx),'solv_direct')
    x(1:n)=xx
  end subroutine
end module


module solv_bicg_m
  use solv_res_m
  use solv_precond_m
  use solv_direct_m
  implicit none
contains
  subroutine solv_bicg(imat,reduc,n_iter,phi,res0,res0_n,not_r0)
    integer,intent(in)::imat,n_iter
    real(wp),intent(in)::reduc
    logical,intent(in)::not_r0
    type(phi_t),dimension(:),intent(inout)::phi
    real(wp),dimension(:,:),intent(inout)::res0
    real(wp),dimension(:),intent(inout)::res0_n
    real(c_double),allocatable::x(:,:),r(:,:)
    integer::n_vc
    n_vc=size(res0_n)
    call fvk_upload_matrix('solv_bicg')
    allocate(x(n_vc,size(pl)),r(size(pl),n_vc))
    call fvk_pack_phi(phi,n_vc,x)
    r=res0(1:size(pl),1:n_vc)
    call fvk_check(fvk_solv_bicg(fvk_ctx,imat,reduc,n_iter,n_vc,x,r,res0_n,merge(1,0,not_r0), &
      c_null_ptr),'solv_bicg')
    call fvk_unpack_phi(x,n_vc,phi)
    res0(1:size(pl),1:n_vc)=r
  end subroutine
end module


module solv_cg_m
  use solv_res_m
  use solv_precond_m
  use solv_direct_m
  implicit none
contains
  subroutine solv_cg(imat,reduc,n_iter,phi,res0,res0_n,not_r0)
    integer,intent(in)::imat,n_iter
    real(wp),intent(in)::reduc
    logical,intent(in)::not_r0
    type(phi_t),dimension(:),intent(inout)::phi
    real(wp),dimension(:,:),intent(inout)::res0
    real(wp),dimension(:),intent(inout)::res0_n
    real(c_double),allocatable::x(:,:),r(:,:)
    integer::n_vc
    n_vc=size(res0_n)
    call fvk_upload_matrix('solv_cg')
    allocate(x(n_vc,size(pl)),r(size(pl),n_vc))
    call fvk_pack_phi(phi,n_vc,x)
    r=res0(1:size(pl),1:n_vc)
    call fvk_check(fvk_solv_cg(fvk_ctx,imat,reduc,n_iter,n_vc,x,r,res0_n,merge(1,0,not_r0), &
      c_null_ptr),'solv_cg')
    call fvk_unpack_phi(x,n_vc,phi)
    res0(1:size(pl),1:n_vc)=r
  end subroutine
end module
