! Drop-in replacement of the reference's grd_solv_m
! (spray-simulation/grd_solv.f90:1-25): it sets up the same solver topology
! in solv_data_m (v_lst aliases the cell face lists, ng holds the face
! owner/neighbour indices, pl is allocated) and in addition uploads topology
! and geometry of grid igrd into an fvk device context (fvk_c::fvk_ctx) that
! the drop-in solver and assembly modules use.
module grd_solv_m
  use grd_data_m
  use solv_data_m
  use fvk_c
  implicit none
contains

  subroutine grd_solv(igrd)
    integer,intent(in)::igrd
    integer::ivol,ifce
    type(grd_t),pointer::g

    g=>grd(igrd)
    allocate(v_lst(g%n_vol),pl(g%n_vol),ng(g%n_fce))
    do ivol=1,g%n_vol
      v_lst(ivol)%fce_lst=>g%vol(ivol)%fce_lst
      v_lst(ivol)%ng_lst=>g%vol(ivol)%ng_lst
    end do
    do ifce=1,g%n_fce
      ng(ifce)%ip=g%fce(ifce)%ip
      ng(ifce)%in=g%fce(ifce)%in
    end do
    call fvk_grd_context(igrd)
  end subroutine

  ! grd(igrd) -> fvk_ctx_create_arrays (layouts of include/fvk.h)
  subroutine fvk_grd_context(igrd)
    integer,intent(in)::igrd
    type(grd_t),pointer::g
    integer(c_int),allocatable::ip(:),in(:),ib(:),bty(:),cptr(:),lst(:)
    real(c_double),allocatable::vx(:,:),vv(:),fx(:,:),fn(:,:),fdp(:,:),fdn(:,:),fs(:),fl(:),fw(:)
    integer::ivol,ifce,k,nd,ns

    g=>grd(igrd)
    nd=n_dx
    allocate(ip(g%n_fce),in(g%n_fce),ib(g%n_fce),bty(g%n_fce),cptr(g%n_vol+1))
    allocate(vx(nd,g%n_vol),vv(g%n_vol),fx(nd,g%n_fce),fn(nd,g%n_fce),fdp(nd,g%n_fce),fdn(nd,g%n_fce))
    allocate(fs(g%n_fce),fl(g%n_fce),fw(g%n_fce))
    cptr(1)=0
    do ivol=1,g%n_vol
      cptr(ivol+1)=cptr(ivol)+g%vol(ivol)%n_fce
      vx(:,ivol)=g%vol(ivol)%x
      vv(ivol)=g%vol(ivol)%v
    end do
    ns=cptr(g%n_vol+1)
    allocate(lst(ns))
    do ivol=1,g%n_vol
      do k=1,g%vol(ivol)%n_fce
        lst(cptr(ivol)+k)=g%vol(ivol)%fce_lst(k)
      end do
    end do
    do ifce=1,g%n_fce
      ip(ifce)=g%fce(ifce)%ip
      in(ifce)=g%fce(ifce)%in
      ib(ifce)=g%fce(ifce)%ib
      bty(ifce)=g%fce(ifce)%bnd_ty
      fx(:,ifce)=g%fce(ifce)%x
      fn(:,ifce)=g%fce(ifce)%n
      fdp(:,ifce)=g%fce(ifce)%dx_p
      fdn(:,ifce)=0
      if(g%fce(ifce)%in>0)fdn(:,ifce)=g%fce(ifce)%dx_n
      fs(ifce)=g%fce(ifce)%s
      fl(ifce)=g%fce(ifce)%l_pn
      fw(ifce)=g%fce(ifce)%w_fp
    end do
    call fvk_check(fvk_ctx_create_arrays(nd,g%n_vol,g%n_fce,g%n_bf,ip,in,ib,bty,cptr,lst, &
      vx,vv,fx,fn,fdp,fdn,fs,fl,fw,fvk_device,0,fvk_ctx),'grd_solv')
  end subroutine
end module
