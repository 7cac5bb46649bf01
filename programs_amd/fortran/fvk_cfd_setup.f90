! Drop-in replacement of the reference's cfd_setup_m
! (spray-simulation/cfd_setup.f90:1-922) for the single-phase steady flow:
! cfd_setup_eqn (momentum assembly + BiCGStab), cfd_setup_pp (pressure
! correction, velocity and flux corrections, SIMPLER or relaxed pressure
! update), solv_grad (Green-Gauss gradients), test_grad and zero_spd keep
! their names and arguments.  The assembly and the solves run on the GPU
! (fvk flow on the grd_solv context); the grd structure, pl/ng and the
! residual bookkeeping of cfd_data_m are updated on the host as the reference
! leaves them, and the same 'reduc:' / 'redpp:' / 'red_p:' lines are printed.
!
! The reference hard-codes its solver settings as locals
! (cfd_setup.f90:61,528,580-581,662,742); here they come from the
! environment, defaulting to the same values:
!   FVK_SIMPLER (1)  FVK_IMAT_U (2)  FVK_IMAT_P (1)  FVK_N_BICG (5)
!   FVK_N_CG (20)    FVK_REDUC_U (0.1)  FVK_REDUC_P (0.1)
! Branches outside the plain flow (spray phase, turbulence, transient
! terms, axisymmetric 2D) stop with a message.
module cfd_setup_m
  use solv_bicg_m
  use solv_cg_m
  use cfd_define_m
  use fvk_c
  implicit none
  real(wp),save,private::pp_res_n=0.
contains

  subroutine zero_spd(igrd)
    integer,intent(in)::igrd
    integer::ifce
    do ifce=1,size(ng)
      grd(igrd)%fce(ifce)%spd=zero
    end do
  end subroutine

  subroutine fvk_plain_only(iph,ieq,isc,where)
    integer,intent(in)::iph,ieq,isc
    character(len=*),intent(in)::where
    if(iph/=1.or.ieq/=1.or.isc/=1.or.n_tl/=1.or.turb.or.spr.or.(n_dx==2.and.cyl))then
      write(*,*)'fvk: ',where,': only the single-phase steady flow (iph=1, ieq=1) runs on the GPU'
      stop 1
    end if
  end subroutine

  ! solver configuration of the flow from ieq and the environment
  subroutine fvk_flow_config(igrd,iph,ieq,urf_p,cfg)
    integer,intent(in)::igrd,iph,ieq
    real(wp),intent(in)::urf_p
    type(fvk_flow_cfg_t),intent(out)::cfg
    type(ieq_t),pointer::ieq_p
    ieq_p=>grd(igrd)%iph(iph)%ieq(ieq)
    cfg%simpler=fvk_env_int('FVK_SIMPLER',1)
    cfg%imat_u=fvk_env_int('FVK_IMAT_U',2)
    cfg%imat_p=fvk_env_int('FVK_IMAT_P',1)
    cfg%n_bicg=fvk_env_int('FVK_N_BICG',5)
    cfg%n_cg=fvk_env_int('FVK_N_CG',20)
    cfg%reduc_u=fvk_env_real('FVK_REDUC_U',0.1_c_double)
    cfg%reduc_p=fvk_env_real('FVK_REDUC_P',0.1_c_double)
    cfg%n_grad=n_grad
    cfg%c_sch=ieq_p%c_sch
    cfg%urf_u=ieq_p%urf
    cfg%urf_p=urf_p
    cfg%c_bl=ieq_p%c_bl
    ! uniform properties: density prp(2), viscosity prp(3) (cfd_solve.f90:64-66)
    cfg%rho=grd(igrd)%vol(1)%phs(iph)%prp(2)%rho(1)
    cfg%mu=grd(igrd)%vol(1)%phs(iph)%prp(3)%rho(1)
    cfg%p_solver=fvk_env_int('FVK_P_SOLVER',0)
    cfg%p_tol=fvk_env_real('FVK_P_TOL',1.e-8_c_double)
  end subroutine

  subroutine fvk_flow_ensure(igrd,iph,ieq,urf_p)
    integer,intent(in)::igrd,iph,ieq
    real(wp),intent(in)::urf_p
    type(fvk_flow_cfg_t)::cfg
    if(.not.c_associated(fvk_ctx))then
      write(*,*)'fvk: cfd_setup called before grd_solv created the device context'
      stop 1
    end if
    call fvk_flow_config(igrd,iph,ieq,urf_p,cfg)
    if(.not.c_associated(fvk_flow))then
      call fvk_check(fvk_flow_create(fvk_ctx,cfg,fvk_flow),'fvk_flow_create')
    else
      call fvk_check(fvk_flow_configure(fvk_flow,cfg),'fvk_flow_configure')
    end if
  end subroutine

  ! grd structure -> device flow state
  subroutine fvk_push_state(igrd,iph,ieq,isc)
    integer,intent(in)::igrd,iph,ieq,isc
    type(grd_t),pointer::g
    real(c_double),allocatable::uc(:,:),ub(:,:),p(:),dp(:,:),sp(:)
    integer::ivol,ifce,ib
    g=>grd(igrd)
    allocate(uc(n_dx,g%n_vol),ub(n_dx,max(g%n_bf,1)),p(g%n_vol),dp(n_dx,g%n_vol),sp(g%n_fce))
    ub=0
    do ivol=1,g%n_vol
      uc(:,ivol)=g%vol(ivol)%phs(iph)%eqn(ieq)%phi(isc,:,1)
      p(ivol)=g%vol(ivol)%p(1)
      dp(:,ivol)=g%vol(ivol)%dpdx
    end do
    do ifce=1,g%n_fce
      sp(ifce)=g%fce(ifce)%spd
      if(g%fce(ifce)%in==0)then
        ib=g%fce(ifce)%ib-g%n_vol
        ub(:,ib)=g%fce(ifce)%phs(iph)%eqn(ieq)%phi(isc,:,1)
      end if
    end do
    call fvk_check(fvk_flow_set_state(fvk_flow,uc,ub,p,dp,sp),'fvk_flow_set_state')
  end subroutine

  ! device flow state -> grd structure, pl and ng
  subroutine fvk_pull_state(igrd,iph,ieq,isc)
    integer,intent(in)::igrd,iph,ieq,isc
    type(grd_t),pointer::g
    real(c_double),allocatable::uc(:,:),ub(:,:),p(:),dp(:,:),sp(:),cs(:),fd(:),gp(:),gn(:)
    integer::ivol,ifce,ib
    g=>grd(igrd)
    allocate(uc(n_dx,g%n_vol),ub(n_dx,max(g%n_bf,1)),p(g%n_vol),dp(n_dx,g%n_vol),sp(g%n_fce))
    allocate(cs(g%n_vol),fd(g%n_fce),gp(g%n_fce),gn(g%n_fce))
    call fvk_check(fvk_flow_get_state(fvk_flow,uc,ub,p,dp,sp,cs,fd),'fvk_flow_get_state')
    call fvk_check(fvk_flow_get_matrix(fvk_flow,pl,gp,gn),'fvk_flow_get_matrix')
    do ivol=1,g%n_vol
      g%vol(ivol)%phs(iph)%eqn(ieq)%phi(isc,:,1)=uc(:,ivol)
      g%vol(ivol)%p(1)=p(ivol)
      g%vol(ivol)%dpdx=dp(:,ivol)
      g%vol(ivol)%cs=cs(ivol)
      g%vol(ivol)%frc_vol=g%vol(ivol)%phs(iph)%prp(1)%rho(1)*g%vol(ivol)%v
    end do
    do ifce=1,g%n_fce
      g%fce(ifce)%spd=sp(ifce)
      g%fce(ifce)%frc_dns=fd(ifce)
      ng(ifce)%p=gp(ifce)
      ng(ifce)%n=gn(ifce)
      if(g%fce(ifce)%in==0)then
        ib=g%fce(ifce)%ib-g%n_vol
        g%fce(ifce)%phs(iph)%eqn(ieq)%phi(isc,:,1)=ub(:,ib)
      end if
    end do
  end subroutine

  ! momentum equation (cfd_setup.f90:22-544)
  subroutine cfd_setup_eqn(igrd,iph,ieq,isc,it,vof)
    integer,intent(in)::igrd,iph,ieq,isc,it
    logical,intent(in)::vof
    type(ieq_t),pointer::ieq_p
    type(fvk_outer_hist_t)::h
    real(c_double)::cin,cout
    real(wp)::mx_res0_n

    ieq_p=>grd(igrd)%iph(iph)%ieq(ieq)
    ieq_p%c_in(isc)=zero
    ieq_p%c_out(isc)=zero
    if(.not.ieq_p%solv)return
    call fvk_plain_only(iph,ieq,isc,'cfd_setup_eqn')
    call fvk_flow_ensure(igrd,iph,ieq,0.1_wp)
    call fvk_push_state(igrd,iph,ieq,isc)
    cin=c_in; cout=c_out
    call fvk_check(fvk_flow_scalars(fvk_flow,1,cin,cout),'fvk_flow_scalars')
    call fvk_check(fvk_flow_setup_eqn(fvk_flow,it,h),'cfd_setup_eqn')
    call fvk_pull_state(igrd,iph,ieq,isc)
    if(vof)then
      call fvk_check(fvk_flow_scalars(fvk_flow,0,cin,cout),'fvk_flow_scalars')
      c_in=cin; c_out=cout
    end if

    ! residual normalisation against the running maximum (:530-543)
    if(it==1)ieq_p%res_n(isc)=zero
    mx_res0_n=maxval(h%res0_n_u(1:ieq_p%n_vc))
    if(mx_res0_n>small)then
      ieq_p%res_n(isc)=max(ieq_p%res_n(isc),mx_res0_n)
      mx_res0_n=mx_res0_n/ieq_p%res_n(isc)
      mx_res_n=max(mx_res_n,mx_res0_n)
    end if
    print*,'reduc:',iph,ieq,isc,mx_res0_n
    if(abs(mx_res0_n)>large)stop
  end subroutine

  ! pressure correction (cfd_setup.f90:549-760)
  subroutine cfd_setup_pp(igrd,iph,ieq,it,n_pp,urf)
    integer,intent(in)::igrd,iph,ieq,it,n_pp
    real(wp),intent(in)::urf
    type(fvk_outer_hist_t)::h
    real(c_double)::cin,cout
    real(c_double),allocatable::gp(:),gn(:)
    real(wp)::mx_res0_n
    integer::ifce

    call fvk_plain_only(iph,ieq,1,'cfd_setup_pp')
    if(n_pp/=1)then
      write(*,*)'fvk: cfd_setup_pp runs one pressure-correction pass (n_pp=1)'
      stop 1
    end if
    call fvk_flow_ensure(igrd,iph,ieq,urf)
    ! the momentum diagonal the caller holds in pl feeds 1/a_P (:583)
    allocate(gp(size(ng)),gn(size(ng)))
    do ifce=1,size(ng)
      gp(ifce)=ng(ifce)%p
      gn(ifce)=ng(ifce)%n
    end do
    call fvk_check(fvk_set_matrix(fvk_ctx,pl,gp,gn),'cfd_setup_pp')
    call fvk_push_state(igrd,iph,ieq,1)
    cin=c_in; cout=c_out
    call fvk_check(fvk_flow_scalars(fvk_flow,1,cin,cout),'fvk_flow_scalars')
    call fvk_check(fvk_flow_setup_pp(fvk_flow,it,h),'cfd_setup_pp')
    call fvk_pull_state(igrd,iph,ieq,1)

    mx_res0_n=h%res0_n_pp+small
    if(it==1)pp_res_n=zero
    pp_res_n=max(pp_res_n,mx_res0_n)
    mx_res0_n=mx_res0_n/pp_res_n
    mx_res_n=max(mx_res_n,mx_res0_n)
    print*,'redpp:',iph,ieq,1,mx_res0_n,h%redpp_sum
    if(fvk_env_int('FVK_SIMPLER',1)/=0)then
      mx_res0_n=h%res0_n_p+small
      if(it==1)pp_res_n=zero
      pp_res_n=max(pp_res_n,mx_res0_n)
      mx_res0_n=mx_res0_n/pp_res_n
      mx_res_n=max(mx_res_n,mx_res0_n)
      print*,'red_p:',iph,ieq,0,mx_res0_n,h%red_p_sum
    end if
  end subroutine

  ! Green-Gauss gradient (cfd_setup.f90:765-866)
  subroutine solv_grad(igrd,phi,grad,n_iter,neu)
    type(phi_t),dimension(:),intent(inout)::phi
    real(wp),dimension(:,:,:),intent(inout)::grad
    integer,intent(in)::igrd,n_iter,neu
    real(c_double),allocatable::x(:,:),g(:,:,:)
    integer::nv,nb,n_vc,i

    if(n_tl>1.and.size(grad,2)>1.or.(n_dx==2.and.cyl))then
      write(*,*)'fvk: solv_grad: Courant numbers and axisymmetric terms are not on the GPU path'
      stop 1
    end if
    if(.not.c_associated(fvk_ctx))then
      write(*,*)'fvk: solv_grad called before grd_solv created the device context'
      stop 1
    end if
    nv=grd(igrd)%n_vol
    nb=grd(igrd)%n_bf
    n_vc=size(grad,2)
    allocate(x(n_vc,nv+nb),g(n_dx,n_vc,nv))
    x=0
    do i=1,nv
      x(:,i)=phi(i)%vc(1:n_vc)
    end do
    if(neu==0)then
      do i=nv+1,nv+nb
        if(associated(phi(i)%vc))x(:,i)=phi(i)%vc(1:n_vc)
      end do
    end if
    call fvk_check(fvk_solv_grad(fvk_ctx,n_vc,x,g,n_iter,neu),'solv_grad')
    do i=1,nv
      grad(i,1:n_vc,1:n_dx)=transpose(g(:,:,i))
    end do
  end subroutine

  ! convergence of solv_grad on phi = product of the coordinates (:871-921)
  subroutine test_grad(igrd)
    integer,intent(in)::igrd
    type(phi_t),dimension(grd(igrd)%n_vbf)::phi
    real(wp),dimension(grd(igrd)%n_vol,1,n_dx)::grad
    real(wp),dimension(n_dx)::err
    integer::ivol,ifce,ib
    type(grd_t),pointer::g
    g=>grd(igrd)
    do ivol=1,g%n_vol
      allocate(phi(ivol)%vc(1))
      phi(ivol)%vc=product(g%vol(ivol)%x)
    end do
    do ifce=1,g%n_fce
      if(g%fce(ifce)%in==0)then
        ib=g%fce(ifce)%ib
        allocate(phi(ib)%vc(1))
        phi(ib)%vc=product(g%fce(ifce)%x)
      end if
    end do
    do n_grad=1,6
      call solv_grad(igrd,phi,grad,n_grad,neu=0)
      err=zero
      do ivol=1,g%n_vol
        err(1)=err(1)+(grad(ivol,1,1)-g%vol(ivol)%x(2))**2
        err(2)=err(2)+(grad(ivol,1,2)-g%vol(ivol)%x(1))**2
      end do
      err=sqrt(err/g%n_vol)
      print*,'grad err',n_grad,err
    end do
    do ivol=1,g%n_vbf
      if(associated(phi(ivol)%vc))deallocate(phi(ivol)%vc)
    end do
  end subroutine
end module
