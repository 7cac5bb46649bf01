! ISO_C_BINDING interfaces of the fvk C ABI (include/fvk.h), and the state
! the drop-in modules share: the device context built by grd_solv, the flow
! built on first use by cfd_setup_eqn, and the solver configuration that the
! reference hard-codes as locals (cfd_setup.f90:61,528,580-581,662,742).
module fvk_c
  use iso_c_binding
  implicit none

  integer(c_int),parameter::fvk_ok=0

  type,bind(C)::fvk_stats_t
    integer(c_int)::iters(3)
    real(c_double)::res_rel(3)
    integer(c_int)::fallback
    real(c_double)::seconds
  end type

  type,bind(C)::fvk_flow_cfg_t
    integer(c_int)::simpler,imat_u,imat_p,n_bicg,n_cg,n_grad,c_sch
    real(c_double)::urf_u,urf_p,reduc_u,reduc_p,c_bl,rho,mu
    integer(c_int)::p_solver
    real(c_double)::p_tol
  end type

  type,bind(C)::fvk_outer_hist_t
    real(c_double)::reduc,redpp,redpp_sum,red_p,red_p_sum,mx_res_n
    integer(c_int)::it_u(3),it_pp,it_p
    real(c_double)::res0_n_u(3),res0_n_pp,res0_n_p
  end type

  interface
    function fvk_last_error() bind(C,name='fvk_last_error')
      import::c_ptr
      type(c_ptr)::fvk_last_error
    end function

    integer(c_int) function fvk_ctx_create_arrays(n_dx,n_vol,n_fce,n_bf,fce_ip,fce_in,fce_ib,fce_bty, &
        cell_ptr,fce_lst,vol_x,vol_v,fce_x,fce_n,fce_dxp,fce_dxn,fce_s,fce_lpn,fce_wfp,device,flags,ctx) &
        bind(C,name='fvk_ctx_create_arrays')
      import::c_int,c_double,c_ptr
      integer(c_int),value::n_dx,n_vol,n_fce,n_bf,device,flags
      integer(c_int),intent(in)::fce_ip(*),fce_in(*),fce_ib(*),fce_bty(*),cell_ptr(*),fce_lst(*)
      real(c_double),intent(in)::vol_x(*),vol_v(*),fce_x(*),fce_n(*),fce_dxp(*),fce_dxn(*)
      real(c_double),intent(in)::fce_s(*),fce_lpn(*),fce_wfp(*)
      type(c_ptr)::ctx
    end function

    integer(c_int) function fvk_set_matrix(ctx,pl,ng_p,ng_n) bind(C,name='fvk_set_matrix')
      import::c_int,c_double,c_ptr
      type(c_ptr),value::ctx
      real(c_double),intent(in)::pl(*),ng_p(*),ng_n(*)
    end function

    integer(c_int) function fvk_solv_precond(ctx,imat,c,c_r) bind(C,name='fvk_solv_precond')
      import::c_int,c_double,c_ptr
      type(c_ptr),value::ctx
      integer(c_int),value::imat
      real(c_double)::c(*),c_r(*)
    end function

    integer(c_int) function fvk_solv_direct(ctx,imat,c,c_r,b,x) bind(C,name='fvk_solv_direct')
      import::c_int,c_double,c_ptr
      type(c_ptr),value::ctx
      integer(c_int),value::imat
      real(c_double),intent(in)::c(*),c_r(*),b(*)
      real(c_double)::x(*)
    end function

    integer(c_int) function fvk_solv_res(ctx,n_vc,phi,res0,res0_n,not_r0) bind(C,name='fvk_solv_res')
      import::c_int,c_double,c_ptr
      type(c_ptr),value::ctx
      integer(c_int),value::n_vc,not_r0
      real(c_double),intent(in)::phi(*)
      real(c_double)::res0(*),res0_n(*)
    end function

    integer(c_int) function fvk_solv_cg(ctx,imat,reduc,n_iter,n_vc,phi,res0,res0_n,not_r0,st) &
        bind(C,name='fvk_solv_cg')
      import::c_int,c_double,c_ptr
      type(c_ptr),value::ctx,st
      integer(c_int),value::imat,n_iter,n_vc,not_r0
      real(c_double),value::reduc
      real(c_double)::phi(*),res0(*),res0_n(*)
    end function

    integer(c_int) function fvk_solv_bicg(ctx,imat,reduc,n_iter,n_vc,phi,res0,res0_n,not_r0,st) &
        bind(C,name='fvk_solv_bicg')
      import::c_int,c_double,c_ptr
      type(c_ptr),value::ctx,st
      integer(c_int),value::imat,n_iter,n_vc,not_r0
      real(c_double),value::reduc
      real(c_double)::phi(*),res0(*),res0_n(*)
    end function

    integer(c_int) function fvk_solv_grad(ctx,n_vc,phi,grad,n_iter,neu) bind(C,name='fvk_solv_grad')
      import::c_int,c_double,c_ptr
      type(c_ptr),value::ctx
      integer(c_int),value::n_vc,n_iter,neu
      real(c_double),intent(in)::phi(*)
      real(c_double)::grad(*)
    end function

    integer(c_int) function fvk_flow_create(ctx,cfg,flow) bind(C,name='fvk_flow_create')
      import::c_int,c_ptr,fvk_flow_cfg_t
      type(c_ptr),value::ctx
      type(fvk_flow_cfg_t),intent(in)::cfg
      type(c_ptr)::flow
    end function

    integer(c_int) function fvk_flow_configure(flow,cfg) bind(C,name='fvk_flow_configure')
      import::c_int,c_ptr,fvk_flow_cfg_t
      type(c_ptr),value::flow
      type(fvk_flow_cfg_t),intent(in)::cfg
    end function

    integer(c_int) function fvk_flow_scalars(flow,set,c_in,c_out) bind(C,name='fvk_flow_scalars')
      import::c_int,c_double,c_ptr
      type(c_ptr),value::flow
      integer(c_int),value::set
      real(c_double)::c_in,c_out
    end function

    integer(c_int) function fvk_flow_set_state(flow,u_cell,u_bnd,p,dpdx,spd) bind(C,name='fvk_flow_set_state')
      import::c_int,c_double,c_ptr
      type(c_ptr),value::flow
      real(c_double),intent(in)::u_cell(*),u_bnd(*),p(*),dpdx(*),spd(*)
    end function

    integer(c_int) function fvk_flow_get_state(flow,u_cell,u_bnd,p,dpdx,spd,cs,frc_dns) &
        bind(C,name='fvk_flow_get_state')
      import::c_int,c_double,c_ptr
      type(c_ptr),value::flow
      real(c_double)::u_cell(*),u_bnd(*),p(*),dpdx(*),spd(*),cs(*),frc_dns(*)
    end function

    integer(c_int) function fvk_flow_get_matrix(flow,pl,ng_p,ng_n) bind(C,name='fvk_flow_get_matrix')
      import::c_int,c_double,c_ptr
      type(c_ptr),value::flow
      real(c_double)::pl(*),ng_p(*),ng_n(*)
    end function

    integer(c_int) function fvk_flow_setup_eqn(flow,it,h) bind(C,name='fvk_flow_setup_eqn')
      import::c_int,c_ptr,fvk_outer_hist_t
      type(c_ptr),value::flow
      integer(c_int),value::it
      type(fvk_outer_hist_t)::h
    end function

    integer(c_int) function fvk_flow_setup_pp(flow,it,h) bind(C,name='fvk_flow_setup_pp')
      import::c_int,c_ptr,fvk_outer_hist_t
      type(c_ptr),value::flow
      integer(c_int),value::it
      type(fvk_outer_hist_t)::h
    end function
  end interface

  ! shared state of the drop-in modules
  type(c_ptr),save::fvk_ctx=c_null_ptr
  type(c_ptr),save::fvk_flow=c_null_ptr
  integer,save::fvk_device=0

contains

  ! Stop with the library's message on a fatal status (the reference stops on
  ! blow-up too, cfd_setup.f90:543).
  subroutine fvk_check(rc,where)
    integer(c_int),intent(in)::rc
    character(len=*),intent(in)::where
    character(kind=c_char),pointer::msg(:)
    integer::k
    if(rc>=0)return
    call c_f_pointer(fvk_last_error(),msg,(/512/))
    write(*,'(3a,i4,a)',advance='no')' fvk error in ',where,' (status',rc,'): '
    do k=1,512
      if(msg(k)==c_null_char)exit
      write(*,'(a)',advance='no')msg(k)
    end do
    write(*,*)
    stop 1
  end subroutine

  ! Integer/real settings from the environment (FVK_SIMPLER, FVK_IMAT_P, ...),
  ! defaulting to the values the reference hard-codes.
  integer function fvk_env_int(name,default)
    character(len=*),intent(in)::name
    integer,intent(in)::default
    character(len=32)::v
    integer::st,ios
    fvk_env_int=default
    call get_environment_variable(name,v,status=st)
    if(st==0)then
      read(v,*,iostat=ios)fvk_env_int
      if(ios/=0)fvk_env_int=default
    end if
  end function

  real(c_double) function fvk_env_real(name,default)
    character(len=*),intent(in)::name
    real(c_double),intent(in)::default
    character(len=64)::v
    integer::st,ios
    fvk_env_real=default
    call get_environment_variable(name,v,status=st)
    if(st==0)then
      read(v,*,iostat=ios)fvk_env_real
      if(ios/=0)fvk_env_real=default
    end if
  end function
end module
