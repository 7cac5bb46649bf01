! Golden-vector driver for the flang build of the spray-simulation reference.
!
! TEST INFRASTRUCTURE.  This program is the project's own code; it links the
! reference modules compiled from /root/reference/spray-simulation (see
! oracle/build_ref.sh) and replaces the reference's main.f90/cfd_solve.f90
! driver (which needs missing case files) with a plain single-phase, steady
! cavity/channel set-up:
!   cfd_solve.f90:63-72  properties and momentum equation set-up
!   cfd_solve.f90:134-178 outer loop (iph=1 calls only)
! It dumps mesh, fields, residual histories and stand-alone solver/gradient
! results as a binary record stream read by tests/golden/make_golden.py.
!
! usage: fvk_ref_driver VRT NEU OUT N_OUTER RHO MU URF_P
!                       BND_CODE1 U1 V1 W1 [BND_CODE2 U2 V2 W2]
program fvk_ref_driver
  use grd_geom_m
  use grd_solv_m
  use cfd_setup_m
  implicit none
  character(len=512)::vrt_file,neu_file,out_file,arg
  integer::n_outer,iter,ivol,ifce,i,c,nv,nf,imat,k,nslot,narg,ib,nbc
  integer::bcode(2)
  real(wp)::rho,mu,urf_p,bvel(3,2)
  real(wp),allocatable,target::x3(:,:)
  real(wp),allocatable::b3(:,:),res(:,:),rn(:),cc(:),cr(:),z(:),bb(:)
  real(wp),allocatable::hist(:),tmp(:),g(:,:,:)
  integer,allocatable::itmp(:)
  type(phi_t),allocatable::phs(:)
  type(grd_t),pointer::gp
  integer,parameter::uo=77
  integer(8)::t0,t1,t2,t3,crate

  call get_command_argument(1,vrt_file)
  call get_command_argument(2,neu_file)
  call get_command_argument(3,out_file)
  call get_command_argument(4,arg); read(arg,*)n_outer
  call get_command_argument(5,arg); read(arg,*)rho
  call get_command_argument(6,arg); read(arg,*)mu
  call get_command_argument(7,arg); read(arg,*)urf_p
  narg=command_argument_count()
  nbc=(narg-7)/4
  bvel=zero; bcode=-1
  do k=1,nbc
    call get_command_argument(8+4*(k-1),arg); read(arg,*)bcode(k)
    do c=1,3
      call get_command_argument(8+4*(k-1)+c,arg); read(arg,*)bvel(c,k)
    end do
  end do

  call grd_prepr(trim(vrt_file))
  n_tl=1; n_grd=1; l_sc=one; cyl=.false.
  allocate(grd(n_grd))
  call grd_read(1,trim(neu_file))
  call grd_conn(1)
  call grd_geom(1)
  call grd_solv(1)
  call c_std(1)
  gp=>grd(1)
  nv=gp%n_vol; nf=gp%n_fce

  open(uo,file=trim(out_file),access='stream',form='unformatted',status='replace')

  ! ---- mesh ----
  call put_i('n_dx',(/n_dx/))
  call put_i('n_vol',(/nv/))
  call put_i('n_fce',(/nf/))
  call put_i('n_bf',(/gp%n_bf/))
  call put_r('small',(/small/))
  call put_r('large',(/large/))
  allocate(tmp(max(nv,nf)*3),itmp(nf*6+8))
  do ivol=1,nv; tmp((ivol-1)*n_dx+1:ivol*n_dx)=gp%vol(ivol)%x; end do
  call put_r('vol_x',tmp(1:nv*n_dx))
  call put_r('vol_v',(/(gp%vol(ivol)%v,ivol=1,nv)/))
  call put_r('vol_vr',(/(gp%vol(ivol)%v_r,ivol=1,nv)/))
  call put_i('fce_ip',(/(gp%fce(ifce)%ip,ifce=1,nf)/))
  call put_i('fce_in',(/(gp%fce(ifce)%in,ifce=1,nf)/))
  call put_i('fce_ib',(/(gp%fce(ifce)%ib,ifce=1,nf)/))
  call put_i('fce_bty',(/(gp%fce(ifce)%bnd_ty,ifce=1,nf)/))
  call put_r('fce_s',(/(gp%fce(ifce)%s,ifce=1,nf)/))
  call put_r('fce_lpn',(/(gp%fce(ifce)%l_pn,ifce=1,nf)/))
  call put_r('fce_wfp',(/(gp%fce(ifce)%w_fp,ifce=1,nf)/))
  do ifce=1,nf; tmp((ifce-1)*n_dx+1:ifce*n_dx)=gp%fce(ifce)%x; end do
  call put_r('fce_x',tmp(1:nf*n_dx))
  do ifce=1,nf; tmp((ifce-1)*n_dx+1:ifce*n_dx)=gp%fce(ifce)%n; end do
  call put_r('fce_n',tmp(1:nf*n_dx))
  do ifce=1,nf; tmp((ifce-1)*n_dx+1:ifce*n_dx)=gp%fce(ifce)%dx_p; end do
  call put_r('fce_dxp',tmp(1:nf*n_dx))
  do ifce=1,nf
    tmp((ifce-1)*n_dx+1:ifce*n_dx)=zero
    if(gp%fce(ifce)%in>0)tmp((ifce-1)*n_dx+1:ifce*n_dx)=gp%fce(ifce)%dx_n
  end do
  call put_r('fce_dxn',tmp(1:nf*n_dx))
  nslot=0
  do ivol=1,nv
    do k=1,gp%vol(ivol)%n_fce
      nslot=nslot+1; itmp(nslot)=gp%vol(ivol)%fce_lst(k)
    end do
  end do
  call put_i('fce_lst',itmp(1:nslot))
  nslot=0
  do ivol=1,nv
    do k=1,gp%vol(ivol)%n_fce
      nslot=nslot+1; itmp(nslot)=gp%vol(ivol)%ng_lst(k)
    end do
  end do
  call put_i('ng_lst',itmp(1:nslot))

  ! ---- stand-alone gradient of phi = product of centre coordinates
  !      (the test_grad field, cfd_setup.f90:882-895) ----
  allocate(phs(gp%n_vbf),g(nv,1,n_dx))
  do ivol=1,nv
    allocate(phs(ivol)%vc(1)); phs(ivol)%vc=product(gp%vol(ivol)%x)
  end do
  do ifce=1,nf
    if(gp%fce(ifce)%in==0)then
      ib=gp%fce(ifce)%ib
      allocate(phs(ib)%vc(1)); phs(ib)%vc=product(gp%fce(ifce)%x)
    end if
  end do
  do k=1,3
    call solv_grad(1,phs,g,k,neu=0)
    call put_r('grad0_'//achar(48+k),reshape(g,(/nv*n_dx/)))
    call solv_grad(1,phs,g,k,neu=2)
    call put_r('grad2_'//achar(48+k),reshape(g,(/nv*n_dx/)))
  end do
  do ivol=1,gp%n_vbf
    if(associated(phs(ivol)%vc))deallocate(phs(ivol)%vc)
  end do
  deallocate(phs,g)

  ! ---- single-phase steady flow (cfd_solve.f90:63-72 with turb=.false.) ----
  dbg=.false.; turb=.false.; spr=.false.; n_grad=2
  call c_phs(1,1)
  call c_prp(1,1,5)
  call c_prp_rho(1,1,ipr=1,rho=one,s_tl=.true.)
  call c_prp_rho(1,1,ipr=2,rho=rho,s_tl=.false.)
  call c_prp_rho(1,1,ipr=3,rho=mu,s_tl=.false.)
  call c_prp_rho(1,1,ipr=4,rho=zero,s_tl=.false.)
  call c_prp_rho(1,1,ipr=5,rho=zero,s_tl=.false.)
  call c_eqn(1,1,1)
  call c_eqn_phi(1,1,ieq=1,n_sc=1,n_vc=n_dx,urf=0.7_wp,c_bl=0.9_wp,c_sch=1,rhs=.true.)
  do ifce=1,nf
    if(gp%fce(ifce)%in==0)then
      do k=1,nbc
        if(gp%fce(ifce)%bnd_ty==bcode(k))gp%fce(ifce)%phs(1)%eqn(1)%phi(1,:,1)=bvel(1:n_dx,k)
      end do
    end if
  end do

  allocate(hist(n_outer))
  allocate(x3(n_dx,nv),b3(nv,n_dx),res(nv,n_dx),rn(n_dx),cc(nv),cr(nv),z(nv),bb(nv))
  allocate(phs(nv))
  do ivol=1,nv; phs(ivol)%vc=>x3(:,ivol); end do
  do ivol=1,nv
    do c=1,n_dx
      b3(ivol,c)=real(mod(ivol*7919+c*104729,1013),wp)/1013._wp-0.5_wp
    end do
  end do
  bb=b3(:,1)

  call system_clock(count_rate=crate)
  do iter=1,n_outer
    mx_res_n=zero; c_in=zero; c_out=zero
    call system_clock(t0)
    call cfd_setup_eqn(1,1,1,1,iter,.true.)
    call system_clock(t1)
    if(iter==n_outer)then
      ! momentum matrix and stand-alone solves on it (solv_*.f90)
      call put_r('mom_pl',pl)
      call put_r('mom_ngp',(/(ng(ifce)%p,ifce=1,nf)/))
      call put_r('mom_ngn',(/(ng(ifce)%n,ifce=1,nf)/))
      do imat=0,2
        call solv_precond(imat,cc,cr)
        call put_r('mom_c'//achar(48+imat),cc)
        call put_r('mom_cr'//achar(48+imat),cr)
        call solv_direct(imat,cc,cr,z,bb)
        call put_r('mom_dir'//achar(48+imat),z)
        x3=zero; res=b3
        call solv_bicg(imat,1.e-30_wp,6,phs,res,rn,.false.)
        call put_r('mom_bicg_x'//achar(48+imat),reshape(x3,(/nv*n_dx/)))
        call put_r('mom_bicg_r'//achar(48+imat),reshape(res,(/nv*n_dx/)))
        call put_r('mom_bicg_n'//achar(48+imat),rn)
        x3=zero; res=b3
        call solv_bicg(imat,0.1_wp,5,phs,res,rn,.false.)
        call put_r('mom_bicgr_x'//achar(48+imat),reshape(x3,(/nv*n_dx/)))
      end do
    end if
    call system_clock(t2)
    call cfd_setup_pp(1,1,1,iter,n_pp=1,urf=urf_p)
    call system_clock(t3)
    hist(iter)=mx_res_n
    write(*,'(a,2i6,es26.17e3)')' REDUC17:',1,iter,mx_res_n
    ! wall time of the iteration's cfd_setup_eqn and cfd_setup_pp (the
    ! stand-alone solves of the last iteration excluded)
    write(*,'(a,i6,es14.6)')' TIME17:',iter,real(t1-t0+t3-t2,wp)/real(crate,wp)
  end do
  call put_r('hist_REDUC',hist)

  ! ---- fields ----
  do ivol=1,nv
    tmp((ivol-1)*n_dx+1:ivol*n_dx)=gp%vol(ivol)%phs(1)%eqn(1)%phi(1,:,1)
  end do
  call put_r('u_cell',tmp(1:nv*n_dx))
  deallocate(tmp); allocate(tmp(max(gp%n_bf,1)*n_dx)); tmp=zero
  do ifce=1,nf
    if(gp%fce(ifce)%in==0)then
      ib=gp%fce(ifce)%ib-nv
      tmp((ib-1)*n_dx+1:ib*n_dx)=gp%fce(ifce)%phs(1)%eqn(1)%phi(1,:,1)
    end if
  end do
  call put_r('u_bnd',tmp(1:gp%n_bf*n_dx))
  call put_r('p',(/(gp%vol(ivol)%p(1),ivol=1,nv)/))
  deallocate(tmp); allocate(tmp(nv*n_dx))
  do ivol=1,nv; tmp((ivol-1)*n_dx+1:ivol*n_dx)=gp%vol(ivol)%dpdx; end do
  call put_r('dpdx',tmp)
  call put_r('spd',(/(gp%fce(ifce)%spd,ifce=1,nf)/))
  call put_r('frc_dns',(/(gp%fce(ifce)%frc_dns,ifce=1,nf)/))
  call put_r('cs',(/(gp%vol(ivol)%cs,ivol=1,nv)/))
  call put_r('pp_pl',pl)
  call put_r('pp_ngp',(/(ng(ifce)%p,ifce=1,nf)/))

  ! ---- stand-alone PCG on the final pressure-correction matrix ----
  deallocate(phs); allocate(phs(nv))
  do ivol=1,nv; phs(ivol)%vc=>x3(1:1,ivol); end do
  bb=b3(:,1)-sum(b3(:,1))/nv
  do imat=0,2
    call solv_precond(imat,cc,cr)
    call put_r('pp_c'//achar(48+imat),cc)
    call solv_direct(imat,cc,cr,z,bb)
    call put_r('pp_dir'//achar(48+imat),z)
    x3=zero; res(:,1)=bb
    call solv_cg(imat,1.e-30_wp,15,phs,res(:,1:1),rn(1:1),.false.)
    call put_r('pp_cg_x'//achar(48+imat),x3(1,:))
    call put_r('pp_cg_n'//achar(48+imat),rn(1:1))
    x3=zero; res(:,1)=bb
    call solv_cg(imat,0.1_wp,20,phs,res(:,1:1),rn(1:1),.true.)
    call put_r('pp_cgr_x'//achar(48+imat),x3(1,:))
  end do
  close(uo)

contains

  subroutine put_hdr(name,kind,n)
    character(len=*),intent(in)::name
    integer,intent(in)::kind,n
    character(len=16)::nm
    nm=name
    write(uo)nm,int(kind,4),int(n,8)
  end subroutine

  subroutine put_i(name,a)
    character(len=*),intent(in)::name
    integer,dimension(:),intent(in)::a
    call put_hdr(name,1,size(a))
    write(uo)int(a,4)
  end subroutine

  subroutine put_r(name,a)
    character(len=*),intent(in)::name
    real(wp),dimension(:),intent(in)::a
    call put_hdr(name,2,size(a))
    write(uo)a
  end subroutine
end program
