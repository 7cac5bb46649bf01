! Golden-output driver of the reference's Tecplot writers (cfd_outp.f90).
!
! TEST INFRASTRUCTURE.  This program is the project's own code; it links the
! reference modules compiled from /root/reference/spray-simulation (see
! oracle/build_ref.sh).  It sets up the same single-phase steady flow as
! fvk_ref_driver.f90, runs N_OUTER SIMPLE iterations, and writes the
! velocity with the reference's eqn_outp(1,1,1,0,DIR) and the pressure with
! prp_outp(1,1,1,0,DIR,p=.true.) (grd1%phs1%eqn1.dat, grd1%phs1%prp1.dat).
!
! usage: fvk_ref_outp VRT TECVRT NEU DIR N_OUTER RHO MU URF_P
!                     BND_CODE1 U1 V1 W1 [BND_CODE2 U2 V2 W2]
program fvk_ref_outp
  use grd_geom_m
  use grd_solv_m
  use cfd_setup_m
  use cfd_outp_m
  implicit none
  character(len=512)::vrt_file,tec_file,neu_file,dir,arg
  integer::n_outer,iter,ifce,c,k,narg,nbc
  integer::bcode(2)
  real(wp)::rho,mu,urf_p,bvel(3,2)
  type(grd_t),pointer::gp

  call get_command_argument(1,vrt_file)
  call get_command_argument(2,tec_file)
  call get_command_argument(3,neu_file)
  call get_command_argument(4,dir)
  call get_command_argument(5,arg); read(arg,*)n_outer
  call get_command_argument(6,arg); read(arg,*)rho
  call get_command_argument(7,arg); read(arg,*)mu
  call get_command_argument(8,arg); read(arg,*)urf_p
  narg=command_argument_count()
  nbc=(narg-8)/4
  bvel=zero; bcode=-1
  do k=1,nbc
    call get_command_argument(9+4*(k-1),arg); read(arg,*)bcode(k)
    do c=1,3
      call get_command_argument(9+4*(k-1)+c,arg); read(arg,*)bvel(c,k)
    end do
  end do

  call grd_prepr(trim(vrt_file))
  call grd_postpr(file_nm=trim(tec_file))
  n_tl=1; n_grd=1; l_sc=one; cyl=.false.
  allocate(grd(n_grd))
  call grd_read(1,trim(neu_file))
  call grd_conn(1)
  call grd_geom(1)
  call grd_solv(1)
  call c_std(1)
  gp=>grd(1)

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
  do ifce=1,gp%n_fce
    if(gp%fce(ifce)%in==0)then
      do k=1,nbc
        if(gp%fce(ifce)%bnd_ty==bcode(k))gp%fce(ifce)%phs(1)%eqn(1)%phi(1,:,1)=bvel(1:n_dx,k)
      end do
    end if
  end do
  do iter=1,n_outer
    mx_res_n=zero; c_in=zero; c_out=zero
    call cfd_setup_eqn(1,1,1,1,iter,.true.)
    call cfd_setup_pp(1,1,1,iter,n_pp=1,urf=urf_p)
  end do
  call eqn_outp(1,1,1,0,trim(dir))
  call prp_outp(1,1,1,0,trim(dir),p=.true.)
end program
