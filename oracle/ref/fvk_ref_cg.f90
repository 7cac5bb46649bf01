! Timing driver for the reference's PCG (solv_cg.f90) on a stand-alone
! pressure system: the CPU baseline of bench.py.
!
! TEST INFRASTRUCTURE.  This program is the project's own code; it links the
! reference's solver modules compiled from /root/reference/spray-simulation
! (see oracle/build_ref.sh).  It reads a system from a stream file, fills
! the module data of solv_data_m as grd_solv.f90 and cfd_setup.f90 leave it
! (pl: diagonal; ng: per face ip, in (0 on boundary faces), p = A(ip,in),
! n = A(in,ip)), and times one call of solv_cg with imat = 0 (Jacobi) for a
! fixed number of iterations (reduc = 0: the exit test never passes), or,
! given REDUC, up to N_ITER iterations with that relative L1 exit test.
!
! File layout (little endian): int32 n_vol, n_fce; int32 ip(n_fce),
! in(n_fce) (1-based); float64 p(n_fce), n(n_fce), pl(n_vol), b(n_vol).
!
! usage: fvk_ref_cg SYSTEM_FILE N_ITER [PHI_FILE [REDUC]]
! (PHI_FILE: the solution, float64 stream)
program fvk_ref_cg
  use solv_cg_m
  implicit none
  character(len=512)::sys_file,arg,phi_file
  integer::nv,nf,n_iter,i
  integer(4)::hdr(2)
  integer(4),allocatable::ip(:),in(:)
  real(8),allocatable::pv(:),nv_(:),b(:)
  type(phi_t),allocatable::phi(:)
  real(wp),allocatable::res0(:,:),res0_n(:)
  real(wp)::reduc
  integer(8)::t0,t1,rate

  call get_command_argument(1,sys_file)
  call get_command_argument(2,arg); read(arg,*)n_iter
  reduc=zero
  if(command_argument_count()>=4)then
    call get_command_argument(4,arg); read(arg,*)reduc
  end if
  open(10,file=trim(sys_file),access='stream',form='unformatted',status='old')
  read(10)hdr
  nv=hdr(1); nf=hdr(2)
  allocate(ip(nf),in(nf),pv(nf),nv_(nf),b(nv))
  allocate(pl(nv),ng(nf))
  read(10)ip
  read(10)in
  read(10)pv
  read(10)nv_
  read(10)pl
  read(10)b
  close(10)
  do i=1,nf
    ng(i)%ip=ip(i); ng(i)%in=in(i); ng(i)%p=pv(i); ng(i)%n=nv_(i)
  end do
  allocate(phi(nv),res0(nv,1),res0_n(1))
  do i=1,nv
    allocate(phi(i)%vc(1)); phi(i)%vc=zero
  end do
  res0(:,1)=b

  call system_clock(t0,rate)
  call solv_cg(0,reduc,n_iter,phi,res0,res0_n,.false.)
  call system_clock(t1)
  print '(a,i0,a,es16.8,a,es16.8)','iterations ',n_iter,' seconds ',dble(t1-t0)/dble(rate), &
    ' res0_n ',res0_n(1)
  if(command_argument_count()>=3)then
    call get_command_argument(3,phi_file)
    open(11,file=trim(phi_file),access='stream',form='unformatted',status='replace')
    write(11)(phi(i)%vc(1),i=1,nv)
    close(11)
  end if
end program
