! Golden-vector driver for flowlab's serial CSR SIMPLE.  TEST
! INFRASTRUCTURE: linked against flowlab's own Fortran modules, compiled
! where they lie (constants_m, hash_m, graph_m, vector_utils_m,
! matrix_utils_m, solv_m, geometry_m, transport_m); built by
! oracle/build_ref.sh in the development container only.
!
! It sets up the mesh, geometry and equation structures the way
! navierstokes (flowlab/src/navierstokes.f90:161-515) does for one serial
! block, from arrays read from a stream file instead of flowlab's C++ mesh
! pipeline, and then runs the outer loop of navierstokes.f90:515-641
! statement for statement, calling the reference's own routines.  Inputs
! (int32 / float64, little endian, in this order):
!   n_dim n_vrt n_bnd n_dmn n_fce nnz_ng n_iter n_dump
!   cc cd urf_u urf_p bl_c res_drop
!   x_vrt(3, n_vrt)
!   dmn elm_vrt: row(n_dmn+1) (1-based) col(row(n_dmn+1)-1) (vertex ids)
!   bnd physical tags (n_bnd)
!   dmn fce_elm: col(2, n_fce) (global element numbers; boundary -k)
!   dmn fce_vrt: row(n_fce+1) col(row(n_fce+1)-1)
!   dmn elm_neig: row(n_dmn+1) col(nnz_ng) (domain-local, 1-based)
!   vel%phi%bnd(n_dim, n_bnd) (boundary values; navierstokes sets them in
!     its contained set_boundary_conditions)
! Elements are numbered as flowlab numbers them: boundary elements 1..n_bnd,
! domain elements n_bnd+1 .. n_bnd+n_dmn.
!
! Output (stream file): geometry (unless n_dump < 0), then per outer
! iteration the four residual norms and norm_gl (fp_iter_ctrl), and for the
! first n_dump iterations the intermediate fields; the final fields at the
! end.
program flowlab_ref_driver
  use constants_m
  use hash_m
  use graph_m
  use geometry_m
  use transport_m
  use solv_m
  use vector_utils_m
  implicit none

  character(len=1024) :: fin, fout
  integer :: u, v, i, k, iter, n_dim, n_vrt, n_bnd, n_dmn, n_fce, nnz_ng, n_iter, n_dump, fp_exit, m
  real(rp) :: cc, cd, urf_u, urf_p, bl_c, res_drop, res0(10), gl_res0
  integer, dimension(:), allocatable :: tags
  type(mesh_t) :: bnd, dmn
  type(geom_t) :: xbnd, xdmn
  type(pde_t) :: vel, pp, pres

  call get_command_argument(1, fin)
  call get_command_argument(2, fout)
  open(newunit=u, file=trim(fin), access='stream', form='unformatted', status='old')
  read(u) n_dim, n_vrt, n_bnd, n_dmn, n_fce, nnz_ng, n_iter, n_dump
  read(u) cc, cd, urf_u, urf_p, bl_c, res_drop

  ! mesh: one block, boundary elements first (navierstokes.f90:205-300)
  dmn%n_dim = n_dim
  dmn%n_vrt = n_vrt
  allocate(xdmn%x_vrt(n_dim, n_vrt))
  block
    real(rp), allocatable :: x3(:,:)
    allocate(x3(3, n_vrt))
    read(u) x3
    xdmn%x_vrt = x3(:n_dim, :)
  end block
  allocate(dmn%vrt_idx(n_vrt))
  dmn%vrt_idx = (/(i, i = 1, n_vrt)/)

  bnd%blk = 1; bnd%n_blk = 1
  allocate(bnd%elm_blk(2)); bnd%elm_blk = (/1, n_bnd + 1/)
  dmn%blk = 1; dmn%n_blk = 1
  allocate(dmn%elm_blk(2)); dmn%elm_blk = (/n_bnd + 1, n_bnd + n_dmn + 1/)
  bnd%n_elm = n_bnd; dmn%n_elm = n_dmn
  bnd%n_ghst = 0; dmn%n_ghst = 0; bnd%n_locl = 0; dmn%n_locl = 0
  allocate(bnd%elm_idx(n_bnd)); bnd%elm_idx = (/(i, i = 1, n_bnd)/)
  allocate(dmn%elm_idx(n_dmn)); dmn%elm_idx = (/(n_bnd + i, i = 1, n_dmn)/)

  allocate(dmn%elm_vrt%row(n_dmn + 1))
  read(u) dmn%elm_vrt%row
  allocate(dmn%elm_vrt%col(dmn%elm_vrt%row(n_dmn + 1) - 1))
  read(u) dmn%elm_vrt%col

  ! boundary element tags (element type, physical tag) as flowlab lists them
  allocate(tags(n_bnd))
  read(u) tags
  allocate(bnd%elm_tag%row(n_bnd + 1), bnd%elm_tag%col(2 * n_bnd))
  do i = 1, n_bnd
    bnd%elm_tag%row(i) = 2 * (i - 1) + 1
    bnd%elm_tag%col(2 * (i - 1) + 1) = 0
    bnd%elm_tag%col(2 * (i - 1) + 2) = tags(i)
  end do
  bnd%elm_tag%row(n_bnd + 1) = 2 * n_bnd + 1

  dmn%n_fce = n_fce
  allocate(dmn%fce_elm%row(n_fce + 1), dmn%fce_elm%col(2 * n_fce))
  dmn%fce_elm%row = (/(2 * (i - 1) + 1, i = 1, n_fce + 1)/)
  read(u) dmn%fce_elm%col
  allocate(dmn%fce_vrt%row(n_fce + 1))
  read(u) dmn%fce_vrt%row
  allocate(dmn%fce_vrt%col(dmn%fce_vrt%row(n_fce + 1) - 1))
  read(u) dmn%fce_vrt%col

  allocate(dmn%elm_neig%row(n_dmn + 1), dmn%elm_neig%col(nnz_ng))
  read(u) dmn%elm_neig%row
  read(u) dmn%elm_neig%col

  ! vertex index map; no ghosts (navierstokes.f90:302-331)
  call hash_initialise(dmn%vi2i, dmn%n_vrt)
  do i = 1, dmn%n_vrt
    call hash_insert(dmn%vi2i, dmn%vrt_idx(i), i)
  end do
  call hash_initialise(bnd%gi2i, 0); call hash_initialise(bnd%gi2b, 0)
  call hash_initialise(dmn%gi2i, 0); call hash_initialise(dmn%gi2b, 0)

  ! geometry arrays (navierstokes.f90:338-371), zeroed so that entries the
  ! reference leaves unset are defined
  allocate(xdmn%x_vc%dmn(n_dim, n_dmn), xdmn%x_vc%ldmn(n_dim, 0), xdmn%x_vc%gdmn(n_dim, 0))
  allocate(xdmn%vol%dmn(n_dmn), xdmn%vol%ldmn(0), xdmn%vol%gdmn(0))
  allocate(xdmn%gl_min%dmn(n_dim, n_dmn), xdmn%gl_min%ldmn(n_dim, 0), xdmn%gl_min%gdmn(n_dim, 0))
  allocate(xdmn%gl_max%dmn(n_dim, n_dmn), xdmn%gl_max%ldmn(n_dim, 0), xdmn%gl_max%gdmn(n_dim, 0))
  allocate(xdmn%gl%dmn(n_dim, n_dmn), xdmn%gl%ldmn(n_dim, 0), xdmn%gl%gdmn(n_dim, 0))
  allocate(xdmn%ls%dmn(n_dim, n_dim, n_dmn), xdmn%ls%ldmn(n_dim, n_dim, 0), xdmn%ls%gdmn(n_dim, n_dim, 0))
  allocate(xdmn%gg%dmn(n_dim, n_dim, n_dmn), xdmn%gg%ldmn(n_dim, n_dim, 0), xdmn%gg%gdmn(n_dim, n_dim, 0))
  allocate(xdmn%x_fc(n_dim, n_fce), xdmn%x_pc(n_dim, n_fce), xdmn%x_nc(n_dim, n_fce))
  allocate(xdmn%norm(n_dim, n_fce), xdmn%volr(n_dmn), xdmn%area(n_fce), xdmn%l_pn(n_fce), xdmn%w_fp(n_fce))
  xdmn%x_vc%dmn = zero; xdmn%vol%dmn = zero; xdmn%x_fc = zero; xdmn%x_pc = zero; xdmn%x_nc = zero
  xdmn%norm = zero; xdmn%volr = zero; xdmn%area = zero; xdmn%l_pn = zero; xdmn%w_fp = zero

  call geometry(bnd, dmn, xbnd, xdmn)
  call ls_grad_matrix(bnd, dmn, xbnd, xdmn)

  ! equations (navierstokes.f90:385-500)
  vel%n_cmp = n_dim
  allocate(vel%bnd_c(vel%n_cmp, 49), vel%bnd_d(vel%n_cmp, 49))
  allocate(vel%grad%dmn(n_dim, vel%n_cmp, n_dmn), vel%grad%ldmn(n_dim, vel%n_cmp, 0), &
           vel%grad%gdmn(n_dim, vel%n_cmp, 0))
  allocate(vel%phi%dmn(vel%n_cmp, n_dmn), vel%phi%ldmn(vel%n_cmp, 0), vel%phi%gdmn(vel%n_cmp, 0), &
           vel%phi%bnd(vel%n_cmp, n_bnd))
  allocate(vel%rhs(vel%n_cmp, n_dmn), vel%spd(n_fce))
  allocate(vel%cc%dmn(n_dmn), vel%cc%ldmn(0), vel%cc%gdmn(0), vel%cc%bnd(n_bnd))
  allocate(vel%cd%dmn(n_dmn), vel%cd%ldmn(0), vel%cd%gdmn(0), vel%cd%bnd(n_bnd))
  allocate(vel%diag%dmn(n_dmn), vel%diag%ldmn(0), vel%diag%gdmn(0))
  m = n_dmn
  allocate(vel%mat%idx(m), vel%mat%val(nnz_ng))
  vel%mat%idx = dmn%elm_idx - dmn%elm_blk(1) + 1
  vel%mat%row => dmn%elm_neig%row
  vel%mat%col => dmn%elm_neig%col
  vel%urf = urf_u
  vel%bl_c = bl_c
  vel%calc_spd = 1
  vel%bl_d = 1
  vel%spd = zero
  vel%cc%dmn = cc; vel%cc%bnd = cc
  vel%cd%dmn = cd; vel%cd%bnd = cd
  vel%phi%dmn = zero; vel%phi%bnd = zero
  vel%rhs = zero
  vel%mat%val = zero
  vel%diag%dmn = zero
  vel%grad%dmn = zero

  pp%n_cmp = 1
  allocate(pp%grad%dmn(n_dim, 1, n_dmn), pp%grad%ldmn(n_dim, 1, 0), pp%grad%gdmn(n_dim, 1, 0))
  allocate(pp%phi%dmn(1, n_dmn), pp%phi%ldmn(1, 0), pp%phi%gdmn(1, 0), pp%phi%bnd(1, n_bnd))
  allocate(pp%ref(1))
  pp%spd => vel%spd
  pp%cc%dmn => vel%cc%dmn; pp%cc%ldmn => vel%cc%ldmn; pp%cc%gdmn => vel%cc%gdmn; pp%cc%bnd => vel%cc%bnd
  pp%rhs => vel%rhs
  pp%diag%dmn => vel%diag%dmn; pp%diag%ldmn => vel%diag%ldmn; pp%diag%gdmn => vel%diag%gdmn
  pp%mat%idx => vel%mat%idx; pp%mat%val => vel%mat%val
  pp%mat%row => vel%mat%row; pp%mat%col => vel%mat%col
  pp%urf = urf_p
  pv_scheme = 1
  pp%ref = zero
  pp%phi%dmn = zero; pp%phi%bnd = zero
  pp%grad%dmn = zero

  pres%n_cmp = 1
  allocate(pres%grad%dmn(n_dim, 1, n_dmn), pres%grad%ldmn(n_dim, 1, 0), pres%grad%gdmn(n_dim, 1, 0))
  allocate(pres%phi%dmn(1, n_dmn), pres%phi%ldmn(1, 0), pres%phi%gdmn(1, 0), pres%phi%bnd(1, n_bnd))
  pres%phi%dmn = zero; pres%phi%bnd = zero
  pres%grad%dmn = zero

  read(u) vel%phi%bnd
  close(u)

  open(newunit=v, file=trim(fout), access='stream', form='unformatted', status='replace')
  ! n_dump < 0: timing runs, no geometry written
  if (n_dump >= 0) write(v) xdmn%x_vc%dmn, xdmn%vol%dmn, xdmn%x_fc, xdmn%norm, xdmn%area, xdmn%x_pc, &
                            xdmn%x_nc, xdmn%l_pn, xdmn%w_fp, xdmn%ls%dmn

  ! the outer loop (navierstokes.f90:515-641); pARMS calls are no-ops in a
  ! serial build and are left out; res0 starts defined (the reference
  ! leaves res0(5:10) undefined until fp_iter_ctrl zeroes it)
  res0 = zero
  fp_exit = 0
  do iter = 1, n_iter
    vel%rhs = zero
    vel%mat%val = zero

    call ls_gradient(bnd, dmn, xbnd, xdmn, pres%phi, pres%grad, 1)
    call ls_gradient(bnd, dmn, xbnd, xdmn, vel%phi, vel%grad, 0)
    if (iter <= n_dump) write(v) pres%grad%dmn, pres%phi%bnd, vel%grad%dmn

    call ns_source(bnd, dmn, xbnd, xdmn, pres, vel)
    call construct_pde(bnd, dmn, xbnd, xdmn, vel)
    if (iter <= n_dump) write(v) vel%mat%val, vel%rhs, vel%diag%dmn, vel%phi%bnd
    do i = 1, vel%n_cmp
      call lin_eqns_solver(vel%mat%val, vel%phi%dmn(i,:), vel%rhs(i,:), &
                           res0(i), vel%mat%col, vel%mat%row, 25, 0.01_rp, 2, i)
    end do
    if (iter <= n_dump) write(v) vel%phi%dmn, vel%rhs, res0(1:3)

    pp%rhs = zero
    pp%mat%val = zero
    call construct_pp(bnd, dmn, xbnd, xdmn, vel, pres, pp)
    if (iter <= n_dump) write(v) pp%mat%val, vel%rhs(1,:), vel%spd

    i = 4
    pp%phi%dmn(1,:) = zero
    call lin_eqns_solver(pp%mat%val, pp%phi%dmn(1,:), pp%rhs(1,:), &
                         res0(i), vel%mat%col, vel%mat%row, 25, 0.01_rp, 1, i)
    if (iter <= n_dump) write(v) pp%phi%dmn, res0(4)

    call ls_gradient(bnd, dmn, xbnd, xdmn, pp%phi, pp%grad, 1)
    call ns_correction(bnd, dmn, xbnd, xdmn, pp, vel, pres)
    if (iter <= n_dump) write(v) pp%grad%dmn, pp%phi%bnd, vel%phi%dmn, pres%phi%dmn

    write(v) res0(1:4)
    call fp_iter_ctrl('F', 1, iter, n_iter, -1, fp_exit, res_drop, res0, gl_res0)
    write(v) gl_res0
    if (fp_exit > 1) exit
  end do
  write(v) min(iter, n_iter)
  write(v) vel%phi%dmn, pres%phi%dmn, pp%phi%dmn, vel%spd, vel%phi%bnd, pres%phi%bnd
  close(v)
end program
