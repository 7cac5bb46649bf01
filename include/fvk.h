/*
 * fvk — MI355X (gfx950) implementation of the spray-simulation SIMPLE hot
 * path behind a plain C ABI.
 *
 * Conventions
 *   - Every function returns an int status: FVK_OK (0), > 0 recoverable
 *     (a fallback was taken), < 0 fatal (nothing was changed that the caller
 *     can rely on).  fvk_last_error() describes the last failure on the
 *     calling thread.  No exception crosses the ABI.
 *   - Entity numbering on the ABI is the reference's: cells 1..n_vol,
 *     faces 1..n_fce in grd_conn order, in = 0 for a boundary face,
 *     boundary values addressed by ib = n_vol+1 .. n_vol+n_bf
 *     (spray-simulation/grd_data.f90:22-53, grd_geom.f90:46-69).
 *   - Vectors per entity are contiguous in the reference's order
 *     (x(1:n_dx) of face 1, then face 2, ...).  Multi-component fields
 *     phi(i)%vc(c) are passed entity-major, stride n_vc; residual arrays
 *     res0(n_vol, n_vc) column-major as the reference declares them.
 *   - Host buffers are borrowed for the duration of a call; device memory is
 *     owned by the context.  A context is used from one host thread at a
 *     time and runs on its own HIP stream.
 */
#ifndef FVK_H
#define FVK_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define FVK_OK 0
#define FVK_FALLBACK 1        /* learned/mixed path fell back to plain PCG */
#define FVK_NOT_CONVERGED 2   /* iteration cap reached before the tolerance */
#define FVK_ERR_ARG (-1)
#define FVK_ERR_HIP (-2)
#define FVK_ERR_STATE (-3)
#define FVK_ERR_DIVERGED (-4) /* |normalised residual| > large: the reference stops here */
#define FVK_ERR_IO (-5)
#define FVK_ERR_NODEV (-6)

const char *fvk_last_error(void);
int fvk_version(void);

/* ------------------------------------------------------------------------
 * Mesh (host): the grd_read / grd_conn / grd_geom layer.  Output arrays are
 * bit-identical to the reference's grd structure; grd_conn's O(n^2) vertex
 * matching is replaced by an O(N) hash with the same face ordering.
 * ---------------------------------------------------------------------- */
typedef struct fvk_mesh fvk_mesh;

typedef struct fvk_mesh_info {
  int n_dx, n_vrt, n_vol, n_fce, n_bf, n_slot, max_fce;
} fvk_mesh_info;

/* grd_read (spray-simulation/grd_read.f90:9-110) + grd_conn + grd_geom.
   l_sc scales coordinates (grd_data.f90:9). */
int fvk_mesh_read_neu(const char *path, double l_sc, fvk_mesh **out);
/* Box of nx*ny*nz hexes (nz = 0: nx*ny quads) on [0,lx]x[0,ly]x[0,lz].
   Interior vertices are moved by U(-jitter, jitter)*h per coordinate, drawn
   from SplitMix64(seed) in vertex order x, y, z.  Faces on z = lz (y = ly in
   2D) get boundary code lid_code, the others wall_code. */
int fvk_mesh_box(int nx, int ny, int nz, double lx, double ly, double lz, double jitter,
                 uint64_t seed, int lid_code, int wall_code, fvk_mesh **out);
/* Cells given directly: cell i (1-based) has Gambit element type v_ty[i-1]
   and vertices v_vrt[v_vptr[i-1] .. v_vptr[i]-1] (1-based vertex ids);
   bnd holds one boundary code per (cell, local face), 0 = none. */
int fvk_mesh_from_cells(int n_dx, int n_vrt, const double *vx3, int n_vol, const int *v_ty,
                        const int *v_vptr, const int *v_vrt, const int *bnd, fvk_mesh **out);
/* gmsh 2.2 ASCII mesh (flowlab/src/mesh.cpp:39-187 read_gmsh_mesh, serial
   mode) + grd_conn + grd_geom: elements of dimension n_dx (n_dx <= 0: the
   highest present) are the cells, in file order (hexahedra, prisms and
   tetrahedra in 3D, quadrilaterals and triangles in 2D; vertex orders
   converted to Gambit's); elements of dimension n_dx - 1 mark boundary
   faces with their physical tag, mapped through code_map (n_map pairs
   physical tag -> boundary code, e.g. 11 inlet, 21 outlet, 31 wall; an
   unmapped tag is used as the code).  l_sc scales coordinates.  Each
   cell's gmsh element number is kept (fvk_mesh_get_elm_idx). */
int fvk_mesh_read_gmsh(const char *path, int n_dx, double l_sc, const int *code_map, int n_map,
                       fvk_mesh **out);
/* Where the mesh constructors (fvk_mesh_read_neu, fvk_mesh_read_gmsh,
   fvk_mesh_box, fvk_mesh_from_cells) build connectivity and geometry: on
   GPU `device`, or on the host (-1, the default).  Both produce the same
   arrays bit for bit; the GPU build takes a fraction of the host's time on
   meshes of millions of cells.  No counterpart in the reference. */
int fvk_mesh_set_build_device(int device);
/* Source element number of each cell (n_vol; 1..n_vol for meshes that
   have none), e.g. for fvk_write_gmsh_field. */
int fvk_mesh_get_elm_idx(const fvk_mesh *m, int *elm_idx);
/* Vertex coordinates (n_vrt*3) and cell vertex lists (cell_vptr n_vol+1,
   cell_vrt 1-based, Gambit order), as the reference's grd structure holds
   them. */
int fvk_mesh_get_cells(const fvk_mesh *m, double *vx3, int *v_ty, int *cell_vptr, int *cell_vrt);
int fvk_mesh_get_info(const fvk_mesh *m, fvk_mesh_info *info);
/* Topology as the reference holds it: fce%ip, fce%in, fce%ib, fce%bnd_ty
   (n_fce each), cell_ptr (n_vol+1, 0-based offsets) and the per-cell
   fce_lst / ng_lst (n_slot each, grd_conn.f90:128-152, 218-238). */
int fvk_mesh_get_topology(const fvk_mesh *m, int *fce_ip, int *fce_in, int *fce_ib, int *fce_bty,
                          int *cell_ptr, int *fce_lst, int *ng_lst);
/* Geometry (grd_geom.f90): vol_x (n_vol*n_dx), vol_v, fce_x, fce_n, fce_dxp,
   fce_dxn (n_fce*n_dx; dx_n is 0 on boundary faces), fce_s, fce_lpn, fce_wfp. */
int fvk_mesh_get_geometry(const fvk_mesh *m, double *vol_x, double *vol_v, double *fce_x,
                          double *fce_n, double *fce_dxp, double *fce_dxn, double *fce_s,
                          double *fce_lpn, double *fce_wfp);
void fvk_mesh_free(fvk_mesh *m);

/* ------------------------------------------------------------------------
 * Device context: topology and geometry resident in HBM.
 * Replaces grd_solv (spray-simulation/grd_solv.f90:7-24): the cell-face
 * lists become a slot-major ELL table, plus level sets for the triangular
 * sweeps of solv_precond / solv_direct.
 * ---------------------------------------------------------------------- */
typedef struct fvk_ctx fvk_ctx;

#define FVK_CTX_RCM 1 /* store cells in reverse Cuthill-McKee order */

int fvk_ctx_create(const fvk_mesh *m, int device, int flags, fvk_ctx **out);
/* Same, from the arrays the reference's grd structure holds after grd_geom
   (the Fortran drop-in path; layouts as fvk_mesh_get_topology/geometry). */
int fvk_ctx_create_arrays(int n_dx, int n_vol, int n_fce, int n_bf, const int *fce_ip,
                          const int *fce_in, const int *fce_ib, const int *fce_bty,
                          const int *cell_ptr, const int *fce_lst, const double *vol_x,
                          const double *vol_v, const double *fce_x, const double *fce_n,
                          const double *fce_dxp, const double *fce_dxn, const double *fce_s,
                          const double *fce_lpn, const double *fce_wfp, int device, int flags,
                          fvk_ctx **out);
void fvk_ctx_destroy(fvk_ctx *ctx);
int fvk_ctx_sync(fvk_ctx *ctx);
/* The reverse Cuthill-McKee order FVK_CTX_RCM gives a context on this
   mesh (perm[k] = reference cell id, 1-based, of row k); host only, no
   device needed.  No counterpart in the reference (cells keep the
   Gambit order there). */
int fvk_mesh_rcm(const fvk_mesh *m, int *perm);
/* Cell order used on the device: perm[k] = reference cell id (1-based) of
   device row k; identity unless FVK_CTX_RCM. */
int fvk_ctx_get_perm(const fvk_ctx *ctx, int *perm);
/* Number of forward level sets of the lower-triangular sweep. */
int fvk_ctx_levels(const fvk_ctx *ctx, int *n_fwd, int *n_bwd);

/* ------------------------------------------------------------------------
 * Linear solvers on the face-based matrix: diagonal pl(n_vol) and per face
 * ng%p / ng%n (spray-simulation/solv_data.f90:14-22).
 * ---------------------------------------------------------------------- */
/* Upload pl, ng%p, ng%n (host, reference order). */
int fvk_set_matrix(fvk_ctx *ctx, const double *pl, const double *ng_p, const double *ng_n);
/* solv_precond (spray-simulation/solv_precond.f90:5-30): imat 0 Jacobi,
   1 DIC, 2 DILU; c = 1/(d+small), c_r = d. */
int fvk_solv_precond(fvk_ctx *ctx, int imat, double *c, double *c_r);
/* solv_direct (spray-simulation/solv_direct.f90:5-43): x = M^-1 b with the
   factor c, c_r as solv_precond returns it; c = c_r = NULL builds the factor
   of the current matrix. */
int fvk_solv_direct(fvk_ctx *ctx, int imat, const double *c, const double *c_r, const double *b,
                    double *x);
/* solv_res (spray-simulation/solv_res.f90:5-30): res0 <- res0 - A phi unless
   not_r0, res0_n = per-component L1 norm. */
int fvk_solv_res(fvk_ctx *ctx, int n_vc, const double *phi, double *res0, double *res0_n,
                 int not_r0);
/* fvk_solv_res on device memory, allocating nothing.  No counterpart in the
   reference: for callers that cannot allocate device memory (code under
   HIP graph capture, tight serving loops).  phi_dev and res0_dev are device
   arrays of n_vc components, component-major in the context's row order
   (element (row, q) at [q*n_vol + row]; the reference's order unless the
   context was built with FVK_CTX_RCM); res0_n_dev is a device array of n_vc
   doubles.  The work is queued on the context's stream and nothing waits
   for it (fvk_ctx_sync does).  Its reduction partials live in a fixed
   48 KiB static device array per GPU (every block's partials are written
   before they are read: nothing carries over between calls); n_vc is at
   most 3.  Calls that overlap on one GPU (contexts on other streams) race
   on that array. */
int fvk_solv_res_dev(fvk_ctx *ctx, int n_vc, const double *phi_dev, double *res0_dev, double *res0_n_dev,
                     int not_r0);

typedef struct fvk_stats {
  int iters[3];      /* iterations per component */
  double res_rel[3]; /* final relative L1 residual per component */
  int fallback;      /* learned/mixed path fell back */
  double seconds;    /* device time of the solve */
} fvk_stats;

/* solv_cg (spray-simulation/solv_cg.f90:7-87) and solv_bicg
   (spray-simulation/solv_bicg.f90:7-121).  phi: n_vol x n_vc entity-major
   (in/out), res0: column-major (in: RHS, out: initial residual),
   res0_n: out.  The n_vc components share the matrix and factor; they are
   solved as one batched launch sequence, each with its own scalars. */
int fvk_solv_cg(fvk_ctx *ctx, int imat, double reduc, int n_iter, int n_vc, double *phi,
                double *res0, double *res0_n, int not_r0, fvk_stats *st);
int fvk_solv_bicg(fvk_ctx *ctx, int imat, double reduc, int n_iter, int n_vc, double *phi,
                  double *res0, double *res0_n, int not_r0, fvk_stats *st);

/* solv_grad (spray-simulation/cfd_setup.f90:765-866): phi has n_vol+n_bf
   entries of n_vc components (cells, then boundary slots ib); grad is
   n_vol x n_vc x n_dx, entity-major (grad[(i*n_vc+c)*n_dx+d]). */
int fvk_solv_grad(fvk_ctx *ctx, int n_vc, const double *phi, double *grad, int n_iter, int neu);

/* ------------------------------------------------------------------------
 * Stand-alone pressure systems held in HBM (no host transfer inside a
 * solve).  The matrix is the symmetric face-based form of the pressure
 * correction (cfd_setup.f90:623-634): A_ij = -a_f, A_ii = sum of a_f.
 * ---------------------------------------------------------------------- */
typedef struct fvk_solve_info {
  int iters;
  int fallback;          /* learned/mixed path fell back to plain PCG */
  double seconds;        /* device time of the solve (HIP events) */
  double res_rec;        /* final relative L1 norm of the recurrence residual */
  double res_true_l1;    /* |b - A x|_1 / |b|_1 in FP64 */
  double res_true_l2;    /* |b - A x|_2 / |b|_2 in FP64 */
  double kern_seconds;   /* mean duration of the dominant kernel (time_kernel != 0) */
  double kern_bytes;     /* its algorithmic bytes per launch */
  long kern_launches;
  int outer;             /* defect-correction steps (DC methods) */
  int inner_iters[8];    /* iterations of the first 8 inner solves */
  double outer_res[8];   /* FP64 relative L1 residual before each of them */
  int breakdowns;        /* inner solves that broke down (DC methods) */
} fvk_solve_info;

#define FVK_SOLVE_PCG64 0 /* the reference PCG recurrence in FP64 (solv_cg.f90) */
#define FVK_SOLVE_DC32 1  /* FP32-inner Jacobi PCG in FP64 defect correction */
#define FVK_SOLVE_DC32_FACTOR 2 /* same, preconditioned by z = G G^T r with the
                                   factor of fvk_factor_set / fvk_factor_build;
                                   falls back to FVK_SOLVE_PCG64 (Jacobi, from
                                   x = 0, with a budget of max_iter of its
                                   own: a call does at most 2 max_iter
                                   iterations) when it does not converge
                                   within max_iter or breaks down again after
                                   "dc_max_breakdowns" restarts (default 1;
                                   p.q <= 0 or non-finite; a breakdown keeps
                                   the finite correction made so far and
                                   restarts the inner solve):
                                   FVK_FALLBACK */

/* a_f = s/l_pn * exp(sigma*z_f), z_f ~ N(0,1) from SplitMix64(seed) by face
   index; b ~ N(0,1) by cell, projected to zero mean.  b_out (reference
   order, may be NULL) receives the RHS. */
int fvk_sys_random(fvk_ctx *ctx, uint64_t seed, double sigma, double *b_out);
/* Symmetric system from host arrays: pl, a (= -ng%p = -ng%n per face), b. */
int fvk_sys_set(fvk_ctx *ctx, const double *pl, const double *a, const double *b);
/* Solve from x = 0 until the method's relative L1 residual is below tol (or
   max_iter iterations; see the fallback above).  imat selects the plain
   preconditioner (0 Jacobi, 1 DIC).  A non-finite right-hand side is
   FVK_ERR_ARG. */
int fvk_sys_solve(fvk_ctx *ctx, int method, int imat, double tol, int max_iter, int time_kernel,
                  fvk_solve_info *info);
int fvk_sys_get_x(fvk_ctx *ctx, double *x);
/* Solver settings of a context: "dc_replace_every" (default 160: every that
   many inner iterations the FP32 correction goes into the FP64 solution and
   the inner solve continues on the FP64 residual, keeping its search
   direction; each inner solve then aims at max(sqrt(1e-12/n),
   dc_eta_margin tol / current residual); 0: no replacement, and the floor
   is "dc_inner_reduc", default 1e-5), "dc_eta_margin" (default 0.8),
   "dc_max_outer" (defect-correction steps, default 30),
   "dc_max_breakdowns" (inner breakdowns followed by a restart from the
   current x before the solve gives up, default 1), "dc_carry" (1: a
   correction step continues the previous step's search direction; 0:
   restart, default).  DIC / DILU factor and sweeps (solv_precond,
   solv_direct): "sweep_one_launch" (1, default when no row has more than 8
   slots: one launch each, rows waiting for their values; 0: one launch per
   level), "sweep_parts" (1, default: each block owns ranges of
   "sweep_part_rows" rows consecutive in the reference numbering, default
   4096, 256..6400, and walks them level by level, values inside a range
   passing through LDS; 0: rows taken in level order in chunks by
   "sweep_blocks_per_cu" resident blocks per CU, default 1).  All give the
   same bits.  "sweep_fail" 1 (a test of the failure path): records that a
   sweep's wait ran past its bound, as a wait that never ends would; rows
   of the ranges form that wait give up within 256 passes, blocks stop
   taking ranges, and this and every later factorisation or sweep of the
   context returns FVK_ERR_HIP. */
int fvk_ctx_set_param(fvk_ctx *ctx, const char *name, double value);

/* The factorised preconditioner z = G G^T r, G = D^-1/2 (I + L) Delta, with
   D the diagonal of the current matrix, L strictly lower triangular on the
   matrix's pattern in the reference numbering (L[in, ip] = l_f for interior
   face f) and Delta_c = exp(sum of sig_f over the faces c owns).  No
   counterpart in the reference.  l, sig: n_fce entries in reference face
   order (boundary faces ignored). */
int fvk_factor_set(fvk_ctx *ctx, const float *l, const float *sig);
/* The factor of the current matrix by FSAI in closed form (Kolotilina and
   Yeremin's factorised sparse approximate inverse on A's own pattern:
   l_f = -Ah_f, Delta_c = (1 - sum of Ah_f^2 over the faces c owns)^-1/2,
   Ah = D^-1/2 A D^-1/2), exact when no two higher neighbours of a cell are
   adjacent (hexahedral and quadrilateral meshes).  The method the network
   learns, built without it; no counterpart in the reference. */
int fvk_factor_fsai(fvk_ctx *ctx);
/* z = G G^T r (FP32 inside; host arrays in reference order). */
int fvk_factor_apply(fvk_ctx *ctx, const double *r, double *z);
/* Per-face l_f, sig_f of the current factor (n_fce entries, 0 on boundary
   faces). */
int fvk_factor_get(fvk_ctx *ctx, float *l, float *sig);

/* ------------------------------------------------------------------------
 * The learned factor: a per-face MLP (12 -> 64 -> 64 -> 2, tanh; bfloat16
 * MFMA with float32 accumulation) maps local features of the Jacobi-scaled
 * matrix and the face geometry to (l_f, sig_f).  Trained on the GPU on the
 * systems the solver generates, by Adam on Kaporin's measure
 * log(tr(K)/N) - log(det K)/N of K = Gh^T Ah Gh (Ah = D^-1/2 A D^-1/2,
 * Gh = D^1/2 G), whose minimiser over the pattern is FSAI.  No counterpart
 * in the reference.
 *
 * Parameters (float32, fvk_model_nparam() = 5122): W1[12][64], b1[64],
 * W2[64][64], b2[64], W3[64][2], b3[2], input-major.  Features per interior
 * face f = (p, n): a/sqrt(a_pp a_nn), a/a_pp, a/a_nn, s/h^2, normal (3),
 * l_pn/h, w_fp, |dx_p|/h, |dx_n|/h, log(|V_p|/|V_n|), with a = -A_pn and
 * h = ((|V_p| + |V_n|)/2)^(1/3) (the reference's cell volumes can be
 * negative on warped cells).
 * ---------------------------------------------------------------------- */
typedef struct fvk_model fvk_model;

int fvk_model_nparam(void);
/* Hidden layers Xavier-uniform from SplitMix64(seed), output layer zero:
   the untrained factor is Jacobi. */
int fvk_model_create(int device, uint64_t seed, fvk_model **out);
void fvk_model_destroy(fvk_model *m);
int fvk_model_get(const fvk_model *m, float *theta);
/* The twelve features the network reads for each interior face of the
   context's current matrix, as it reads them (rounded to bfloat16),
   X[12 f + k] in reference face order, 0 on boundary faces.  A diagnostic
   of the inputs; no counterpart in the reference. */
int fvk_factor_features(fvk_ctx *ctx, float *X);
/* Sets the parameters and resets the optimiser state. */
int fvk_model_set(fvk_model *m, const float *theta);
/* Factor of the context's current matrix from the network (then
   FVK_SOLVE_DC32_FACTOR uses it). */
int fvk_factor_build(fvk_ctx *ctx, const fvk_model *m);
/* Loss J of the network on the context's current matrix; grad (5122, may be
   NULL): dJ/dtheta; dl, dsig (n_fce, may be NULL): dJ/dl_f, dJ/dsig_f. */
int fvk_model_loss(fvk_model *m, fvk_ctx *ctx, double *J, float *grad, float *dl, float *dsig);
/* steps Adam steps (beta 0.9, 0.999) of rate lr on the context's current
   matrix; loss (steps entries, may be NULL) receives J before each step. */
int fvk_model_train(fvk_model *m, fvk_ctx *ctx, int steps, double lr, double *loss);
/* The loaded system in reference order: pl (n_vol), a (n_fce, 0 on boundary
   faces), b (n_vol); NULL pointers skip. */
int fvk_sys_get(fvk_ctx *ctx, double *pl, double *a, double *b);
/* Mean duration of one FP64 SpMV launch over reps launches (HIP events). */
int fvk_bench_spmv(fvk_ctx *ctx, int reps, double *sec_per_launch);
/* Seconds per application of the DIC (imat 1) / DILU (imat 2) factor of
   the context's current matrix (forward and backward sweep) to nvc
   right-hand sides, as the solvers apply it.  No counterpart in the
   reference. */
int fvk_bench_sweep(fvk_ctx *ctx, int imat, int nvc, int reps, double *sec_per_apply);
/* Mean duration of a device copy of nbytes (kernel k_copy, 16-byte vector
   loads and stores): the achievable-bandwidth reference and the calibration
   kernel for hardware byte counters.  Allocates two scratch buffers. */
int fvk_bench_copy(int device, long nbytes, int reps, double *sec_per_launch);
/* Seconds per pass of a streaming probe over nbytes (kind 0: copy, nbytes
   read and nbytes written; 1: read only; 2: write only; every thread
   keeping four independent 16-byte loads in flight; 3: copy and 4: read
   with eight non-temporal 16-byte loads in flight and one resident grid):
   the bandwidth ceiling the kernels are measured against.  No counterpart
   in the reference. */
int fvk_bench_stream(int device, long nbytes, int reps, int kind, double *sec);

/* ------------------------------------------------------------------------
 * SIMPLE: momentum assembly + BiCGStab (cfd_setup_eqn,
 * spray-simulation/cfd_setup.f90:22-544) and pressure correction
 * (cfd_setup_pp, :549-760), plain single-phase steady branch.
 * ---------------------------------------------------------------------- */
typedef struct fvk_flow fvk_flow;

typedef struct fvk_flow_cfg {
  int simpler;      /* 1: SIMPLER pressure solve (shipped, cfd_setup.f90:581) */
  int imat_u;       /* momentum preconditioner (shipped 2 = DILU) */
  int imat_p;       /* pressure preconditioner (shipped 1 = DIC) */
  int n_bicg;       /* shipped 5 */
  int n_cg;         /* shipped 20 */
  int n_grad;       /* shipped 2 */
  int c_sch;        /* 0 CDS, 1 minmod, 2 van Leer, 3 superbee, >3 HRIC */
  double urf_u, urf_p, reduc_u, reduc_p, c_bl;
  double rho, mu;
  int p_solver;     /* pressure solves: 0 the reference PCG recurrence
                       (imat_p, reduc_p, n_cg; parity); to the FP64 relative
                       L1 residual p_tol from x = 0: 1 FP32-inner defect
                       correction (Jacobi), 2 the same with the learned
                       factor of the model set by fvk_flow_set_model (plain
                       FP64 PCG when it does not converge), 3 plain FP64
                       Jacobi PCG; solutions of singular (pure Neumann)
                       systems are shifted to mean zero */
  double p_tol;
} fvk_flow_cfg;

typedef struct fvk_outer_hist {
  double reduc;     /* 'reduc:' line, cfd_setup.f90:540 */
  double redpp;     /* 'redpp:' line, :672 */
  double redpp_sum; /* sum(rhs) on that line */
  double red_p;     /* 'red_p:' line, :751 (SIMPLER only) */
  double red_p_sum;
  double mx_res_n;  /* 'REDUC:' line, cfd_solve.f90:176-177 */
  int it_u[3], it_pp, it_p;
  double res0_n_u[3]; /* initial L1 norms of the momentum solve */
  double res0_n_pp;   /* ... of the p' solve */
  double res0_n_p;    /* ... of the SIMPLER pressure solve */
} fvk_outer_hist;

int fvk_flow_create(fvk_ctx *ctx, const fvk_flow_cfg *cfg, fvk_flow **out);
void fvk_flow_destroy(fvk_flow *f);
/* Replace the configuration (e.g. the urf argument of cfd_setup_pp). */
int fvk_flow_configure(fvk_flow *f, const fvk_flow_cfg *cfg);
/* The module variables c_in / c_out (cfd_data.f90:15): set = 1 copies the
   given values to the device, set = 0 reads them back. */
int fvk_flow_scalars(fvk_flow *f, int set, double *c_in, double *c_out);
/* Boundary velocity of every boundary face whose bnd_ty equals code. */
int fvk_flow_set_bnd_u(fvk_flow *f, int code, const double *u);
/* Full state in reference order: u_cell (n_vol*n_dx), u_bnd (n_bf*n_dx, by
   ib), p (n_vol), dpdx (n_vol*n_dx), spd (n_fce).  NULL pointers skip. */
int fvk_flow_set_state(fvk_flow *f, const double *u_cell, const double *u_bnd, const double *p,
                       const double *dpdx, const double *spd);
int fvk_flow_get_state(fvk_flow *f, double *u_cell, double *u_bnd, double *p, double *dpdx,
                       double *spd, double *cs, double *frc_dns);
/* Last assembled matrix (pl, ng%p, ng%n in reference order). */
int fvk_flow_get_matrix(fvk_flow *f, double *pl, double *ng_p, double *ng_n);
/* Right-hand side of the last solve of the first component (after
   cfd_setup_pp: the pressure-correction residual, the res0 the reference
   passes to solv_cg with not_r0), reference order.  With fvk_flow_get_matrix
   it harvests the pressure systems the flow generates. */
int fvk_flow_get_rhs(fvk_flow *f, double *rhs);
int fvk_flow_setup_eqn(fvk_flow *f, int it, fvk_outer_hist *h);
int fvk_flow_setup_pp(fvk_flow *f, int it, fvk_outer_hist *h);
/* One outer iteration (cfd_solve.f90:134-178, plain branch). */
int fvk_flow_outer(fvk_flow *f, int it, fvk_outer_hist *h);
/* n outer iterations it0 .. it0+n-1 with the state kept in HBM; h has n
   entries (may be NULL). */
int fvk_flow_run(fvk_flow *f, int it0, int n, fvk_outer_hist *h);
/* The same with iteration it0 run eagerly and the others as replays of HIP
   graphs of one outer iteration (one graph, or one per segment between
   pressure solves to tolerance, which check convergence from the host).
   Same results as fvk_flow_run. */
int fvk_flow_run_graph(fvk_flow *f, int it0, int n, fvk_outer_hist *h);
/* Model of the learned factor for p_solver 2 (not owned; NULL detaches);
   train_steps > 0: that many Adam steps of rate lr on every
   pressure-correction matrix before its factor is built. */
int fvk_flow_set_model(fvk_flow *f, fvk_model *m, int train_steps, double lr);

/* ------------------------------------------------------------------------
 * Output formats.
 * ---------------------------------------------------------------------- */
/* Tecplot FEBLOCK file of the flow's velocity, as eqn_outp writes it
   (spray-simulation/cfd_outp.f90:8-98): DIR/grd<igrd>%phs<iph>%eqn<ieq>.dat
   (with @<ict, 4 digits> when n_tl > 1 and ict > 0), vertex coordinates of
   the mesh, the n_dx velocity components per cell (cell-centred), cell
   vertex lists in Tecplot's brick / quadrilateral order (tecplot.vrt); the
   bytes of the flang-built reference's list-directed output.  m is the mesh
   the flow's context was built on. */
int fvk_eqn_outp(const fvk_mesh *m, fvk_flow *f, int igrd, int iph, int ieq, int ict, int n_tl,
                 const char *dir);
/* The same for the pressure, as prp_outp(..., p=.true.) writes it
   (cfd_outp.f90:102-189): DIR/grd<igrd>%phs<iph>%prp<ipr>.dat. */
int fvk_prp_outp_p(const fvk_mesh *m, fvk_flow *f, int igrd, int iph, int ipr, int ict, int n_tl,
                   const char *dir);
/* gmsh ElementData files of cell values, as flowlab writes them
   (flowlab/src/mesh.cpp:298-412): elm_idx[i] is the element number of
   value i; the vector form writes n_dim components padded to 3. */
int fvk_write_gmsh_field(const char *fname, int n, const int *elm_idx, const double *phi);
int fvk_write_gmsh_vector_field(const char *fname, int n, int n_dim, const int *elm_idx, const double *phi);
/* flang's list-directed editing of one REAL(8) value (what the Tecplot
   writers print), into buf (len bytes, NUL-terminated). */
int fvk_fmt_list_real(double x, char *buf, int len);

/* ------------------------------------------------------------------------
 * flowlab's serial CSR SIMPLE (flowlab/src/navierstokes.f90:515-641,
 * geometry_m.f90, transport_m.f90, solv_m.f90).  Arrays as flowlab's
 * Fortran receives them from its C++ mesh pipeline (navierstokes.f90:1-300),
 * one block: boundary elements 1..n_bnd, domain elements n_bnd+1 ..
 * n_bnd+n_dmn; all indices 1-based.  Fields are column-major (n_cmp, m) as
 * Fortran declares them.
 * ---------------------------------------------------------------------- */
typedef struct fvk_fl_mesh fvk_fl_mesh;
typedef struct fvk_fl_ctx fvk_fl_ctx;
typedef struct fvk_fl_flow fvk_fl_flow;

/* gmsh 2.2 file -> flowlab-layout arrays (the job of flowlab's read_gmsh_mesh,
   connectivity and mesh_to_dual, flowlab/src/main.cpp:146-480): elements of
   dimension n_dim are the domain, of n_dim - 1 the boundary (physical tag =
   boundary type, e.g. 11 inlet, 21 outlet, 31 wall); boundary faces first,
   interior faces in reverse order of discovery, each interior face with the
   vertex order of the element that met it first (its owner); the cell
   matrix pattern with its diagonal, columns ascending. */
int fvk_fl_mesh_read_gmsh(const char *path, int n_dim, double l_sc, fvk_fl_mesh **out);
/* The same from element records in memory: node coordinates x3 (3 x
   n_node), per element its gmsh type, physical tag and nodes (e_vptr 0-based
   offsets, n_elm+1; e_vrt node numbers 1..n_node, gmsh vertex order). */
int fvk_fl_mesh_from_records(int n_dim, int n_node, const double *x3, int n_elm, const int *e_ty, const int *e_tag,
                             const int *e_vptr, const int *e_vrt, fvk_fl_mesh **out);
/* info: n_dim, n_vrt, n_bnd, n_dmn, n_fce, nnz, element vertex entries,
   face vertex entries */
int fvk_fl_mesh_get_info(const fvk_fl_mesh *m, int *info);
/* x3 (3 x n_vrt), elm_vptr (n_dmn+1), elm_vrt, bnd_tag (n_bnd), fce_elm
   (2 x n_fce), fce_vptr (n_fce+1), fce_vrt, ng_ptr (n_dmn+1), ng_col (nnz),
   elm_idx (gmsh element number of each domain element); NULL skips */
int fvk_fl_mesh_get_arrays(const fvk_fl_mesh *m, double *x3, int *elm_vptr, int *elm_vrt, int *bnd_tag,
                           int *fce_elm, int *fce_vptr, int *fce_vrt, int *ng_ptr, int *ng_col, int *elm_idx);
void fvk_fl_mesh_free(fvk_fl_mesh *m);

/* Device context from flowlab's arrays: topology in HBM, then geometry
   (geometry, geometry_m.f90:97-330) and the inverted least-squares matrices
   (ls_grad_matrix, :592-660) on the GPU, bit for bit the reference's. */
int fvk_fl_ctx_create(int n_dim, int n_vrt, const double *x3, int n_bnd, const int *bnd_tag, int n_dmn,
                      const int *elm_vptr, const int *elm_vrt, int n_fce, const int *fce_elm, const int *fce_vptr,
                      const int *fce_vrt, const int *ng_ptr, const int *ng_col, int device, fvk_fl_ctx **out);
/* A context holding only a cell matrix pattern (ia n+1, ja nnz, 1-based
   CSR as solv_m receives them), for fvk_fl_lin_eqns_solver alone: the
   drop-in solv_m of flowlab, cfdpack and pcns (one source, byte-identical in
   the three).  fvk_fl_csr_matches: 1 when (n, ia, ja) is the context's
   pattern. */
int fvk_fl_csr_create(int n, const int *ia, const int *ja, int device, fvk_fl_ctx **out);
int fvk_fl_csr_matches(const fvk_fl_ctx *ctx, int n, const int *ia, const int *ja);
void fvk_fl_ctx_destroy(fvk_fl_ctx *ctx);
/* "sweep_one_launch" 1 (default when every row has at most 8 lower and 8
   upper entries): the factor and direct_solve's sweeps in one launch each,
   rows waiting for their values; 0: one launch per level.  "sweep_parts" 1
   (default with one launch): each block of the launch owns ranges of
   "sweep_part_rows" consecutive rows (default 4096, 256..6400) and walks them
   level by level, values inside a range passing through LDS; 0: rows taken
   in level order in chunks.  "sweep_blocks" (0: one per CU): blocks of the
   one-launch kernels.  All give the same bits.  "sweep_fail" 1 (a test of
   the failure path): as fvk_ctx_set_param's; the context's solves then
   return FVK_ERR_HIP.  No counterpart in the reference. */
int fvk_fl_ctx_set_param(fvk_fl_ctx *ctx, const char *name, double value);
/* xdmn%x_vc (n_dim x n_dmn), vol, x_fc, norm, area, x_pc, x_nc, l_pn, w_fp
   (per face), ls (n_dim x n_dim x n_dmn); NULL skips */
int fvk_fl_geometry(fvk_fl_ctx *ctx, double *x_vc, double *vol, double *x_fc, double *norm, double *area,
                    double *x_pc, double *x_nc, double *l_pn, double *w_fp, double *ls);
/* ls_gradient (geometry_m.f90:665-770): phi (n_cmp x n_dmn), phi_bnd
   (n_cmp x n_bnd; extrapolated in place when neu_bnd), grad (n_dim x n_cmp
   x n_dmn) */
int fvk_fl_ls_gradient(fvk_fl_ctx *ctx, int n_cmp, const double *phi, double *phi_bnd, double *grad, int neu_bnd);
/* lin_eqns_solver (solv_m.f90:10-30) on the context's pattern (its ia, ja):
   isol 1 cg_solver, 2 bicg_solver, with solv_m's preconditioner and
   direct_solve; rhs returns the initial residual as the reference leaves
   it, norm the final L1 residual */
int fvk_fl_lin_eqns_solver(fvk_fl_ctx *ctx, const double *a_ij, double *phi, double *rhs, double *norm, int n_iter,
                           double reduc, int isol);

/* SIMPLE state (velocity, pressure, pressure correction) on a context.
   params: density, viscosity, velocity relaxation, pressure relaxation,
   convection blending, momentum iterations and reduction, pressure
   iterations and reduction (navierstokes.f90:161-175, 566, 598: 25, 0.01). */
int fvk_fl_flow_create(fvk_fl_ctx *ctx, const double *params, fvk_fl_flow **out);
void fvk_fl_flow_destroy(fvk_fl_flow *f);
/* vel (n_dim x n_dmn), vel_bnd (n_dim x n_bnd; the values
   set_boundary_conditions sets), pres, pres_bnd; NULL keeps */
int fvk_fl_flow_set(fvk_fl_flow *f, const double *vel, const double *vel_bnd, const double *pres,
                    const double *pres_bnd);
int fvk_fl_flow_get(fvk_fl_flow *f, double *vel, double *vel_bnd, double *pres, double *pres_bnd, double *pp,
                    double *spd);
/* the current matrix values (nnz), right-hand side (n_dim x n_dmn), 1/a_P,
   velocity (n_dim x n_dim x n_dmn), pressure and pressure-correction
   gradients, pressure-correction boundary values */
int fvk_fl_flow_get_system(fvk_fl_flow *f, double *val, double *rhs, double *diag, double *vel_grad,
                           double *pres_grad, double *pp_grad, double *pp_bnd);
/* One step of the outer loop by the routine it replaces: 0 zero rhs and
   matrix, 1 ls_gradient(pres, 1), 2 ls_gradient(vel, 0), 3 ns_source
   (transport_m.f90:726-752), 4 construct_pde (:58-348), 5 momentum
   lin_eqns_solver calls (res0[0..n_dim)), 6 construct_pp (:430-629),
   7 pressure-correction lin_eqns_solver (res0[0]), 8 ls_gradient(pp, 1),
   9 ns_correction (:757-800) */
int fvk_fl_flow_step(fvk_fl_flow *f, int step, double *res0);
/* Outer iterations with fp_iter_ctrl's exit rule (transport_m.f90:806-845):
   res0 (4 per iteration), norm_gl (per iteration), *n_done */
int fvk_fl_flow_run(fvk_fl_flow *f, int n_iter, double res_drop, double *res0, double *norm_gl, int *n_done);

#ifdef __cplusplus
}
#endif
#endif
