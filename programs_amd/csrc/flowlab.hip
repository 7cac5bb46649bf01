// flowlab's serial CSR SIMPLE on gfx950: geometry (geometry_m.f90:97-330),
// least-squares gradients (ls_grad_matrix / ls_gradient, :592-730),
// momentum and pressure-correction assembly (transport_m.f90:58-348,
// :430-629, ns_source :726-752, ns_correction :757-800) and the CSR Krylov
// solvers with their factor (solv_m.f90:10-289), on one device context
// built from the arrays flowlab's Fortran receives (flmesh.h).
//
// Data layout: cells in the reference's order (domain-local index), faces
// in the reference's face order, structure-of-arrays.  A sum the reference
// builds by scattering over faces in face order is gathered per cell over
// the cell's faces in ascending face order: the same additions in the same
// order, so the geometry, the gradients and the assembled matrices and
// right-hand sides equal the reference's bit for bit (compiled without FMA
// contraction).  The Krylov reductions run in a fixed tree (rounding-level
// differences from the reference's sequential sums).
#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "ctx.h"
#include "flmesh.h"
#include "sweep.h"

namespace fvk {

// ---------------------------------------------------------------- context
struct FlKS {  // scalars of one component's Krylov solve (solv_m.f90 cof(1:6))
  double cof[7];
  double norm, norm0;
  int iter, done, pad[2];
};

struct FlCtx {
  int device = 0;
  hipStream_t stream = nullptr;
  int nd = 0, n = 0, nb = 0, nf = 0, nnz = 0, n_vrt = 0;
  std::vector<void *> owned;
  std::vector<int> h_ia, h_ja, h_j1, h_jn, h_btag;
  // geometry
  double *xvc = nullptr, *vol = nullptr, *volr = nullptr, *ls = nullptr;
  double *xfc = nullptr, *nrm = nullptr, *area = nullptr, *xpc = nullptr, *xnc = nullptr, *lpn = nullptr,
         *wfp = nullptr, *fvol = nullptr;
  int *j1 = nullptr, *jn = nullptr, *btag = nullptr;
  // matrix pattern
  int *ia = nullptr, *ja = nullptr, *kii = nullptr, *kij = nullptr, *kji = nullptr, *tpos = nullptr;
  // a row's lower entries are [ia, klo) (the reference's loops stop at the
  // first column >= the row), its upper entries [kup, ia+1) (taken from the
  // end down to the first column <= the row)
  int *klo = nullptr, *kup = nullptr;
  int *cptr = nullptr, *cfl = nullptr;
  // triangular sweeps by levels (lower / upper dependencies)
  std::vector<int> lf_ptr, lb_ptr;
  int *lf_rows = nullptr, *lb_rows = nullptr;
  // one launch per factorisation / sweep (rows wait for their values) when
  // no row has more than kSweepSlots lower or upper entries
  bool one_launch = true;
  int n_cu = 256;
  int sweep_blocks = 0;  // blocks of the one-launch kernels (0: one per CU)
  // partitioned one-launch sweeps (sweep_parts): forward / backward level
  // order of each row range, for ranges of R1 rows (one component, the
  // factor) and R3 rows (two or three components)
  struct Parts {
    int P = 0, R = 0, W = 0, wmax = 0;  // wmax: the most rows of one level of a range
    int *hd = nullptr, *jc = nullptr, *lc = nullptr;  // descriptors (sweep_pack), 16-byte words
    int *step = nullptr, *sp = nullptr;
    SweepTabs tabs() const { return SweepTabs{(const int4 *)hd, (const int4 *)jc, (const int4 *)lc}; }
  };
  Parts pf, pb;
  bool parts = false;
  int part_rows = 4096;  // R: LDS of one block = components x R doubles
  std::vector<int> h_levf, h_levb, h_kii, h_klo, h_kup;
  int *sweep_int = nullptr;  // tickets: forward, backward, factor; [3] wait-bound error
  // Krylov work: 6 vectors x 3 components, partials, scalars
  double *wk = nullptr, *pc = nullptr, *part = nullptr;
  FlKS *ks = nullptr;
  double *scal = nullptr;  // c_rat and scratch scalars
};

// zeroed allocation; the clear has finished when this returns (an upload
// that follows on another stream must not race it)
template <class T>
static int fl_alloc(FlCtx &c, T **p, size_t n) {
  void *q = nullptr;
  FVK_HIP(hipMalloc(&q, (n ? n : 1) * sizeof(T)));
  c.owned.push_back(q);
  FVK_HIP(hipMemsetAsync(q, 0, (n ? n : 1) * sizeof(T), c.stream));
  FVK_HIP(hipStreamSynchronize(c.stream));
  *p = (T *)q;
  return 0;
}
template <class T>
static int fl_up(FlCtx &c, T **p, const std::vector<T> &v) {
  int rc = fl_alloc(c, p, v.size());
  if (rc) return rc;
  if (!v.empty()) FVK_HIP(hipMemcpy(*p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice));
  return 0;
}

static void fl_free(FlCtx &c, void *p) {
  if (!p) return;
  for (size_t k = 0; k < c.owned.size(); ++k)
    if (c.owned[k] == p) {
      hipFree(p);
      c.owned.erase(c.owned.begin() + (long)k);
      return;
    }
}

static double *wkv(FlCtx &c, int k) { return c.wk + (size_t)k * 3 * c.n; }

// ---------------------------------------------------------------- geometry
// cross_prod (vector_utils_m.f90:35-41)
__device__ __forceinline__ void fl_cross(const double *a, const double *b, double *v) {
  v[0] = a[1] * b[2] - b[1] * a[2];
  v[1] = a[2] * b[0] - b[2] * a[0];
  v[2] = a[0] * b[1] - b[0] * a[1];
}

__global__ __launch_bounds__(BS) void k_fl_centre(int n, int nd, const double *__restrict__ xv,
                                                  const int *__restrict__ evp, const int *__restrict__ evr,
                                                  double *__restrict__ xvc) {
  const int i = blockIdx.x * BS + threadIdx.x;
  if (i >= n) return;
  double x[3] = {0.0, 0.0, 0.0};
  const int a = evp[i], b = evp[i + 1];
  for (int k = a; k < b; ++k)
    for (int d = 0; d < nd; ++d) x[d] = x[d] + xv[3 * (size_t)evr[k] + d];
  const double cnt = (double)(b - a);
  for (int d = 0; d < nd; ++d) xvc[(size_t)d * n + i] = x[d] / cnt;
}

// geometry_m.f90:140-300 for one face
__global__ __launch_bounds__(BS) void k_fl_face_geom(int nf, int n, int nd, const double *__restrict__ xv,
                                                     const int *__restrict__ fvp, const int *__restrict__ fvr,
                                                     const int *__restrict__ j1, const int *__restrict__ jn,
                                                     const double *__restrict__ xvc, double *xfc, double *nrm,
                                                     double *area, double *xpc, double *xnc, double *lpn,
                                                     double *wfp, double *fvol) {
  const int f = blockIdx.x * BS + threadIdx.x;
  if (f >= nf) return;
  const int a = fvp[f], b = fvp[f + 1], nv = b - a;
  double x1[3] = {0.0, 0.0, 0.0};
  for (int k = a; k < b; ++k)
    for (int d = 0; d < nd; ++d) x1[d] = x1[d] + xv[3 * (size_t)fvr[k] + d];
  double xf[3] = {0.0, 0.0, 0.0};
  for (int d = 0; d < nd; ++d) xf[d] = x1[d] / (double)nv;
  // face normal and area
  const double *v0 = xv + 3 * (size_t)fvr[a];
  double s[3] = {0.0, 0.0, 0.0}, e1[3] = {0.0, 0.0, 0.0};
  for (int k = a + 1; k < b; ++k) {
    double e[3] = {0.0, 0.0, 0.0};
    const double *vk = xv + 3 * (size_t)fvr[k];
    for (int d = 0; d < nd; ++d) e[d] = vk[d] - v0[d];
    if (nv == 2) {
      const double uz[3] = {0.0, 0.0, 1.0};
      double cp[3];
      fl_cross(e, uz, cp);
      for (int d = 0; d < nd; ++d) s[d] = cp[d];
    } else {
      if (k - (a + 1) >= 1) {
        double cp[3];
        fl_cross(e1, e, cp);
        for (int d = 0; d < 3; ++d) s[d] = s[d] + 0.5 * cp[d];
      }
      for (int d = 0; d < 3; ++d) e1[d] = e[d];
    }
  }
  double ss = 0.0;
  for (int d = 0; d < nd; ++d) ss = ss + s[d] * s[d];
  const double ar = sqrt(ss);
  double nv3[3] = {0.0, 0.0, 0.0};
  for (int d = 0; d < nd; ++d) nv3[d] = s[d] / ar;
  area[f] = ar;
  for (int d = 0; d < nd; ++d) {
    xfc[(size_t)d * nf + f] = xf[d];
    nrm[(size_t)d * nf + f] = nv3[d];
  }
  fvol[f] = xf[0] * nv3[0] * ar;
  const int p = j1[f];
  double dx[3], t = 0.0;
  for (int d = 0; d < nd; ++d) dx[d] = xf[d] - xvc[(size_t)d * n + p];
  for (int d = 0; d < nd; ++d) t = t + dx[d] * nv3[d];
  double xp[3];
  for (int d = 0; d < nd; ++d) {
    xp[d] = xf[d] - nv3[d] * t;
    xpc[(size_t)d * nf + f] = xp[d];
  }
  const int q = jn[f];
  if (q >= 0) {
    double u = 0.0, xn[3];
    for (int d = 0; d < nd; ++d) dx[d] = xf[d] - xvc[(size_t)d * n + q];
    for (int d = 0; d < nd; ++d) u = u + dx[d] * nv3[d];
    for (int d = 0; d < nd; ++d) {
      xn[d] = xf[d] - nv3[d] * u;
      xnc[(size_t)d * nf + f] = xn[d];
    }
    double l = 0.0, w = 0.0;
    for (int d = 0; d < nd; ++d) l = l + (xn[d] - xp[d]) * (xn[d] - xp[d]);
    const double l_pn = sqrt(l);
    for (int d = 0; d < nd; ++d) w = w + (xf[d] - xp[d]) * (xf[d] - xp[d]);
    wfp[f] = sqrt(w) / l_pn;
    lpn[f] = l_pn;
  } else {
    double l = 0.0;
    for (int d = 0; d < nd; ++d) l = l + (xf[d] - xp[d]) * (xf[d] - xp[d]);
    lpn[f] = sqrt(l);
  }
}

// volume: the face scatter of geometry_m.f90:212-232 gathered per cell
__global__ __launch_bounds__(BS) void k_fl_vol(int n, const int *__restrict__ cptr, const int *__restrict__ cfl,
                                               const int *__restrict__ j1, const double *__restrict__ fvol,
                                               double *vol, double *volr) {
  const int i = xcd_block() * BS + threadIdx.x;
  if (i >= n) return;
  double v = 0.0;
  for (int k = cptr[i]; k < cptr[i + 1]; ++k) {
    const int f = cfl[k];
    v = j1[f] == i ? v + fvol[f] : v - fvol[f];
  }
  vol[i] = v;
  volr[i] = 1.0 / (v + SMALL);
}

// matrix_inv (matrix_utils_m.f90:8-52) of a column-major nd x nd matrix
__device__ __forceinline__ void fl_inv(int nd, const double *m, double *inv) {
#define M(i, j) m[((j) - 1) * nd + (i) - 1]
#define I(i, j) inv[((j) - 1) * nd + (i) - 1]
  double det;
  if (nd == 2) {
    det = M(1, 1) * M(2, 2) - M(1, 2) * M(2, 1);
  } else {
    det = (M(2, 2) * M(3, 3) - M(2, 3) * M(3, 2)) * M(1, 1);
    det = det + (M(2, 3) * M(3, 1) - M(2, 1) * M(3, 3)) * M(1, 2);
    det = det + (M(2, 1) * M(3, 2) - M(2, 2) * M(3, 1)) * M(1, 3);
  }
  if (fabs(det) < SMALL) {
    for (int k = 0; k < nd * nd; ++k) inv[k] = 0.0;
    return;
  }
  if (nd == 2) {
    I(1, 1) = M(2, 2);
    I(1, 2) = -M(1, 2);
    I(2, 1) = -M(2, 1);
    I(2, 2) = M(1, 1);
  } else {
    I(1, 1) = M(2, 2) * M(3, 3) - M(2, 3) * M(3, 2);
    I(2, 1) = M(2, 3) * M(3, 1) - M(2, 1) * M(3, 3);
    I(3, 1) = M(2, 1) * M(3, 2) - M(2, 2) * M(3, 1);
    I(1, 2) = M(3, 2) * M(1, 3) - M(3, 3) * M(1, 2);
    I(2, 2) = M(3, 3) * M(1, 1) - M(3, 1) * M(1, 3);
    I(3, 2) = M(3, 1) * M(1, 2) - M(3, 2) * M(1, 1);
    I(1, 3) = M(1, 2) * M(2, 3) - M(1, 3) * M(2, 2);
    I(2, 3) = M(1, 3) * M(2, 1) - M(1, 1) * M(2, 3);
    I(3, 3) = M(1, 1) * M(2, 2) - M(1, 2) * M(2, 1);
  }
  for (int k = 0; k < nd * nd; ++k) inv[k] = inv[k] / det;
#undef M
#undef I
}

// ls_grad_matrix (geometry_m.f90:592-660): sum of dx dx^T over the cell's
// faces (ascending), inverted; ls[(jc*nd + ic)*n + i] = inverse(ic, jc)
__global__ __launch_bounds__(BS) void k_fl_ls(int n, int nf, int nd, const int *__restrict__ cptr,
                                              const int *__restrict__ cfl, const int *__restrict__ j1,
                                              const int *__restrict__ jn, const double *__restrict__ xvc,
                                              const double *__restrict__ xfc, double *ls) {
  const int i = xcd_block() * BS + threadIdx.x;
  if (i >= n) return;
  double m[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  for (int k = cptr[i]; k < cptr[i + 1]; ++k) {
    const int f = cfl[k], p = j1[f], q = jn[f];
    double dx[3];
    for (int d = 0; d < nd; ++d)
      dx[d] = q >= 0 ? xvc[(size_t)d * n + q] - xvc[(size_t)d * n + p] : xfc[(size_t)d * nf + f] - xvc[(size_t)d * n + p];
    for (int jc = 0; jc < nd; ++jc)
      for (int ic = 0; ic < nd; ++ic) m[jc * nd + ic] = m[jc * nd + ic] + dx[ic] * dx[jc];
  }
  double inv[9];
  fl_inv(nd, m, inv);
  for (int k = 0; k < nd * nd; ++k) ls[(size_t)k * n + i] = inv[k];
}

// ---------------------------------------------------------------- gradients
// ls_gradient (geometry_m.f90:665-770): g(:, q) = LS * sum over faces of
// dx dphi(q); neu: boundary differences are zero and the boundary values
// are extrapolated afterwards (k_fl_grad_bnd)
__global__ __launch_bounds__(BS) void k_fl_grad(int n, int nf, int nb, int nd, int ncmp, const int *__restrict__ cptr,
                                                const int *__restrict__ cfl, const int *__restrict__ j1,
                                                const int *__restrict__ jn, const double *__restrict__ xvc,
                                                const double *__restrict__ xfc, const double *__restrict__ ls,
                                                const double *__restrict__ phi, const double *__restrict__ phib,
                                                int neu, double *__restrict__ g) {
  const int i = xcd_block() * BS + threadIdx.x;
  if (i >= n) return;
  double a[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};  // a[q][d]
  for (int k = cptr[i]; k < cptr[i + 1]; ++k) {
    const int f = cfl[k], p = j1[f], q = jn[f];
    double dx[3], dp[3];
    if (q >= 0) {
      for (int d = 0; d < nd; ++d) dx[d] = xvc[(size_t)d * n + q] - xvc[(size_t)d * n + p];
      for (int c = 0; c < ncmp; ++c) dp[c] = phi[(size_t)c * n + q] - phi[(size_t)c * n + p];
    } else {
      const int b = -q - 1;
      for (int d = 0; d < nd; ++d) dx[d] = xfc[(size_t)d * nf + f] - xvc[(size_t)d * n + p];
      for (int c = 0; c < ncmp; ++c) dp[c] = neu ? 0.0 : phib[(size_t)c * nb + b] - phi[(size_t)c * n + p];
    }
    for (int c = 0; c < ncmp; ++c)
      for (int d = 0; d < nd; ++d) a[c][d] = a[c][d] + dx[d] * dp[c];
  }
  double L[9];
  for (int k = 0; k < nd * nd; ++k) L[k] = ls[(size_t)k * n + i];
  for (int c = 0; c < ncmp; ++c)
    for (int r = 0; r < nd; ++r) {
      double s = 0.0;
      for (int j = 0; j < nd; ++j) s = s + L[j * nd + r] * a[c][j];
      g[(size_t)(c * nd + r) * n + i] = s;
    }
}

__global__ __launch_bounds__(BS) void k_fl_grad_bnd(int nbf, const int *__restrict__ bfl, int n, int nf, int nb,
                                                    int nd, int ncmp, const int *__restrict__ j1,
                                                    const int *__restrict__ jn, const double *__restrict__ xvc,
                                                    const double *__restrict__ xfc, const double *__restrict__ phi,
                                                    const double *__restrict__ g, double *phib) {
  const int t = blockIdx.x * BS + threadIdx.x;
  if (t >= nbf) return;
  const int f = bfl[t], p = j1[f], b = -jn[f] - 1;
  double dx[3];
  for (int d = 0; d < nd; ++d) dx[d] = xfc[(size_t)d * nf + f] - xvc[(size_t)d * n + p];
  for (int c = 0; c < ncmp; ++c) {
    double s = 0.0;
    for (int d = 0; d < nd; ++d) s = s + g[(size_t)(c * nd + d) * n + p] * dx[d];
    phib[(size_t)c * nb + b] = phi[(size_t)c * n + p] + s;
  }
}

// ---------------------------------------------------------------- Krylov
// csr_matvec_prod (matrix_utils_m.f90:70-82): y(i) = 0 + sum a x in CSR order
__device__ __forceinline__ double fl_row_mv(const int *ia, const int *ja, const double *a, const double *x, int i) {
  double y = 0.0;
  for (int k = ia[i]; k < ia[i + 1]; ++k) y = y + a[k] * x[ja[k]];
  return y;
}

// preconditioner (solv_m.f90:224-250) over one level of rows
__global__ __launch_bounds__(BS) void k_fl_pc_level(const int *__restrict__ rows, int cnt, const int *__restrict__ ia,
                                                    const int *__restrict__ ja, const int *__restrict__ kii,
                                                    const int *__restrict__ tpos, const double *__restrict__ a,
                                                    double *pc) {
  const int t = blockIdx.x * BS + threadIdx.x;
  if (t >= cnt) return;
  const int i = rows[t];
  double s = a[kii[i]];
  for (int k = ia[i]; k < ia[i + 1]; ++k) {
    const int j = ja[k];
    if (j >= i) break;
    s = s - a[k] * pc[j] * a[tpos[k]];
  }
  pc[i] = 1.0 / (s + SMALL);
}

// direct_solve forward sweep (solv_m.f90:255-268) for ncmp right-hand sides
__global__ __launch_bounds__(BS) void k_fl_fwd_level(const int *__restrict__ rows, int cnt, int n, int ncmp,
                                                     const int *__restrict__ ia, const int *__restrict__ ja,
                                                     const double *__restrict__ a, const double *__restrict__ pc,
                                                     const double *__restrict__ b, double *x,
                                                     const FlKS *__restrict__ ks) {
  const int t = blockIdx.x * BS + threadIdx.x;
  if (t >= cnt) return;
  const int i = rows[t];
  for (int c = 0; c < ncmp; ++c) {
    if (ks && ks[c].done) continue;
    double s = b[(size_t)c * n + i];
    for (int k = ia[i]; k < ia[i + 1]; ++k) {
      const int j = ja[k];
      if (j >= i) break;
      s = s - a[k] * x[(size_t)c * n + j];
    }
    x[(size_t)c * n + i] = s * pc[i];
  }
}

// direct_solve scaling and backward sweep (solv_m.f90:270-286): t holds the
// forward result, y receives x
__global__ __launch_bounds__(BS) void k_fl_bwd_level(const int *__restrict__ rows, int cnt, int n, int ncmp,
                                                     const int *__restrict__ ia, const int *__restrict__ ja,
                                                     const double *__restrict__ a, const double *__restrict__ pc,
                                                     const double *__restrict__ tf, double *y,
                                                     const FlKS *__restrict__ ks) {
  const int t = blockIdx.x * BS + threadIdx.x;
  if (t >= cnt) return;
  const int i = rows[t];
  const double cof = 1.0 / (pc[i] + SMALL);
  for (int c = 0; c < ncmp; ++c) {
    if (ks && ks[c].done) continue;
    double s = cof * tf[(size_t)c * n + i];
    for (int k = ia[i + 1] - 1; k >= ia[i]; --k) {
      const int j = ja[k];
      if (j <= i) break;
      s = s - a[k] * y[(size_t)c * n + j];
    }
    y[(size_t)c * n + i] = s * pc[i];
  }
}

// Block partial sums of NV per-row terms per component, stored at
// part[(c*NV + j)*RG + block]
template <int NV, class F>
__device__ __forceinline__ void fl_partials(int n, int ncmp, const FlKS *ks, double *part, F term) {
  int r0, r1;
  block_rows<RB>(n, r0, r1);
  for (int c = 0; c < 3; ++c) {
    if (c >= ncmp) break;
    double v[NV];
#pragma unroll
    for (int j = 0; j < NV; ++j) v[j] = 0.0;
    if (!ks[c].done)
      for (int i = r0 + (int)threadIdx.x; i < r1; i += RB) term(c, i, v);
    rblock_sum<NV, RB>(v);
    if (threadIdx.x == 0)
#pragma unroll
      for (int j = 0; j < NV; ++j) part[(size_t)(c * NV + j) * RG + blockIdx.x] = v[j];
  }
}

// initial residual: wk1 = A phi; res = res - wk1; wk1 = res; partial |res|
__global__ __launch_bounds__(RB) void k_fl_init(int n, int ncmp, const int *__restrict__ ia, const int *__restrict__ ja,
                                                const double *__restrict__ a, const double *__restrict__ phi,
                                                double *res, double *w1, const FlKS *ks, double *part) {
  fl_partials<1>(n, ncmp, ks, part, [&](int c, int i, double *v) {
    const double y = fl_row_mv(ia, ja, a, phi + (size_t)c * n, i);
    const double r = res[(size_t)c * n + i] - y;
    res[(size_t)c * n + i] = r;
    w1[(size_t)c * n + i] = r;
    v[0] += fabs(r);
  });
}

// s = sum x*y (two vectors)
__global__ __launch_bounds__(RB) void k_fl_dot(int n, int ncmp, const double *__restrict__ x,
                                               const double *__restrict__ y, const FlKS *ks, double *part) {
  fl_partials<1>(n, ncmp, ks, part, [&](int c, int i, double *v) {
    v[0] += x[(size_t)c * n + i] * y[(size_t)c * n + i];
  });
}

// q = A p and sum p*q (CG)
__global__ __launch_bounds__(RB) void k_fl_mv_dot(int n, int ncmp, const int *__restrict__ ia,
                                                  const int *__restrict__ ja, const double *__restrict__ a,
                                                  const double *__restrict__ p, double *q, const double *__restrict__ d,
                                                  const FlKS *ks, double *part) {
  fl_partials<1>(n, ncmp, ks, part, [&](int c, int i, double *v) {
    const double y = fl_row_mv(ia, ja, a, p + (size_t)c * n, i);
    q[(size_t)c * n + i] = y;
    v[0] += d[(size_t)c * n + i] * y;
  });
}

// v = A z, sum v*r and sum (v+small)^2 (BiCGStab, solv_m.f90:190-196)
__global__ __launch_bounds__(RB) void k_fl_mv_dot2(int n, int ncmp, const int *__restrict__ ia,
                                                   const int *__restrict__ ja, const double *__restrict__ a,
                                                   const double *__restrict__ z, double *vv,
                                                   const double *__restrict__ r, const FlKS *ks, double *part) {
  fl_partials<2>(n, ncmp, ks, part, [&](int c, int i, double *v) {
    const double y = fl_row_mv(ia, ja, a, z + (size_t)c * n, i);
    vv[(size_t)c * n + i] = y;
    v[0] += y * r[(size_t)c * n + i];
    v[1] += (y + SMALL) * (y + SMALL);
  });
}

// phi += alpha p; r -= alpha q; partial |r| (alpha = cof[ai])
__global__ __launch_bounds__(RB) void k_fl_upd(int n, int ncmp, int ai, double *phi, const double *__restrict__ p,
                                               double *r, const double *__restrict__ q, const FlKS *ks, double *part,
                                               int with_norm) {
  fl_partials<1>(n, ncmp, ks, part, [&](int c, int i, double *v) {
    const double al = ks[c].cof[ai];
    phi[(size_t)c * n + i] = phi[(size_t)c * n + i] + al * p[(size_t)c * n + i];
    const double rr = r[(size_t)c * n + i] - al * q[(size_t)c * n + i];
    r[(size_t)c * n + i] = rr;
    if (with_norm) v[0] += fabs(rr);
  });
}

// p = z + beta p (CG) or p = r + omega (p - alpha u) (BiCGStab)
__global__ __launch_bounds__(BS) void k_fl_dir(int n, int ncmp, int bicg, double *p, const double *__restrict__ z,
                                               const double *__restrict__ u, const FlKS *ks) {
  const int i = blockIdx.x * BS + threadIdx.x;
  if (i >= n) return;
  for (int c = 0; c < ncmp; ++c) {
    if (ks[c].done) continue;
    const size_t o = (size_t)c * n + i;
    if (bicg) p[o] = z[o] + ks[c].cof[6] * (p[o] - ks[c].cof[2] * u[o]);
    else p[o] = z[o] + ks[c].cof[3] * p[o];
  }
}

__device__ __forceinline__ double fl_total(const double *part, int P, int comp) {
  double v[1];
  v[0] = 0.0;
  for (int b = threadIdx.x; b < P; b += RB) v[0] += part[(size_t)comp * RG + b];
  rblock_sum<1, RB>(v);
  return v[0];
}

// scalar steps of cg_solver / bicg_solver; one block, thread 0 writes.
// stage: 0 initial norm, 10 CG s/beta, 11 CG alpha, 12 CG norm/exit,
// 20 BiCG beta/omega, 21 BiCG gamma, 22 BiCG alpha, 23 BiCG norm/exit
__global__ __launch_bounds__(RB) void k_fl_scal(int stage, int ncmp, int P, int NV, const double *part, FlKS *ks,
                                                double reduc, int bicg) {
  for (int c = 0; c < ncmp; ++c) {
    const double t0 = fl_total(part, P, c * NV);
    const double t1 = NV > 1 ? fl_total(part, P, c * NV + 1) : 0.0;
    if (threadIdx.x != 0) continue;
    FlKS &k = ks[c];
    if (stage == 0) {
      for (int j = 0; j < 7; ++j) k.cof[j] = 0.0;
      k.norm = t0;
      k.norm0 = t0;
      k.iter = 0;
      k.done = 0;
      if (bicg) {
        k.cof[2] = 1.0;
        k.cof[3] = 1.0;
        k.cof[5] = 1.0;
      } else {
        k.cof[4] = 1.0 / SMALL;
      }
      continue;
    }
    if (k.done) continue;
    switch (stage) {
      case 10: {
        k.cof[1] = 0.0;
        k.cof[3] = 1.0 / (k.cof[4] + SMALL);
        k.cof[4] = t0;
        k.cof[3] = k.cof[3] * k.cof[4];
        break;
      }
      case 11: k.cof[2] = k.cof[4] / (t0 + SMALL); break;
      case 12:
      case 23: {
        k.norm = t0;
        k.cof[1] = t0 / (k.norm0 + SMALL);
        k.iter += 1;
        if (k.cof[1] < reduc) k.done = 1;
        break;
      }
      case 20: {
        k.cof[1] = 0.0;
        k.cof[4] = t0;
        k.cof[6] = k.cof[4] * k.cof[5] / (k.cof[2] * k.cof[3] + SMALL);
        k.cof[3] = k.cof[4];
        break;
      }
      case 21: k.cof[5] = k.cof[4] / (t0 + SMALL); break;
      case 22: {
        k.cof[1] = t0;
        k.cof[2] = t1;
        k.cof[2] = k.cof[1] / (k.cof[2] + SMALL);
        break;
      }
    }
  }
}

// a new solve: every component open (the initial residual kernel is gated
// like the others)
__global__ void k_fl_ks_open(FlKS *ks, int ncmp) {
  if (threadIdx.x == 0)
    for (int c = 0; c < ncmp; ++c) ks[c].done = 0;
}

static int fl_pgrid(int n) { return red_grid(n); }

// ---------------------------------------------------------------- one-launch sweeps
__device__ __forceinline__ bool fl_all_done(const FlKS *ks, int ncmp) {
  if (!ks) return false;
  for (int c = 0; c < ncmp; ++c)
    if (!ks[c].done) return false;
  return true;
}

__global__ __launch_bounds__(BS) void k_fl_mark(size_t cnt, double *out, int *ticket, const FlKS *ks, int ncmp) {
  if (fl_all_done(ks, ncmp)) return;
  unsigned long long *o = (unsigned long long *)out;
  for (size_t i = blockIdx.x * (size_t)BS + threadIdx.x; i < cnt; i += (size_t)gridDim.x * BS) o[i] = kSweepMark;
  if (blockIdx.x == 0 && threadIdx.x == 0) *ticket = 0;
}

// preconditioner (solv_m.f90:224-250) for one row: the lower entries in
// column order
struct FlPcRow {
  const int *ia, *ja, *kii, *klo, *tpos, *order;
  const double *a;
  double *pc;
  int r, k, nl;
  double s;
  int j[kSweepSlots];
  double a1[kSweepSlots], a2[kSweepSlots];
  __device__ void start(int pos) {
    r = order[pos];
    const int lo = ia[r];
    nl = klo[r] - lo;
    s = a[kii[r]];
    k = 0;
#pragma unroll
    for (int t = 0; t < kSweepSlots; ++t) {
      j[t] = 0;
      a1[t] = a2[t] = 0.0;
      if (t < nl) {
        j[t] = ja[lo + t];
        a1[t] = a[lo + t];
        a2[t] = a[tpos[lo + t]];
      }
    }
  }
  __device__ bool advance() {
    unsigned long long v[kSweepSlots];
#pragma unroll
    for (int t = 0; t < kSweepSlots; ++t)
      if (t >= k && t < nl) v[t] = ld_agent(pc + j[t]);
#pragma unroll
    for (int t = 0; t < kSweepSlots; ++t) {
      if (t < k || t >= nl) continue;
      if (v[t] == kSweepMark) break;
      s = s - a1[t] * __longlong_as_double((long long)v[t]) * a2[t];
      k = t + 1;
    }
    if (k < nl) return false;
    st_agent(pc + r, 1.0 / (s + SMALL));
    return true;
  }
};

// direct_solve for NC right-hand sides: forward (solv_m.f90:255-268) writes
// out from src = b; backward (:270-286) scales src = the forward result and
// writes out from the upper entries, last column first
template <bool FWD, int NC>
struct FlSweepRow {
  const int *ia, *ja, *klo, *kup, *order;
  const double *a, *pc, *src;
  double *out;
  int n, r, k, nl;
  double s[NC];
  int j[kSweepSlots];
  double av[kSweepSlots];
  __device__ void start(int pos) {
    r = order[pos];
    int first, step;
    if (FWD) {
      first = ia[r];
      nl = klo[r] - first;
      step = 1;
    } else {
      first = ia[r + 1] - 1;
      nl = first - kup[r] + 1;
      step = -1;
    }
    k = 0;
#pragma unroll
    for (int t = 0; t < kSweepSlots; ++t) {
      j[t] = 0;
      av[t] = 0.0;
      if (t < nl) {
        j[t] = ja[first + step * t];
        av[t] = a[first + step * t];
      }
    }
    if (FWD) {
#pragma unroll
      for (int q = 0; q < NC; ++q) s[q] = src[(size_t)q * n + r];
    } else {
      const double cof = 1.0 / (pc[r] + SMALL);
#pragma unroll
      for (int q = 0; q < NC; ++q) s[q] = cof * src[(size_t)q * n + r];
    }
  }
  __device__ bool advance() {
    unsigned long long v[kSweepSlots][NC];
#pragma unroll
    for (int t = 0; t < kSweepSlots; ++t)
      if (t >= k && t < nl)
#pragma unroll
        for (int q = 0; q < NC; ++q) v[t][q] = ld_agent(out + (size_t)q * n + j[t]);
#pragma unroll
    for (int t = 0; t < kSweepSlots; ++t) {
      if (t < k || t >= nl) continue;
      bool ready = true;
#pragma unroll
      for (int q = 0; q < NC; ++q) ready = ready && v[t][q] != kSweepMark;
      if (!ready) break;
#pragma unroll
      for (int q = 0; q < NC; ++q) s[q] = s[q] - av[t] * __longlong_as_double((long long)v[t][q]);
      k = t + 1;
    }
    if (k < nl) return false;
    const double cc = pc[r];
#pragma unroll
    for (int q = 0; q < NC; ++q) st_agent(out + (size_t)q * n + r, s[q] * cc);
    return true;
  }
};

__global__ __launch_bounds__(BS) void k_fl_pc_sf(int n, const int *__restrict__ order, const int *__restrict__ ia,
                                                 const int *__restrict__ ja, const int *__restrict__ kii,
                                                 const int *__restrict__ klo, const int *__restrict__ tpos,
                                                 const double *__restrict__ a, double *pc, int *ticket, int *err) {
  FlPcRow row;
  row.ia = ia; row.ja = ja; row.kii = kii; row.klo = klo; row.tpos = tpos; row.order = order; row.a = a; row.pc = pc;
  row.r = 0; row.k = 0; row.nl = 0; row.s = 0.0;
  sweep_chunks(n, ticket, err, row);
}

template <bool FWD, int NC>
__global__ __launch_bounds__(BS) void k_fl_sweep_sf(int n, const int *__restrict__ order, const int *__restrict__ ia,
                                                    const int *__restrict__ ja, const int *__restrict__ klo,
                                                    const int *__restrict__ kup, const double *__restrict__ a,
                                                    const double *__restrict__ pc, const double *__restrict__ src,
                                                    double *out, int *ticket, int *err, const FlKS *ks) {
  if (fl_all_done(ks, NC)) return;
  FlSweepRow<FWD, NC> row;
  row.ia = ia; row.ja = ja; row.klo = klo; row.kup = kup; row.order = order; row.a = a; row.pc = pc; row.src = src;
  row.out = out; row.n = n; row.r = 0; row.k = 0; row.nl = 0;
  sweep_chunks(n, ticket, err, row);
}

// the same rows for sweep_parts: a position's descriptor holds the row, its
// first entry (forward: ia; backward: the last entry, walked down), the
// entry count and (factor) the diagonal entry
template <bool FWD, int NC>
struct FlPtSweep {
  const double *a, *pc, *src;
  int n;
  double cc, s[NC];
  double av[kSweepSlots];
  __device__ void fetch(const SweepDesc &d, int MW) {
    constexpr int st = FWD ? 1 : -1;
#pragma unroll
    for (int t = 0; t < kSweepSlots; ++t)
      if (t < MW) av[t] = a[t < d.nl ? d.first + st * t : d.first];
    cc = pc[d.r];
#pragma unroll
    for (int q = 0; q < NC; ++q) s[q] = src[(size_t)q * n + d.r];
  }
  __device__ void begin() {
    if (!FWD) {
      const double cof = 1.0 / (cc + SMALL);
#pragma unroll
      for (int q = 0; q < NC; ++q) s[q] = cof * s[q];
    }
  }
  __device__ void take(int t, const double *w) {
#pragma unroll
    for (int q = 0; q < NC; ++q) s[q] = s[q] - av[t] * w[q];
  }
  __device__ double result(int q) const { return s[q] * cc; }
  __device__ void done(const SweepDesc &) {}
};

struct FlPtPc {
  const double *a;
  const int *tpos;
  double s;
  double a1[kSweepSlots], a2[kSweepSlots];
  __device__ void fetch(const SweepDesc &d, int MW) {
    s = a[d.aux];
#pragma unroll
    for (int t = 0; t < kSweepSlots; ++t)
      if (t < MW) {
        const int k = t < d.nl ? d.first + t : d.aux;
        a1[t] = a[k];
        a2[t] = a[t < d.nl ? tpos[k] : k];
      }
  }
  __device__ void begin() {}
  __device__ void take(int t, const double *w) { s = s - a1[t] * w[0] * a2[t]; }
  __device__ double result(int) const { return 1.0 / (s + SMALL); }
  __device__ void done(const SweepDesc &) {}
};

template <int MW, int TB>
__global__ __launch_bounds__(TB) void k_fl_pc_pt(int n, int R, int P, SweepTabs T, const int *__restrict__ step,
                                                 const int *__restrict__ sp, const int *__restrict__ tpos,
                                                 const double *__restrict__ a, double *pc, int *ticket, int *err) {
  FlPtPc row;
  row.a = a; row.tpos = tpos;
  sweep_parts<1, MW, FlPtPc, TB>(n, R, P, T, step, sp, pc, ticket, err, row);
}

template <bool FWD, int NC, int MW, int TB>
__global__ __launch_bounds__(TB) void k_fl_sweep_pt(int n, int R, int P, SweepTabs T, const int *__restrict__ step,
                                                    const int *__restrict__ sp, const double *__restrict__ a,
                                                    const double *__restrict__ pc, const double *__restrict__ src,
                                                    double *out, int *ticket, int *err, const FlKS *ks) {
  if (fl_all_done(ks, NC)) return;
  FlPtSweep<FWD, NC> row;
  row.a = a; row.pc = pc; row.src = src; row.n = n;
  sweep_parts<NC, MW, FlPtSweep<FWD, NC>, TB>(n, R, P, T, step, sp, out, ticket, err, row);
}

// ---------------------------------------------------------------- host
// Ranges of R rows in the sweep's order and, in each, its rows by level
// (ties by row), with each position's descriptor: the tables of sweep_parts.
static int fl_parts_build(FlCtx &c, const std::vector<int> &lev, bool fwd, int R, FlCtx::Parts &pt) {
  const int n = c.n;
  const std::vector<int> &ia = c.h_ia, &ja = c.h_ja;
  pt.R = R;
  pt.P = (n + R - 1) / R;
  std::vector<int> rows, first(n), cnt(n), aux(n), lrow(n), step, sp;
  rows.reserve(n);
  std::vector<std::pair<int, int>> buf;
  for (int q = 0; q < pt.P; ++q) {
    const int lo = fwd ? q * R : std::max(0, n - (q + 1) * R);
    const int hi = fwd ? std::min(n, (q + 1) * R) : n - q * R;
    buf.clear();
    for (int r = lo; r < hi; ++r) buf.emplace_back(lev[r], r);
    std::sort(buf.begin(), buf.end());
    sp.push_back((int)step.size());
    for (size_t t = 0; t < buf.size(); ++t) {
      if (t == 0 || buf[t].first != buf[t - 1].first) step.push_back((int)rows.size());
      rows.push_back(buf[t].second);
    }
    step.push_back((int)rows.size());
  }
  sp.push_back((int)step.size());
  pt.wmax = 0;
  for (int q = 0; q < pt.P; ++q)
    for (int k = sp[q]; k + 1 < sp[q + 1]; ++k) pt.wmax = std::max(pt.wmax, step[k + 1] - step[k]);
  int W = 1;
  for (int p = 0; p < n; ++p) {
    const int r = rows[p];
    first[p] = fwd ? ia[r] : ia[r + 1] - 1;
    cnt[p] = fwd ? c.h_klo[r] - ia[r] : ia[r + 1] - c.h_kup[r];
    aux[p] = c.h_kii[r];
    W = std::max(W, cnt[p]);
  }
  if (W > kSweepSlots) { set_error("a row has more lower or upper entries than one launch holds"); return FVK_ERR_ARG; }
  pt.W = W;
  // a range's rows: forward [qR, (q+1)R), backward counted from the end
  auto range_lo = [&](int r) { return fwd ? r / R * R : std::max(0, n - (n - 1 - r) / R * R - R); };
  std::vector<int> cols((size_t)W * n, 0), lcol((size_t)W * n, -1);
  for (int p = 0; p < n; ++p) {
    const int r = rows[p], lo = range_lo(r);
    lrow[p] = r - lo;
    for (int t = 0; t < cnt[p]; ++t) {
      const int j = ja[fwd ? first[p] + t : first[p] - t];
      cols[(size_t)t * n + p] = j;
      if (range_lo(j) == lo) lcol[(size_t)t * n + p] = j - lo;
    }
  }
  std::vector<int> hd, jc, lc;
  sweep_pack(n, W, rows, first, cnt, aux, lrow, cols, lcol, hd, jc, lc);
  int rc = fl_up(c, &pt.hd, hd);
  if (!rc) rc = fl_up(c, &pt.jc, jc);
  if (!rc) rc = fl_up(c, &pt.lc, lc);
  if (!rc) rc = fl_up(c, &pt.step, step);
  if (!rc) rc = fl_up(c, &pt.sp, sp);
  return rc;
}

static int fl_parts_setup(FlCtx &c) {
  for (FlCtx::Parts *pt : {&c.pf, &c.pb}) {
    for (int *q : {pt->hd, pt->jc, pt->lc, pt->step, pt->sp}) fl_free(c, q);
    *pt = FlCtx::Parts();
  }
  int rc = fl_parts_build(c, c.h_levf, true, c.part_rows, c.pf);
  if (!rc) rc = fl_parts_build(c, c.h_levb, false, c.part_rows, c.pb);
  if (rc) return rc;
  const int bytes = 3 * c.part_rows * 8;
  if (bytes > 65536) {
    const void *k[] = {
#define K_(D, C, MW) (const void *)k_fl_sweep_pt<D, C, MW, 64>, (const void *)k_fl_sweep_pt<D, C, MW, BS>
#define K2_(C, MW) K_(true, C, MW), K_(false, C, MW)
        K2_(1, 4), K2_(2, 4), K2_(3, 4), K2_(1, 8), K2_(2, 8), K2_(3, 8)
#undef K2_
#undef K_
    };
    for (const void *f : k) FVK_HIP(hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, bytes));
  }
  return 0;
}

static int fl_levels(const std::vector<int> &ia, const std::vector<int> &ja, int n, bool fwd,
                     std::vector<int> &ptr, std::vector<int> &rows, std::vector<int> &lev) {
  lev.assign(n, 0);
  int nl = 0;
  for (int s = 0; s < n; ++s) {
    const int i = fwd ? s : n - 1 - s;
    int l = 0;
    for (int k = ia[i]; k < ia[i + 1]; ++k) {
      const int j = ja[k];
      if (fwd ? j < i : j > i) l = std::max(l, lev[j] + 1);
    }
    lev[i] = l;
    nl = std::max(nl, l + 1);
  }
  ptr.assign(nl + 1, 0);
  for (int i = 0; i < n; ++i) ptr[lev[i] + 1]++;
  for (int l = 0; l < nl; ++l) ptr[l + 1] += ptr[l];
  rows.assign(n, 0);
  std::vector<int> pos(ptr.begin(), ptr.end() - 1);
  for (int i = 0; i < n; ++i) rows[pos[lev[i]]++] = i;
  return 0;
}

static int fl_sweep_grid(const FlCtx &c) {
  return std::max(1, std::min((c.n + BS - 1) / BS, c.sweep_blocks > 0 ? c.sweep_blocks : c.n_cu));
}
static int fl_mark_grid(size_t cnt) { return (int)std::max<size_t>(1, std::min<size_t>((cnt + BS - 1) / BS, 4096)); }

// blocks of the partitioned kernels (sweep_blocks 0: as many per CU as the
// LDS of a block lets reside, at most four of 64 threads or two of 256)
static int fl_pt_grid(const FlCtx &c, const FlCtx::Parts &pt, size_t lds, int tb) {
  const int per_cu = std::max(1, std::min<int>(tb == 64 ? 4 : 2, (int)(163840 / (lds + 1024))));
  return std::max(1, std::min(pt.P, c.sweep_blocks > 0 ? c.sweep_blocks : c.n_cu * per_cu));
}

// one partitioned sweep: one wave a block when no level of a range has more
// than 64 rows
template <bool FWD, int C, int MW>
static void fl_sweep_launch(FlCtx &c, const FlCtx::Parts &X, size_t lds, const double *a, const double *src,
                            double *out, int *ticket, const FlKS *ks) {
  if (X.wmax <= 64)
    k_fl_sweep_pt<FWD, C, MW, 64><<<fl_pt_grid(c, X, lds, 64), 64, lds, c.stream>>>(
        c.n, X.R, X.P, X.tabs(), X.step, X.sp, a, c.pc, src, out, ticket, c.sweep_int + 3, ks);
  else
    k_fl_sweep_pt<FWD, C, MW, BS><<<fl_pt_grid(c, X, lds, BS), BS, lds, c.stream>>>(
        c.n, X.R, X.P, X.tabs(), X.step, X.sp, a, c.pc, src, out, ticket, c.sweep_int + 3, ks);
}

template <bool FWD>
static void fl_sweep_any(FlCtx &c, int ncmp, const FlCtx::Parts &X, size_t lds, const double *a, const double *src,
                         double *out, int *ticket, const FlKS *ks) {
  if (X.W <= 4) {
    if (ncmp == 1) fl_sweep_launch<FWD, 1, 4>(c, X, lds, a, src, out, ticket, ks);
    else if (ncmp == 2) fl_sweep_launch<FWD, 2, 4>(c, X, lds, a, src, out, ticket, ks);
    else fl_sweep_launch<FWD, 3, 4>(c, X, lds, a, src, out, ticket, ks);
  } else {
    if (ncmp == 1) fl_sweep_launch<FWD, 1, 8>(c, X, lds, a, src, out, ticket, ks);
    else if (ncmp == 2) fl_sweep_launch<FWD, 2, 8>(c, X, lds, a, src, out, ticket, ks);
    else fl_sweep_launch<FWD, 3, 8>(c, X, lds, a, src, out, ticket, ks);
  }
}

static int fl_precond(FlCtx &c, const double *a) {
  if (c.one_launch && c.parts) {
    const FlCtx::Parts &P = c.pf;
    const size_t lds = (size_t)P.R * 8;
    k_fl_mark<<<fl_mark_grid(c.n), BS, 0, c.stream>>>(c.n, c.pc, c.sweep_int + 2, nullptr, 1);
#define PC_(MW, TB) k_fl_pc_pt<MW, TB><<<fl_pt_grid(c, P, lds, TB), TB, lds, c.stream>>>(c.n, P.R, P.P, P.tabs(), \
    P.step, P.sp, c.tpos, a, c.pc, c.sweep_int + 2, c.sweep_int + 3)
    if (P.W <= 4) {
      if (P.wmax <= 64) PC_(4, 64); else PC_(4, BS);
    } else {
      if (P.wmax <= 64) PC_(8, 64); else PC_(8, BS);
    }
#undef PC_
    FVK_HIP(hipGetLastError());
    return 0;
  }
  if (c.one_launch) {
    k_fl_mark<<<fl_mark_grid(c.n), BS, 0, c.stream>>>(c.n, c.pc, c.sweep_int + 2, nullptr, 1);
    k_fl_pc_sf<<<fl_sweep_grid(c), BS, 0, c.stream>>>(c.n, c.lf_rows, c.ia, c.ja, c.kii, c.klo, c.tpos, a, c.pc,
                                                       c.sweep_int + 2, c.sweep_int + 3);
    FVK_HIP(hipGetLastError());
    return 0;
  }
  for (size_t l = 0; l + 1 < c.lf_ptr.size(); ++l) {
    const int cnt = c.lf_ptr[l + 1] - c.lf_ptr[l];
    k_fl_pc_level<<<grid_for(cnt), BS, 0, c.stream>>>(c.lf_rows + c.lf_ptr[l], cnt, c.ia, c.ja, c.kii, c.tpos, a,
                                                       c.pc);
  }
  FVK_HIP(hipGetLastError());
  return 0;
}

// x = M^-1 b (direct_solve), t: scratch
static int fl_direct(FlCtx &c, int ncmp, const double *a, const double *b, double *t, double *x, const FlKS *ks) {
  if (c.one_launch && c.parts) {
    const size_t cnt = (size_t)ncmp * c.n;
    const size_t lds = (size_t)ncmp * c.part_rows * 8;
    k_fl_mark<<<fl_mark_grid(cnt), BS, 0, c.stream>>>(cnt, t, c.sweep_int, ks, ncmp);
    fl_sweep_any<true>(c, ncmp, c.pf, lds, a, b, t, c.sweep_int, ks);
    k_fl_mark<<<fl_mark_grid(cnt), BS, 0, c.stream>>>(cnt, x, c.sweep_int + 1, ks, ncmp);
    fl_sweep_any<false>(c, ncmp, c.pb, lds, a, t, x, c.sweep_int + 1, ks);
    FVK_HIP(hipGetLastError());
    return 0;
  }
  if (c.one_launch) {
    const size_t cnt = (size_t)ncmp * c.n;
    const int G = fl_sweep_grid(c);
    k_fl_mark<<<fl_mark_grid(cnt), BS, 0, c.stream>>>(cnt, t, c.sweep_int, ks, ncmp);
#define FWD(C) k_fl_sweep_sf<true, C><<<G, BS, 0, c.stream>>>(c.n, c.lf_rows, c.ia, c.ja, c.klo, c.kup, a, c.pc, b, \
                                                              t, c.sweep_int, c.sweep_int + 3, ks)
#define BWD(C) k_fl_sweep_sf<false, C><<<G, BS, 0, c.stream>>>(c.n, c.lb_rows, c.ia, c.ja, c.klo, c.kup, a, c.pc, t, \
                                                               x, c.sweep_int + 1, c.sweep_int + 3, ks)
    if (ncmp == 1) FWD(1); else if (ncmp == 2) FWD(2); else FWD(3);
    k_fl_mark<<<fl_mark_grid(cnt), BS, 0, c.stream>>>(cnt, x, c.sweep_int + 1, ks, ncmp);
    if (ncmp == 1) BWD(1); else if (ncmp == 2) BWD(2); else BWD(3);
#undef FWD
#undef BWD
    FVK_HIP(hipGetLastError());
    return 0;
  }
  for (size_t l = 0; l + 1 < c.lf_ptr.size(); ++l) {
    const int cnt = c.lf_ptr[l + 1] - c.lf_ptr[l];
    k_fl_fwd_level<<<grid_for(cnt), BS, 0, c.stream>>>(c.lf_rows + c.lf_ptr[l], cnt, c.n, ncmp, c.ia, c.ja, a, c.pc,
                                                        b, t, ks);
  }
  for (size_t l = 0; l + 1 < c.lb_ptr.size(); ++l) {
    const int cnt = c.lb_ptr[l + 1] - c.lb_ptr[l];
    k_fl_bwd_level<<<grid_for(cnt), BS, 0, c.stream>>>(c.lb_rows + c.lb_ptr[l], cnt, c.n, ncmp, c.ia, c.ja, a, c.pc,
                                                        t, x, ks);
  }
  FVK_HIP(hipGetLastError());
  return 0;
}

// lin_eqns_solver (solv_m.f90:10-30) for ncmp right-hand sides sharing the
// matrix a (device, CSR of the context): phi and res (the right-hand side
// in, the initial residual out) component-major on the device; norms to
// ks[c].norm.  isol 1: cg_solver, 2: bicg_solver.
static int fl_solve(FlCtx &c, int isol, int ncmp, const double *a, double *phi, double *res, int n_iter,
                    double reduc) {
  const int n = c.n, P = fl_pgrid(n), G = grid_for(n);
  double *w1 = wkv(c, 0), *w2 = wkv(c, 1), *w3 = wkv(c, 2), *w4 = wkv(c, 3), *w5 = wkv(c, 4), *wt = wkv(c, 5);
  const bool bicg = isol == 2;
  FVK_HIP(hipMemsetAsync(c.wk, 0, sizeof(double) * 5 * 3 * (size_t)n, c.stream));
  k_fl_ks_open<<<1, 64, 0, c.stream>>>(c.ks, ncmp);
  k_fl_init<<<P, RB, 0, c.stream>>>(n, ncmp, c.ia, c.ja, a, phi, res, w1, c.ks, c.part);
  k_fl_scal<<<1, RB, 0, c.stream>>>(0, ncmp, P, 1, c.part, c.ks, reduc, bicg ? 1 : 0);
  int rc = fl_precond(c, a);
  if (rc) return rc;
  for (int it = 0; it < n_iter; ++it) {
    if (!bicg) {
      if ((rc = fl_direct(c, ncmp, a, w1, wt, w3, c.ks))) return rc;
      k_fl_dot<<<P, RB, 0, c.stream>>>(n, ncmp, w1, w3, c.ks, c.part);
      k_fl_scal<<<1, RB, 0, c.stream>>>(10, ncmp, P, 1, c.part, c.ks, reduc, 0);
      k_fl_dir<<<G, BS, 0, c.stream>>>(n, ncmp, 0, w2, w3, nullptr, c.ks);
      k_fl_mv_dot<<<P, RB, 0, c.stream>>>(n, ncmp, c.ia, c.ja, a, w2, w3, w2, c.ks, c.part);
      k_fl_scal<<<1, RB, 0, c.stream>>>(11, ncmp, P, 1, c.part, c.ks, reduc, 0);
      k_fl_upd<<<P, RB, 0, c.stream>>>(n, ncmp, 2, phi, w2, w1, w3, c.ks, c.part, 1);
      k_fl_scal<<<1, RB, 0, c.stream>>>(12, ncmp, P, 1, c.part, c.ks, reduc, 0);
    } else {
      k_fl_dot<<<P, RB, 0, c.stream>>>(n, ncmp, w1, res, c.ks, c.part);
      k_fl_scal<<<1, RB, 0, c.stream>>>(20, ncmp, P, 1, c.part, c.ks, reduc, 1);
      k_fl_dir<<<G, BS, 0, c.stream>>>(n, ncmp, 1, w2, w1, w3, c.ks);
      if ((rc = fl_direct(c, ncmp, a, w2, wt, w5, c.ks))) return rc;
      k_fl_mv_dot<<<P, RB, 0, c.stream>>>(n, ncmp, c.ia, c.ja, a, w5, w3, res, c.ks, c.part);
      k_fl_scal<<<1, RB, 0, c.stream>>>(21, ncmp, P, 1, c.part, c.ks, reduc, 1);
      k_fl_upd<<<P, RB, 0, c.stream>>>(n, ncmp, 5, phi, w5, w1, w3, c.ks, c.part, 0);
      if ((rc = fl_direct(c, ncmp, a, w1, wt, w5, c.ks))) return rc;
      k_fl_mv_dot2<<<P, RB, 0, c.stream>>>(n, ncmp, c.ia, c.ja, a, w5, w4, w1, c.ks, c.part);
      k_fl_scal<<<1, RB, 0, c.stream>>>(22, ncmp, P, 2, c.part, c.ks, reduc, 1);
      k_fl_upd<<<P, RB, 0, c.stream>>>(n, ncmp, 2, phi, w5, w1, w4, c.ks, c.part, 1);
      k_fl_scal<<<1, RB, 0, c.stream>>>(23, ncmp, P, 1, c.part, c.ks, reduc, 1);
    }
  }
  FVK_HIP(hipGetLastError());
  return 0;
}

// ---------------------------------------------------------------- SIMPLE
struct FlCfg {
  double cc, cd, urf_u, urf_p, bl_c;
  int n_iter_u, n_iter_p;
  double reduc_u, reduc_p;
};

struct FlFlow {
  FlCtx *c = nullptr;
  FlCfg cfg{};
  std::vector<void *> owned;
  double *u = nullptr, *ub = nullptr, *ug = nullptr;   // velocity: cells, boundary, gradient
  double *p = nullptr, *pb = nullptr, *pg = nullptr;   // pressure
  double *q = nullptr, *qb = nullptr, *qg = nullptr;   // pressure correction
  double *rhs = nullptr, *diag = nullptr, *val = nullptr, *spd = nullptr;
  double *cc = nullptr, *ccb = nullptr, *cd = nullptr, *cdb = nullptr;
  double *fa = nullptr, *fb = nullptr, *fd = nullptr, *fcc = nullptr, *fdc = nullptr;  // face terms
  int *bfl = nullptr;          // boundary faces
  int n_bf = 0;
  double res0[4] = {0, 0, 0, 0};
};

template <class T>
static int ff_alloc(FlFlow &F, T **p, size_t n) {
  void *q = nullptr;
  FVK_HIP(hipMalloc(&q, (n ? n : 1) * sizeof(T)));
  F.owned.push_back(q);
  FVK_HIP(hipMemsetAsync(q, 0, (n ? n : 1) * sizeof(T), F.c->stream));
  FVK_HIP(hipStreamSynchronize(F.c->stream));
  *p = (T *)q;
  return 0;
}

// boundary type ranges (transport_m.f90:11-14)
__device__ __forceinline__ bool fl_is(int ty, int lo) { return ty >= lo && ty <= lo + 9; }

__global__ __launch_bounds__(BS) void k_fl_fill(size_t n, double *a, double v) {
  for (size_t i = blockIdx.x * (size_t)BS + threadIdx.x; i < n; i += (size_t)gridDim.x * BS) a[i] = v;
}

// ns_source (transport_m.f90:726-752): rhs = 0 + (-grad p) vol
__global__ __launch_bounds__(BS) void k_fl_source(int n, int nd, const double *__restrict__ pg,
                                                  const double *__restrict__ vol, double *rhs) {
  const int i = blockIdx.x * BS + threadIdx.x;
  if (i >= n) return;
  for (int d = 0; d < nd; ++d) rhs[(size_t)d * n + i] = 0.0 + (-pg[(size_t)d * n + i]) * vol[i];
}

struct FlGeo {
  int n, nf, nb, nd;
  const int *j1, *jn, *btag, *kij, *kji;
  const double *xvc, *xfc, *nrm, *area, *xpc, *lpn, *wfp;
};

// construct_pde face terms (transport_m.f90:97-297)
__global__ __launch_bounds__(BS) void k_fl_pde_face(FlGeo g, double bl_c, const double *__restrict__ u,
                                                    double *ub, const double *__restrict__ ug,
                                                    const double *__restrict__ spd, const double *__restrict__ cc,
                                                    const double *__restrict__ ccb, const double *__restrict__ cd,
                                                    const double *__restrict__ cdb, double *val, double *fa,
                                                    double *fb, double *fd, double *fcc, double *fdc) {
  const int f = xcd_block() * BS + threadIdx.x;
  if (f >= g.nf) return;
  const int n = g.n, nf = g.nf, nd = g.nd, p = g.j1[f], q = g.jn[f];
  const double ar = g.area[f];
  double nv[3] = {0, 0, 0};
  for (int d = 0; d < nd; ++d) nv[d] = g.nrm[(size_t)d * nf + f];
  if (q >= 0) {
    const double w = g.wfp[f], wn = 1.0 - w;
    double del[3] = {0, 0, 0}, dl = 0.0;
    for (int d = 0; d < nd; ++d) del[d] = g.xvc[(size_t)d * n + q] - g.xvc[(size_t)d * n + p];
    for (int d = 0; d < nd; ++d) dl = dl + del[d] * del[d];
    const double dm = sqrt(dl) + SMALL;
    for (int d = 0; d < nd; ++d) del[d] = nv[d] - del[d] / dm;
    double fg[3][3];  // fg[d][c] = grad(d, c)
    for (int c = 0; c < nd; ++c)
      for (int d = 0; d < nd; ++d)
        fg[d][c] = w * ug[(size_t)(c * nd + d) * n + q] + wn * ug[(size_t)(c * nd + d) * n + p];
    double dx[3];
    for (int d = 0; d < nd; ++d)
      dx[d] = g.xfc[(size_t)d * nf + f] - (w * g.xvc[(size_t)d * n + q] + wn * g.xvc[(size_t)d * n + p]);
    double fphi[3];
    for (int c = 0; c < nd; ++c) {
      double s = 0.0;
      for (int d = 0; d < nd; ++d) s = s + fg[d][c] * dx[d];
      fphi[c] = w * u[(size_t)c * n + q] + wn * u[(size_t)c * n + p] + s;
    }
    const double fc = w * cc[q] + wn * cc[p];
    const double fdd = w * cd[q] + wn * cd[p];
    const double cv = fc * ar * spd[f];
    const double dv = fdd * ar / g.lpn[f];
    const double c_ij = fmin(cv, 0.0), c_ji = -fmax(cv, 0.0), d_ij = -dv;
    for (int c = 0; c < nd; ++c) {
      const double imp = cv > 0.0 ? u[(size_t)c * n + p] : u[(size_t)c * n + q];
      const double c_cor = bl_c * cv * (imp - fphi[c]);
      double dc = 0.0;
      for (int d = 0; d < nd; ++d) dc = dc + fg[d][c] * del[d];
      dc = dc * fdd * ar;
      double de = 0.0;
      for (int d = 0; d < nd; ++d) de = de + fg[c][d] * nv[d];
      de = de * fdd * ar;
      dc = dc + de;
      fcc[(size_t)c * nf + f] = c_cor;
      fdc[(size_t)c * nf + f] = dc;
    }
    val[g.kij[f]] = 0.0 + c_ij + d_ij;
    val[g.kji[f]] = 0.0 + c_ji + d_ij;
    fa[f] = c_ij;
    fb[f] = c_ji;
    fd[f] = d_ij;
  } else {
    const int b = -q - 1, ty = g.btag[b];
    double fg[3][3];
    for (int c = 0; c < nd; ++c)
      for (int d = 0; d < nd; ++d) fg[d][c] = ug[(size_t)(c * nd + d) * n + p];
    double fphi[3], vphi[3];
    for (int c = 0; c < nd; ++c) {
      fphi[c] = ub[(size_t)c * g.nb + b];
      vphi[c] = u[(size_t)c * n + p];
    }
    const double fc = ccb[b], fdd = cdb[b];
    const double cv = fc * ar * spd[f];
    const double dv = fdd * ar / g.lpn[f];
    double c_ij = fmin(cv, 0.0), d_ij = -dv;
    double vn = 0.0;
    for (int d = 0; d < nd; ++d) vn = vn + vphi[d] * nv[d];
    double dexp[3];
    for (int c = 0; c < nd; ++c) {
      if (fl_is(ty, 30)) {
        dexp[c] = fphi[c] - (vphi[c] - nv[c] * vn);
        dexp[c] = dexp[c] / g.lpn[f];
      } else if (fl_is(ty, 40)) {
        dexp[c] = 2.0 * (0.0 - nv[c] * vn);
        dexp[c] = dexp[c] / g.lpn[f];
      } else {
        double s = 0.0;
        for (int d = 0; d < nd; ++d) s = s + (fg[c][d] + fg[d][c]) * nv[d];
        dexp[c] = s;
      }
    }
    double dcor[3];
    for (int c = 0; c < nd; ++c) dcor[c] = fdd * ar * dexp[c] - dv * (fphi[c] - vphi[c]);
    if (fl_is(ty, 20)) {
      for (int c = 0; c < nd; ++c) fphi[c] = vphi[c];
      c_ij = 0.0;
      d_ij = 0.0;
    } else if (fl_is(ty, 40)) {
      for (int c = 0; c < nd; ++c) fphi[c] = vphi[c] - nv[c] * vn;
    }
    for (int c = 0; c < nd; ++c) {
      ub[(size_t)c * g.nb + b] = fphi[c];
      fcc[(size_t)c * nf + f] = fphi[c];
      fdc[(size_t)c * nf + f] = dcor[c];
    }
    fa[f] = c_ij;
    fd[f] = d_ij;
  }
}

// construct_pde cell part: diagonal and right-hand side gathered over the
// cell's faces (ascending), then under-relaxation (transport_m.f90:300-347)
__global__ __launch_bounds__(BS) void k_fl_pde_cell(int n, int nf, int nd, const int *__restrict__ cptr,
                                                    const int *__restrict__ cfl, const int *__restrict__ j1,
                                                    const int *__restrict__ jn, const int *__restrict__ kii,
                                                    const double *__restrict__ fa, const double *__restrict__ fb,
                                                    const double *__restrict__ fd, const double *__restrict__ fcc,
                                                    const double *__restrict__ fdc, const double *__restrict__ u,
                                                    double urf, double *val, double *rhs, double *diag) {
  const int i = xcd_block() * BS + threadIdx.x;
  if (i >= n) return;
  double v = 0.0, r[3] = {0, 0, 0};
  for (int c = 0; c < nd; ++c) r[c] = rhs[(size_t)c * n + i];
  for (int k = cptr[i]; k < cptr[i + 1]; ++k) {
    const int f = cfl[k];
    if (jn[f] < 0) {
      v = v - fa[f] - fd[f];
      const double cd = fa[f] + fd[f];
      for (int c = 0; c < nd; ++c) r[c] = r[c] - cd * fcc[(size_t)c * nf + f] + fdc[(size_t)c * nf + f];
    } else if (j1[f] == i) {
      v = v - fa[f] - fd[f];
      for (int c = 0; c < nd; ++c) r[c] = r[c] + fcc[(size_t)c * nf + f] + fdc[(size_t)c * nf + f];
    } else {
      v = v - fb[f] - fd[f];
      for (int c = 0; c < nd; ++c) r[c] = r[c] - fcc[(size_t)c * nf + f] - fdc[(size_t)c * nf + f];
    }
  }
  const double urfr = 1.0 / urf;
  v = v * urfr;
  for (int c = 0; c < nd; ++c) rhs[(size_t)c * n + i] = r[c] + (1.0 - urf) * v * u[(size_t)c * n + i];
  val[kii[i]] = v;
  diag[i] = 1.0 / v;
}

// construct_pp face terms (transport_m.f90:466-570), first pass
__global__ __launch_bounds__(BS) void k_fl_pp_face(FlGeo g, const double *__restrict__ u, const double *__restrict__ ub,
                                                   const double *__restrict__ ug, const double *__restrict__ pr,
                                                   const double *__restrict__ pg, const double *__restrict__ cc,
                                                   const double *__restrict__ ccb, const double *__restrict__ diag,
                                                   double *spd, double *val, double *fa, double *fd, double *fdc) {
  const int f = xcd_block() * BS + threadIdx.x;
  if (f >= g.nf) return;
  const int n = g.n, nf = g.nf, nd = g.nd, p = g.j1[f], q = g.jn[f];
  const double ar = g.area[f];
  double nv[3] = {0, 0, 0};
  for (int d = 0; d < nd; ++d) nv[d] = g.nrm[(size_t)d * nf + f];
  if (q >= 0) {
    const double w = g.wfp[f], wn = 1.0 - w;
    double dx[3];
    for (int d = 0; d < nd; ++d)
      dx[d] = g.xfc[(size_t)d * nf + f] - (w * g.xvc[(size_t)d * n + q] + wn * g.xvc[(size_t)d * n + p]);
    double fv[3];
    for (int c = 0; c < nd; ++c) {
      double s = 0.0;
      for (int d = 0; d < nd; ++d)
        s = s + (w * ug[(size_t)(c * nd + d) * n + q] + wn * ug[(size_t)(c * nd + d) * n + p]) * dx[d];
      fv[c] = w * u[(size_t)c * n + q] + wn * u[(size_t)c * n + p] + s;
    }
    double sp = 0.0;
    for (int d = 0; d < nd; ++d) sp = sp + fv[d] * nv[d];
    spd[f] = sp;
    const double fc = w * cc[q] + wn * cc[p];
    const double fapr = w * diag[q] + wn * diag[p];
    const double d_ij = -(fc * fapr * (ar * ar));
    const double cv = -(fc * ar * sp);
    double t = 0.0;
    for (int d = 0; d < nd; ++d)
      t = t + (w * pg[(size_t)d * n + q] + wn * pg[(size_t)d * n + p]) *
                  (g.xvc[(size_t)d * n + q] - g.xvc[(size_t)d * n + p]);
    double dcor = pr[q] - pr[p];
    dcor = dcor - t;
    dcor = -(dcor * d_ij);
    val[g.kij[f]] = 0.0 + d_ij;
    val[g.kji[f]] = 0.0 + d_ij;
    fa[f] = cv;
    fd[f] = d_ij;
    fdc[f] = dcor;
  } else {
    const int b = -q - 1;
    double sp = 0.0;
    for (int d = 0; d < nd; ++d) sp = sp + ub[(size_t)d * g.nb + b] * nv[d];
    double cv = -(ccb[b] * ar * sp);
    if (fl_is(g.btag[b], 20)) {
      if (sp < 0.0) sp = 0.0;
      cv = 0.0;
    }
    spd[f] = sp;
    fa[f] = cv;
  }
}

// flux_out_corr (transport_m.f90:350-425): inlet and outlet mass fluxes
__global__ __launch_bounds__(RB) void k_fl_inout(int nbf, const int *__restrict__ bfl, const int *__restrict__ jn,
                                                 const int *__restrict__ btag, const double *__restrict__ ccb,
                                                 const double *__restrict__ area, const double *__restrict__ spd,
                                                 double *part) {
  double v[2] = {0.0, 0.0};
  for (int t = blockIdx.x * RB + threadIdx.x; t < nbf; t += gridDim.x * RB) {
    const int f = bfl[t], b = -jn[f] - 1, ty = btag[b];
    const double c = ccb[b] * area[f] * spd[f];
    if (fl_is(ty, 10)) v[0] += c;
    else if (fl_is(ty, 20)) v[1] += c;
  }
  rblock_sum<2, RB>(v);
  if (threadIdx.x == 0) {
    part[blockIdx.x] = v[0];
    part[RG + blockIdx.x] = v[1];
  }
}

__global__ __launch_bounds__(RB) void k_fl_crat(int P, const double *part, double *scal) {
  const double ci = fl_total(part, P, 0);
  const double co = fl_total(part, P, 1);
  if (threadIdx.x == 0) {
    double r = ci / (co + SMALL);
    r = fabs(r);
    scal[0] = fmin(1000.0, r);
  }
}

// outlet fluxes scaled by c_rat (transport_m.f90:599-627)
__global__ __launch_bounds__(BS) void k_fl_outlet(int nbf, const int *__restrict__ bfl, const int *__restrict__ jn,
                                                  const int *__restrict__ btag, const double *__restrict__ scal,
                                                  double *spd) {
  const int t = blockIdx.x * BS + threadIdx.x;
  if (t >= nbf) return;
  const int f = bfl[t], b = -jn[f] - 1;
  if (fl_is(btag[b], 20)) spd[f] = spd[f] * scal[0];
}

// construct_pp cell part: diagonal, right-hand side (first pass over the
// faces, then the outlet faces' corrected fluxes)
__global__ __launch_bounds__(BS) void k_fl_pp_cell(int n, int nb, const int *__restrict__ cptr,
                                                   const int *__restrict__ cfl, const int *__restrict__ j1,
                                                   const int *__restrict__ jn, const int *__restrict__ kii,
                                                   const int *__restrict__ btag, const double *__restrict__ ccb,
                                                   const double *__restrict__ area, const double *__restrict__ spd,
                                                   const double *__restrict__ fa, const double *__restrict__ fd,
                                                   const double *__restrict__ fdc, double *val, double *rhs) {
  const int i = xcd_block() * BS + threadIdx.x;
  if (i >= n) return;
  double v = 0.0, r = 0.0;
  for (int k = cptr[i]; k < cptr[i + 1]; ++k) {
    const int f = cfl[k];
    if (jn[f] < 0) {
      r = r + fa[f];
    } else if (j1[f] == i) {
      v = v - fd[f];
      r = r + fa[f] + fdc[f];
    } else {
      v = v - fd[f];
      r = r - fa[f] - fdc[f];
    }
  }
  for (int k = cptr[i]; k < cptr[i + 1]; ++k) {
    const int f = cfl[k];
    if (jn[f] >= 0) continue;
    const int b = -jn[f] - 1;
    if (!fl_is(btag[b], 20)) continue;
    r = r + -(ccb[b] * area[f] * spd[f]);
  }
  val[kii[i]] = v;
  rhs[i] = r;
}

// ns_correction (transport_m.f90:757-800); the serial build keeps the
// reference pressure at its initial zero (blk is 1 there)
__global__ __launch_bounds__(BS) void k_fl_correct(int n, int nd, const double *__restrict__ qg,
                                                   const double *__restrict__ vol, const double *__restrict__ diag,
                                                   const double *__restrict__ q, double ref, double urf, double *u,
                                                   double *p) {
  const int i = blockIdx.x * BS + threadIdx.x;
  if (i >= n) return;
  for (int d = 0; d < nd; ++d) {
    double dv = -qg[(size_t)d * n + i];
    dv = dv * vol[i] * diag[i];
    u[(size_t)d * n + i] = u[(size_t)d * n + i] + dv;
  }
  const double dp = (q[i] - ref) * urf;
  p[i] = p[i] + dp;
}

static FlGeo fl_geo(const FlCtx &c) {
  FlGeo g;
  g.n = c.n; g.nf = c.nf; g.nb = c.nb; g.nd = c.nd;
  g.j1 = c.j1; g.jn = c.jn; g.btag = c.btag; g.kij = c.kij; g.kji = c.kji;
  g.xvc = c.xvc; g.xfc = c.xfc; g.nrm = c.nrm; g.area = c.area; g.xpc = c.xpc; g.lpn = c.lpn; g.wfp = c.wfp;
  return g;
}

int fl_grad(FlCtx &c, int ncmp, const double *phi, double *phib, double *g, int neu, const int *bfl, int n_bf) {
  k_fl_grad<<<grid_for(c.n), BS, 0, c.stream>>>(c.n, c.nf, c.nb, c.nd, ncmp, c.cptr, c.cfl, c.j1, c.jn, c.xvc, c.xfc,
                                                 c.ls, phi, phib, neu, g);
  if (neu && n_bf)
    k_fl_grad_bnd<<<grid_for(n_bf), BS, 0, c.stream>>>(n_bf, bfl, c.n, c.nf, c.nb, c.nd, ncmp, c.j1, c.jn, c.xvc,
                                                        c.xfc, phi, g, phib);
  FVK_HIP(hipGetLastError());
  return 0;
}

// ---------------------------------------------------------------- host set-up
// The cell matrix pattern (1-based CSR) and the Krylov workspace.
static int fl_csr_setup(FlCtx &c, int n, const int *ngp, const int *ngc) {
  c.n = n;
  c.nnz = ngp[n] - 1;
  if (ngp[0] != 1 || c.nnz < n) { set_error("bad matrix pattern"); return FVK_ERR_ARG; }
  std::vector<int> ia(n + 1), ja(c.nnz);
  for (int i = 0; i <= n; ++i) {
    ia[i] = ngp[i] - 1;
    if (i && ia[i] < ia[i - 1]) { set_error("bad matrix pattern"); return FVK_ERR_ARG; }
  }
  for (int k = 0; k < c.nnz; ++k) {
    ja[k] = ngc[k] - 1;
    if (ja[k] < 0 || ja[k] >= n) { set_error("matrix column out of range"); return FVK_ERR_ARG; }
  }
  auto csr_index = [&](int i, int j) {
    for (int k = ia[i]; k < ia[i + 1]; ++k)
      if (ja[k] == j) return k;
    return -1;
  };
  std::vector<int> kii(n), klo(n), kup(n), tpos(c.nnz, -1);
  int mx = 0;
  for (int i = 0; i < n; ++i) {
    if ((kii[i] = csr_index(i, i)) < 0) { set_error("matrix pattern lacks a diagonal entry"); return FVK_ERR_ARG; }
    int k = ia[i];
    while (k < ia[i + 1] && ja[k] < i) ++k;
    klo[i] = k;
    k = ia[i + 1];
    while (k > ia[i] && ja[k - 1] > i) --k;
    kup[i] = k;
    mx = std::max({mx, klo[i] - ia[i], ia[i + 1] - kup[i]});
    for (int q = ia[i]; q < klo[i]; ++q)
      if ((tpos[q] = csr_index(ja[q], i)) < 0) { set_error("matrix pattern is not symmetric"); return FVK_ERR_ARG; }
  }
  c.one_launch = mx <= kSweepSlots;
  c.h_ia = ia;
  c.h_ja = ja;
  c.h_kii = kii;
  c.h_klo = klo;
  c.h_kup = kup;
  int rc = 0;
#define A(ptr, cnt_) if (!rc) rc = fl_alloc(c, &ptr, cnt_)
#define U(ptr, vec) if (!rc) rc = fl_up(c, &ptr, vec)
  U(c.ia, ia); U(c.ja, ja); U(c.kii, kii); U(c.klo, klo); U(c.kup, kup); U(c.tpos, tpos);
  A(c.wk, (size_t)6 * 3 * n); A(c.pc, n); A(c.part, (size_t)6 * RG); A(c.ks, 3); A(c.scal, 8);
  A(c.sweep_int, 4);
#undef A
#undef U
  if (rc) return rc;
  FVK_HIP(hipDeviceGetAttribute(&c.n_cu, hipDeviceAttributeMultiprocessorCount, c.device));
  std::vector<int> lrows;
  fl_levels(ia, ja, n, true, c.lf_ptr, lrows, c.h_levf);
  if ((rc = fl_up(c, &c.lf_rows, lrows))) return rc;
  fl_levels(ia, ja, n, false, c.lb_ptr, lrows, c.h_levb);
  if ((rc = fl_up(c, &c.lb_rows, lrows))) return rc;
  if (c.one_launch) {
    if ((rc = fl_parts_setup(c))) return rc;
    c.parts = true;
  }
  return 0;
}

static int fl_ctx_build(FlCtx &c, int nd, int n_vrt, const double *x3, int nb, const int *btag, int n,
                        const int *evp, const int *evr, int nf, const int *fce_elm, const int *fvp, const int *fvr,
                        const int *ngp, const int *ngc) {
  c.nd = nd; c.n_vrt = n_vrt; c.nb = nb; c.n = n; c.nf = nf;
  {
    const int rc = fl_csr_setup(c, n, ngp, ngc);
    if (rc) return rc;
  }
  // element and face numbering -> 0-based local indices
  std::vector<int> j1(nf), jn(nf);
  for (int f = 0; f < nf; ++f) {
    const int a = fce_elm[2 * f], b = fce_elm[2 * f + 1];
    if (a <= nb || a > nb + n) { set_error("face " + std::to_string(f + 1) + ": owner is not a domain element"); return FVK_ERR_ARG; }
    j1[f] = a - nb - 1;
    if (b > 0) {
      if (b <= nb || b > nb + n) { set_error("face " + std::to_string(f + 1) + ": bad neighbour"); return FVK_ERR_ARG; }
      jn[f] = b - nb - 1;
    } else {
      if (b == 0 || -b > nb) { set_error("face " + std::to_string(f + 1) + ": bad boundary element"); return FVK_ERR_ARG; }
      jn[f] = b;  // -(boundary element), i.e. -(b0 + 1) with b0 0-based
    }
  }
  const std::vector<int> &ia = c.h_ia, &ja = c.h_ja;
  auto csr_index = [&](int i, int j) {
    for (int k = ia[i]; k < ia[i + 1]; ++k)
      if (ja[k] == j) return k;
    return -1;
  };
  std::vector<int> kij(nf, -1), kji(nf, -1);
  for (int f = 0; f < nf; ++f)
    if (jn[f] >= 0) {
      kij[f] = csr_index(j1[f], jn[f]);
      kji[f] = csr_index(jn[f], j1[f]);
      if (kij[f] < 0 || kji[f] < 0) { set_error("matrix pattern lacks a face's entries"); return FVK_ERR_ARG; }
    }
  // cell face lists, ascending
  std::vector<int> cnt(n + 1, 0), cfl;
  for (int f = 0; f < nf; ++f) {
    cnt[j1[f] + 1]++;
    if (jn[f] >= 0) cnt[jn[f] + 1]++;
  }
  for (int i = 0; i < n; ++i) cnt[i + 1] += cnt[i];
  cfl.resize(cnt[n]);
  std::vector<int> fill(cnt.begin(), cnt.end() - 1);
  for (int f = 0; f < nf; ++f) {
    cfl[fill[j1[f]]++] = f;
    if (jn[f] >= 0) cfl[fill[jn[f]]++] = f;
  }
  std::vector<int> evp0(n + 1), evr0(evp[n] - 1), fvp0(nf + 1), fvr0(fvp[nf] - 1);
  for (int i = 0; i <= n; ++i) evp0[i] = evp[i] - 1;
  for (size_t k = 0; k < evr0.size(); ++k) {
    evr0[k] = evr[k] - 1;
    if (evr0[k] < 0 || evr0[k] >= n_vrt) { set_error("element vertex out of range"); return FVK_ERR_ARG; }
  }
  for (int f = 0; f <= nf; ++f) fvp0[f] = fvp[f] - 1;
  for (size_t k = 0; k < fvr0.size(); ++k) {
    fvr0[k] = fvr[k] - 1;
    if (fvr0[k] < 0 || fvr0[k] >= n_vrt) { set_error("face vertex out of range"); return FVK_ERR_ARG; }
  }
  for (int f = 0; f < nf; ++f)
    if (fvp0[f + 1] - fvp0[f] < 2 || (nd == 3 && fvp0[f + 1] - fvp0[f] < 3)) {
      set_error("face with too few vertices");
      return FVK_ERR_ARG;
    }
  c.h_j1 = j1; c.h_jn = jn;
  c.h_btag.assign(btag, btag + nb);
  int rc = 0;
#define A(ptr, cnt_) if (!rc) rc = fl_alloc(c, &ptr, cnt_)
#define U(ptr, vec) if (!rc) rc = fl_up(c, &ptr, vec)
  std::vector<double> xv(x3, x3 + 3 * (size_t)n_vrt);
  double *dxv = nullptr;
  int *devp = nullptr, *devr = nullptr, *dfvp = nullptr, *dfvr = nullptr;
  U(dxv, xv); U(devp, evp0); U(devr, evr0); U(dfvp, fvp0); U(dfvr, fvr0);
  U(c.j1, j1); U(c.jn, jn); U(c.btag, c.h_btag);
  U(c.kij, kij); U(c.kji, kji);
  U(c.cptr, cnt); U(c.cfl, cfl);
  A(c.xvc, (size_t)nd * n); A(c.vol, n); A(c.volr, n); A(c.ls, (size_t)nd * nd * n);
  A(c.xfc, (size_t)nd * nf); A(c.nrm, (size_t)nd * nf); A(c.area, nf); A(c.xpc, (size_t)nd * nf);
  A(c.xnc, (size_t)nd * nf); A(c.lpn, nf); A(c.wfp, nf); A(c.fvol, nf);
#undef A
#undef U
  if (rc) return rc;
  k_fl_centre<<<grid_for(n), BS, 0, c.stream>>>(n, nd, dxv, devp, devr, c.xvc);
  k_fl_face_geom<<<grid_for(nf), BS, 0, c.stream>>>(nf, n, nd, dxv, dfvp, dfvr, c.j1, c.jn, c.xvc, c.xfc, c.nrm,
                                                     c.area, c.xpc, c.xnc, c.lpn, c.wfp, c.fvol);
  k_fl_vol<<<grid_for(n), BS, 0, c.stream>>>(n, c.cptr, c.cfl, c.j1, c.fvol, c.vol, c.volr);
  k_fl_ls<<<grid_for(n), BS, 0, c.stream>>>(n, nf, nd, c.cptr, c.cfl, c.j1, c.jn, c.xvc, c.xfc, c.ls);
  FVK_HIP(hipGetLastError());
  FVK_HIP(hipStreamSynchronize(c.stream));
  return 0;
}

static void fl_ctx_free(FlCtx &c) {
  if (c.stream) hipStreamSynchronize(c.stream);
  for (void *p : c.owned) hipFree(p);
  c.owned.clear();
  if (c.stream) hipStreamDestroy(c.stream);
  c.stream = nullptr;
}

// device (component-major, SoA) <-> Fortran column-major a(ncmp, m)
static int fl_to_dev(hipStream_t s, const double *h, int ncmp, int m, double *d) {
  std::vector<double> t((size_t)ncmp * m);
  for (int i = 0; i < m; ++i)
    for (int q = 0; q < ncmp; ++q) t[(size_t)q * m + i] = h[(size_t)i * ncmp + q];
  FVK_HIP(hipStreamSynchronize(s));
  FVK_HIP(hipMemcpy(d, t.data(), t.size() * sizeof(double), hipMemcpyHostToDevice));
  return 0;
}
static int fl_to_host(hipStream_t s, const double *d, int ncmp, int m, double *h) {
  std::vector<double> t((size_t)ncmp * m);
  FVK_HIP(hipStreamSynchronize(s));
  FVK_HIP(hipMemcpy(t.data(), d, t.size() * sizeof(double), hipMemcpyDeviceToHost));
  for (int i = 0; i < m; ++i)
    for (int q = 0; q < ncmp; ++q) h[(size_t)i * ncmp + q] = t[(size_t)q * m + i];
  return 0;
}

static int ff_build(FlFlow &F) {
  FlCtx &c = *F.c;
  const size_t n = c.n, nb = std::max(c.nb, 1), nf = c.nf, nd = c.nd;
  int rc = 0;
#define A(ptr, cnt_) if (!rc) rc = ff_alloc(F, &ptr, cnt_)
  A(F.u, nd * n); A(F.ub, nd * nb); A(F.ug, nd * nd * n);
  A(F.p, n); A(F.pb, nb); A(F.pg, nd * n);
  A(F.q, n); A(F.qb, nb); A(F.qg, nd * n);
  A(F.rhs, nd * n); A(F.diag, n); A(F.val, (size_t)c.nnz); A(F.spd, nf);
  A(F.cc, n); A(F.ccb, nb); A(F.cd, n); A(F.cdb, nb);
  A(F.fa, nf); A(F.fb, nf); A(F.fd, nf); A(F.fcc, nd * nf); A(F.fdc, nd * nf);
#undef A
  if (rc) return rc;
  std::vector<int> bfl;
  for (int f = 0; f < c.nf; ++f)
    if (c.h_jn[f] < 0) bfl.push_back(f);
  F.n_bf = (int)bfl.size();
  {
    void *q = nullptr;
    FVK_HIP(hipMalloc(&q, std::max<size_t>(bfl.size(), 1) * sizeof(int)));
    F.owned.push_back(q);
    F.bfl = (int *)q;
    if (!bfl.empty()) FVK_HIP(hipMemcpy(F.bfl, bfl.data(), bfl.size() * sizeof(int), hipMemcpyHostToDevice));
  }
  const int g = 1024;
  k_fl_fill<<<g, BS, 0, c.stream>>>(n, F.cc, F.cfg.cc);
  k_fl_fill<<<g, BS, 0, c.stream>>>(nb, F.ccb, F.cfg.cc);
  k_fl_fill<<<g, BS, 0, c.stream>>>(n, F.cd, F.cfg.cd);
  k_fl_fill<<<g, BS, 0, c.stream>>>(nb, F.cdb, F.cfg.cd);
  FVK_HIP(hipGetLastError());
  FVK_HIP(hipStreamSynchronize(c.stream));
  return 0;
}

static void ff_free(FlFlow &F) {
  if (F.c && F.c->stream) hipStreamSynchronize(F.c->stream);
  for (void *p : F.owned) hipFree(p);
  F.owned.clear();
}

// the operators of one outer iteration (navierstokes.f90:518-620)
enum FlField { FL_VEL = 0, FL_PRES = 1, FL_PP = 2 };

static int ff_gradient(FlFlow &F, int which, int neu) {
  FlCtx &c = *F.c;
  if (which == FL_VEL) return fl_grad(c, c.nd, F.u, F.ub, F.ug, neu, F.bfl, F.n_bf);
  if (which == FL_PRES) return fl_grad(c, 1, F.p, F.pb, F.pg, neu, F.bfl, F.n_bf);
  return fl_grad(c, 1, F.q, F.qb, F.qg, neu, F.bfl, F.n_bf);
}

static int ff_zero(FlFlow &F) {
  FlCtx &c = *F.c;
  k_fl_fill<<<1024, BS, 0, c.stream>>>((size_t)c.nd * c.n, F.rhs, 0.0);
  k_fl_fill<<<1024, BS, 0, c.stream>>>((size_t)c.nnz, F.val, 0.0);
  FVK_HIP(hipGetLastError());
  return 0;
}

static int ff_source(FlFlow &F) {
  FlCtx &c = *F.c;
  k_fl_source<<<grid_for(c.n), BS, 0, c.stream>>>(c.n, c.nd, F.pg, c.vol, F.rhs);
  FVK_HIP(hipGetLastError());
  return 0;
}

static int ff_construct_pde(FlFlow &F) {
  FlCtx &c = *F.c;
  k_fl_pde_face<<<grid_for(c.nf), BS, 0, c.stream>>>(fl_geo(c), F.cfg.bl_c, F.u, F.ub, F.ug, F.spd, F.cc, F.ccb, F.cd,
                                                      F.cdb, F.val, F.fa, F.fb, F.fd, F.fcc, F.fdc);
  k_fl_pde_cell<<<grid_for(c.n), BS, 0, c.stream>>>(c.n, c.nf, c.nd, c.cptr, c.cfl, c.j1, c.jn, c.kii, F.fa, F.fb,
                                                     F.fd, F.fcc, F.fdc, F.u, F.cfg.urf_u, F.val, F.rhs, F.diag);
  FVK_HIP(hipGetLastError());
  return 0;
}

// FVK_ERR_HIP when a one-launch factorisation or sweep exceeded its wait
// bound (its results are then invalid)
static int fl_sweep_status(FlCtx &c) {
  int e = 0;
  FVK_HIP(hipMemcpyAsync(&e, c.sweep_int + 3, sizeof e, hipMemcpyDeviceToHost, c.stream));
  FVK_HIP(hipStreamSynchronize(c.stream));
  if (e) {
    set_error("triangular sweep: a wait exceeded its bound; the preconditioner's results are invalid");
    return FVK_ERR_HIP;
  }
  return 0;
}

static int ff_solve(FlFlow &F, int isol, int ncmp, double *phi, double *rhs, int n_iter, double reduc, double *res0) {
  FlCtx &c = *F.c;
  int rc = fl_solve(c, isol, ncmp, F.val, phi, rhs, n_iter, reduc);
  if (rc) return rc;
  FlKS k[3];
  FVK_HIP(hipMemcpyAsync(k, c.ks, sizeof(FlKS) * ncmp, hipMemcpyDeviceToHost, c.stream));
  FVK_HIP(hipStreamSynchronize(c.stream));
  for (int q = 0; q < ncmp; ++q) res0[q] = k[q].norm;
  return fl_sweep_status(c);
}

static int ff_construct_pp(FlFlow &F) {
  FlCtx &c = *F.c;
  k_fl_pp_face<<<grid_for(c.nf), BS, 0, c.stream>>>(fl_geo(c), F.u, F.ub, F.ug, F.p, F.pg, F.cc, F.ccb, F.diag,
                                                     F.spd, F.val, F.fa, F.fd, F.fdc);
  const int P = std::max(1, std::min(RG, (F.n_bf + RB - 1) / RB));
  k_fl_inout<<<P, RB, 0, c.stream>>>(F.n_bf, F.bfl, c.jn, c.btag, F.ccb, c.area, F.spd, c.part);
  k_fl_crat<<<1, RB, 0, c.stream>>>(P, c.part, c.scal);
  if (F.n_bf) k_fl_outlet<<<grid_for(F.n_bf), BS, 0, c.stream>>>(F.n_bf, F.bfl, c.jn, c.btag, c.scal, F.spd);
  k_fl_pp_cell<<<grid_for(c.n), BS, 0, c.stream>>>(c.n, c.nb, c.cptr, c.cfl, c.j1, c.jn, c.kii, c.btag, F.ccb, c.area,
                                                    F.spd, F.fa, F.fd, F.fdc, F.val, F.rhs);
  FVK_HIP(hipGetLastError());
  return 0;
}

static int ff_correct(FlFlow &F) {
  FlCtx &c = *F.c;
  k_fl_correct<<<grid_for(c.n), BS, 0, c.stream>>>(c.n, c.nd, F.qg, c.vol, F.diag, F.q, 0.0, F.cfg.urf_p, F.u, F.p);
  FVK_HIP(hipGetLastError());
  return 0;
}

// one outer iteration; res0[0..3]: the momentum and pressure-correction norms
static int ff_outer(FlFlow &F, double *res0) {
  FlCtx &c = *F.c;
  int rc;
  if ((rc = ff_zero(F))) return rc;
  if ((rc = ff_gradient(F, FL_PRES, 1))) return rc;
  if ((rc = ff_gradient(F, FL_VEL, 0))) return rc;
  if ((rc = ff_source(F))) return rc;
  if ((rc = ff_construct_pde(F))) return rc;
  if ((rc = ff_solve(F, 2, c.nd, F.u, F.rhs, F.cfg.n_iter_u, F.cfg.reduc_u, res0))) return rc;
  if ((rc = ff_zero(F))) return rc;
  if ((rc = ff_construct_pp(F))) return rc;
  k_fl_fill<<<1024, BS, 0, c.stream>>>((size_t)c.n, F.q, 0.0);
  if ((rc = ff_solve(F, 1, 1, F.q, F.rhs, F.cfg.n_iter_p, F.cfg.reduc_p, res0 + 3))) return rc;
  if ((rc = ff_gradient(F, FL_PP, 1))) return rc;
  if ((rc = ff_correct(F))) return rc;
  for (int k = 0; k < 4; ++k) F.res0[k] = res0[k];
  return 0;
}

}  // namespace fvk

// ---------------------------------------------------------------- C ABI
struct fvk_fl_mesh {
  fvk::FlMesh m;
};
struct fvk_fl_ctx {
  fvk::FlCtx c;
};
struct fvk_fl_flow {
  fvk::FlFlow f;
};

using namespace fvk;

extern "C" {

int fvk_fl_mesh_read_gmsh(const char *path, int n_dim, double l_sc, fvk_fl_mesh **out) {
  if (!path || !out) { set_error("null argument"); return FVK_ERR_ARG; }
  GmshRaw g;
  std::string e = parse_gmsh(g, path);
  if (!e.empty()) { set_error(e); return FVK_ERR_IO; }
  fvk_fl_mesh *m = new fvk_fl_mesh;
  e = fl_mesh_from_gmsh(m->m, g, n_dim, l_sc);
  if (!e.empty()) { delete m; set_error(e); return FVK_ERR_ARG; }
  *out = m;
  return FVK_OK;
}

int fvk_fl_mesh_from_records(int n_dim, int n_node, const double *x3, int n_elm, const int *e_ty, const int *e_tag,
                             const int *e_vptr, const int *e_vrt, fvk_fl_mesh **out) {
  if (!x3 || !e_ty || !e_tag || !e_vptr || !e_vrt || !out || n_node < 1 || n_elm < 1) {
    set_error("bad argument");
    return FVK_ERR_ARG;
  }
  static const int kNvrt[8] = {0, 2, 3, 4, 4, 8, 6, 5};
  GmshRaw g;
  g.node_id.resize(n_node);
  for (int i = 0; i < n_node; ++i) g.node_id[i] = i + 1;
  g.node_x.assign(x3, x3 + 3 * (size_t)n_node);
  g.e_vptr.assign(1, 0);
  for (int e = 0; e < n_elm; ++e) {
    if (e_ty[e] < 1 || e_ty[e] > 7 || e_vptr[e + 1] - e_vptr[e] != kNvrt[e_ty[e]]) {
      set_error("element " + std::to_string(e + 1) + ": bad type or vertex count");
      return FVK_ERR_ARG;
    }
    g.e_idx.push_back(e + 1);
    g.e_ty.push_back(e_ty[e]);
    g.e_tag0.push_back(e_tag[e]);
    for (int k = e_vptr[e]; k < e_vptr[e + 1]; ++k) g.e_vrt.push_back(e_vrt[k]);
    g.e_vptr.push_back((int)g.e_vrt.size());
  }
  fvk_fl_mesh *m = new fvk_fl_mesh;
  const std::string e = fl_mesh_from_gmsh(m->m, g, n_dim, 1.0);
  if (!e.empty()) { delete m; set_error(e); return FVK_ERR_ARG; }
  *out = m;
  return FVK_OK;
}

int fvk_fl_mesh_get_info(const fvk_fl_mesh *mm, int *info) {
  if (!mm || !info) { set_error("null argument"); return FVK_ERR_ARG; }
  const FlMesh &m = mm->m;
  info[0] = m.n_dim; info[1] = m.n_vrt; info[2] = m.n_bnd; info[3] = m.n_dmn; info[4] = m.n_fce;
  info[5] = (int)m.ng_col.size(); info[6] = (int)m.elm_vrt.size(); info[7] = (int)m.fce_vrt.size();
  return FVK_OK;
}

int fvk_fl_mesh_get_arrays(const fvk_fl_mesh *mm, double *x3, int *elm_vptr, int *elm_vrt, int *bnd_tag,
                           int *fce_elm, int *fce_vptr, int *fce_vrt, int *ng_ptr, int *ng_col, int *elm_idx) {
  if (!mm) { set_error("null argument"); return FVK_ERR_ARG; }
  const FlMesh &m = mm->m;
  auto cp = [](const auto &v, auto *dst) { if (dst) std::copy(v.begin(), v.end(), dst); };
  cp(m.x, x3); cp(m.elm_vptr, elm_vptr); cp(m.elm_vrt, elm_vrt); cp(m.bnd_tag, bnd_tag); cp(m.fce_elm, fce_elm);
  cp(m.fce_vptr, fce_vptr); cp(m.fce_vrt, fce_vrt); cp(m.ng_ptr, ng_ptr); cp(m.ng_col, ng_col);
  cp(m.elm_idx, elm_idx);
  return FVK_OK;
}

void fvk_fl_mesh_free(fvk_fl_mesh *m) { delete m; }

int fvk_fl_ctx_create(int n_dim, int n_vrt, const double *x3, int n_bnd, const int *bnd_tag, int n_dmn,
                      const int *elm_vptr, const int *elm_vrt, int n_fce, const int *fce_elm, const int *fce_vptr,
                      const int *fce_vrt, const int *ng_ptr, const int *ng_col, int device, fvk_fl_ctx **out) {
  if (!x3 || !elm_vptr || !elm_vrt || !fce_elm || !fce_vptr || !fce_vrt || !ng_ptr || !ng_col || !out ||
      (n_bnd > 0 && !bnd_tag) || (n_dim != 2 && n_dim != 3) || n_dmn < 1 || n_fce < 1 || n_vrt < 1 || n_bnd < 0) {
    set_error("bad argument");
    return FVK_ERR_ARG;
  }
  fvk_fl_ctx *cc = new fvk_fl_ctx;
  FlCtx &c = cc->c;
  c.device = device;
  hipError_t e = hipSetDevice(device);
  if (e == hipSuccess) e = hipStreamCreateWithFlags(&c.stream, hipStreamNonBlocking);
  if (e != hipSuccess) { delete cc; set_error(hipGetErrorString(e)); return FVK_ERR_HIP; }
  const int rc = fl_ctx_build(c, n_dim, n_vrt, x3, n_bnd, bnd_tag, n_dmn, elm_vptr, elm_vrt, n_fce, fce_elm, fce_vptr,
                              fce_vrt, ng_ptr, ng_col);
  if (rc) { fl_ctx_free(c); delete cc; return rc; }
  *out = cc;
  return FVK_OK;
}

int fvk_fl_csr_create(int n, const int *ia, const int *ja, int device, fvk_fl_ctx **out) {
  if (!ia || !ja || !out || n < 1) { set_error("bad argument"); return FVK_ERR_ARG; }
  fvk_fl_ctx *cc = new fvk_fl_ctx;
  FlCtx &c = cc->c;
  c.device = device;
  hipError_t e = hipSetDevice(device);
  if (e == hipSuccess) e = hipStreamCreateWithFlags(&c.stream, hipStreamNonBlocking);
  if (e != hipSuccess) { delete cc; set_error(hipGetErrorString(e)); return FVK_ERR_HIP; }
  const int rc = fl_csr_setup(c, n, ia, ja);
  if (rc) { fl_ctx_free(c); delete cc; return rc; }
  *out = cc;
  return FVK_OK;
}

int fvk_fl_csr_matches(const fvk_fl_ctx *cc, int n, const int *ia, const int *ja) {
  if (!cc || !ia || !ja) return 0;
  const FlCtx &c = cc->c;
  if (n != c.n || ia[n] - 1 != c.nnz) return 0;
  for (int i = 0; i <= n; ++i)
    if (ia[i] - 1 != c.h_ia[i]) return 0;
  for (int k = 0; k < c.nnz; ++k)
    if (ja[k] - 1 != c.h_ja[k]) return 0;
  return 1;
}

void fvk_fl_ctx_destroy(fvk_fl_ctx *cc) {
  if (!cc) return;
  fl_ctx_free(cc->c);
  delete cc;
}

int fvk_fl_geometry(fvk_fl_ctx *cc, double *x_vc, double *vol, double *x_fc, double *norm, double *area, double *x_pc,
                    double *x_nc, double *l_pn, double *w_fp, double *ls) {
  if (!cc) { set_error("null argument"); return FVK_ERR_ARG; }
  FlCtx &c = cc->c;
  const int nd = c.nd;
  int rc = 0;
  if (x_vc && (rc = fl_to_host(c.stream, c.xvc, nd, c.n, x_vc))) return rc;
  if (vol && (rc = fl_to_host(c.stream, c.vol, 1, c.n, vol))) return rc;
  if (x_fc && (rc = fl_to_host(c.stream, c.xfc, nd, c.nf, x_fc))) return rc;
  if (norm && (rc = fl_to_host(c.stream, c.nrm, nd, c.nf, norm))) return rc;
  if (area && (rc = fl_to_host(c.stream, c.area, 1, c.nf, area))) return rc;
  if (x_pc && (rc = fl_to_host(c.stream, c.xpc, nd, c.nf, x_pc))) return rc;
  if (x_nc && (rc = fl_to_host(c.stream, c.xnc, nd, c.nf, x_nc))) return rc;
  if (l_pn && (rc = fl_to_host(c.stream, c.lpn, 1, c.nf, l_pn))) return rc;
  if (w_fp && (rc = fl_to_host(c.stream, c.wfp, 1, c.nf, w_fp))) return rc;
  if (ls && (rc = fl_to_host(c.stream, c.ls, nd * nd, c.n, ls))) return rc;
  return FVK_OK;
}

int fvk_fl_ls_gradient(fvk_fl_ctx *cc, int n_cmp, const double *phi, double *phi_bnd, double *grad, int neu_bnd) {
  if (!cc || !phi || !grad || (!phi_bnd && cc->c.nb > 0) || n_cmp < 1 || n_cmp > 3) {
    set_error("bad argument");
    return FVK_ERR_ARG;
  }
  FlCtx &c = cc->c;
  double *d = wkv(c, 0), *db = nullptr, *dg = wkv(c, 2);
  void *q = nullptr;
  FVK_HIP(hipMalloc(&q, sizeof(double) * 3 * (size_t)std::max(c.nb, 1)));
  db = (double *)q;
  std::vector<int> bfl;
  for (int f = 0; f < c.nf; ++f)
    if (c.h_jn[f] < 0) bfl.push_back(f);
  int *dbfl = nullptr;
  int rc = 0;
  if (hipMalloc(&q, sizeof(int) * std::max<size_t>(bfl.size(), 1)) != hipSuccess) {
    hipFree(db);
    set_error("hipMalloc failed");
    return FVK_ERR_HIP;
  }
  dbfl = (int *)q;
  if (!bfl.empty()) hipMemcpy(dbfl, bfl.data(), bfl.size() * sizeof(int), hipMemcpyHostToDevice);
  if (!rc) rc = fl_to_dev(c.stream, phi, n_cmp, c.n, d);
  if (!rc && c.nb) rc = fl_to_dev(c.stream, phi_bnd, n_cmp, c.nb, db);
  if (!rc) rc = fl_grad(c, n_cmp, d, db, dg, neu_bnd, dbfl, (int)bfl.size());
  if (!rc) rc = fl_to_host(c.stream, dg, c.nd * n_cmp, c.n, grad);
  if (!rc && neu_bnd && c.nb) rc = fl_to_host(c.stream, db, n_cmp, c.nb, phi_bnd);
  hipFree(dbfl);
  hipFree(db);
  return rc;
}

int fvk_fl_lin_eqns_solver(fvk_fl_ctx *cc, const double *a_ij, double *phi, double *rhs, double *norm, int n_iter,
                           double reduc, int isol) {
  if (!cc || !a_ij || !phi || !rhs || !norm || n_iter < 0 || (isol != 1 && isol != 2)) {
    set_error("bad argument");
    return FVK_ERR_ARG;
  }
  FlCtx &c = cc->c;
  void *q = nullptr;
  FVK_HIP(hipMalloc(&q, sizeof(double) * ((size_t)c.nnz + 2 * (size_t)c.n)));
  double *da = (double *)q, *dphi = da + c.nnz, *drhs = dphi + c.n;
  hipError_t e = hipStreamSynchronize(c.stream);
  if (e == hipSuccess) e = hipMemcpy(da, a_ij, sizeof(double) * c.nnz, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(dphi, phi, sizeof(double) * c.n, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(drhs, rhs, sizeof(double) * c.n, hipMemcpyHostToDevice);
  int rc = e == hipSuccess ? fl_solve(c, isol, 1, da, dphi, drhs, n_iter, reduc) : FVK_ERR_HIP;
  if (e != hipSuccess) set_error(hipGetErrorString(e));
  FlKS k;
  if (!rc) {
    e = hipMemcpyAsync(&k, c.ks, sizeof k, hipMemcpyDeviceToHost, c.stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c.stream);
    if (e == hipSuccess) e = hipMemcpy(phi, dphi, sizeof(double) * c.n, hipMemcpyDeviceToHost);
    if (e == hipSuccess) e = hipMemcpy(rhs, drhs, sizeof(double) * c.n, hipMemcpyDeviceToHost);
    if (e != hipSuccess) { set_error(hipGetErrorString(e)); rc = FVK_ERR_HIP; }
    else *norm = k.norm;
  }
  hipFree(q);
  return rc ? rc : fl_sweep_status(c);
}

int fvk_fl_ctx_set_param(fvk_fl_ctx *cc, const char *name, double value) {
  if (!cc || !name) { set_error("null argument"); return FVK_ERR_ARG; }
  FlCtx &c = cc->c;
  if (!std::strcmp(name, "sweep_one_launch") && (value == 0 || value == 1)) {
    if (value == 1) {
      int mx = 0;
      for (int i = 0; i < c.n; ++i) {
        int lo = c.h_ia[i], up = c.h_ia[i + 1];
        while (lo < c.h_ia[i + 1] && c.h_ja[lo] < i) ++lo;
        while (up > c.h_ia[i] && c.h_ja[up - 1] > i) --up;
        mx = std::max({mx, lo - c.h_ia[i], c.h_ia[i + 1] - up});
      }
      if (mx > kSweepSlots) { set_error("a row has more lower or upper entries than one launch holds"); return FVK_ERR_ARG; }
    }
    c.one_launch = value != 0;
    return FVK_OK;
  }
  if (!std::strcmp(name, "sweep_parts") && (value == 0 || value == 1)) {
    if (value == 1 && !c.pf.hd) { set_error("a row has more lower or upper entries than one launch holds"); return FVK_ERR_ARG; }
    c.parts = value != 0;
    return FVK_OK;
  }
  if (!std::strcmp(name, "sweep_part_rows") && value >= 256 && value <= 6400 && value == (int)value) {
    if (!c.pf.hd) { set_error("a row has more lower or upper entries than one launch holds"); return FVK_ERR_ARG; }
    FVK_HIP(hipStreamSynchronize(c.stream));
    c.part_rows = (int)value;
    return fl_parts_setup(c);
  }
  if (!std::strcmp(name, "sweep_blocks") && value >= 0 && value <= 65536 && value == (int)value) {
    c.sweep_blocks = (int)value;
    return FVK_OK;
  }
  if (!std::strcmp(name, "sweep_fail") && value == 1 && c.sweep_int) {
    const int one = 1;
    FVK_HIP(hipMemcpyAsync(c.sweep_int + 3, &one, sizeof one, hipMemcpyHostToDevice, c.stream));
    FVK_HIP(hipStreamSynchronize(c.stream));
    return FVK_OK;
  }
  set_error(std::string("unknown parameter or value out of range: ") + name);
  return FVK_ERR_ARG;
}

int fvk_fl_flow_create(fvk_fl_ctx *cc, const double *params, fvk_fl_flow **out) {
  if (!cc || !params || !out) { set_error("null argument"); return FVK_ERR_ARG; }
  FlCfg cfg;
  cfg.cc = params[0]; cfg.cd = params[1]; cfg.urf_u = params[2]; cfg.urf_p = params[3]; cfg.bl_c = params[4];
  cfg.n_iter_u = (int)params[5]; cfg.reduc_u = params[6]; cfg.n_iter_p = (int)params[7]; cfg.reduc_p = params[8];
  if (!(cfg.urf_u > 0.0) || cfg.n_iter_u < 0 || cfg.n_iter_p < 0) { set_error("bad flow parameters"); return FVK_ERR_ARG; }
  fvk_fl_flow *ff = new fvk_fl_flow;
  ff->f.c = &cc->c;
  ff->f.cfg = cfg;
  const int rc = ff_build(ff->f);
  if (rc) { ff_free(ff->f); delete ff; return rc; }
  *out = ff;
  return FVK_OK;
}

void fvk_fl_flow_destroy(fvk_fl_flow *ff) {
  if (!ff) return;
  ff_free(ff->f);
  delete ff;
}

int fvk_fl_flow_set(fvk_fl_flow *ff, const double *vel, const double *vel_bnd, const double *pres,
                    const double *pres_bnd) {
  if (!ff) { set_error("null argument"); return FVK_ERR_ARG; }
  FlFlow &F = ff->f;
  FlCtx &c = *F.c;
  int rc = 0;
  if (vel && (rc = fl_to_dev(c.stream, vel, c.nd, c.n, F.u))) return rc;
  if (vel_bnd && c.nb && (rc = fl_to_dev(c.stream, vel_bnd, c.nd, c.nb, F.ub))) return rc;
  if (pres && (rc = fl_to_dev(c.stream, pres, 1, c.n, F.p))) return rc;
  if (pres_bnd && c.nb && (rc = fl_to_dev(c.stream, pres_bnd, 1, c.nb, F.pb))) return rc;
  return FVK_OK;
}

int fvk_fl_flow_get(fvk_fl_flow *ff, double *vel, double *vel_bnd, double *pres, double *pres_bnd, double *pp,
                    double *spd) {
  if (!ff) { set_error("null argument"); return FVK_ERR_ARG; }
  FlFlow &F = ff->f;
  FlCtx &c = *F.c;
  int rc = 0;
  if (vel && (rc = fl_to_host(c.stream, F.u, c.nd, c.n, vel))) return rc;
  if (vel_bnd && c.nb && (rc = fl_to_host(c.stream, F.ub, c.nd, c.nb, vel_bnd))) return rc;
  if (pres && (rc = fl_to_host(c.stream, F.p, 1, c.n, pres))) return rc;
  if (pres_bnd && c.nb && (rc = fl_to_host(c.stream, F.pb, 1, c.nb, pres_bnd))) return rc;
  if (pp && (rc = fl_to_host(c.stream, F.q, 1, c.n, pp))) return rc;
  if (spd && (rc = fl_to_host(c.stream, F.spd, 1, c.nf, spd))) return rc;
  return FVK_OK;
}

// intermediate arrays of the current iteration: the matrix values (nnz),
// the right-hand side (n_dim x n_dmn), 1/a_P (n_dmn), the gradients of
// velocity, pressure and pressure correction
int fvk_fl_flow_get_system(fvk_fl_flow *ff, double *val, double *rhs, double *diag, double *vel_grad,
                           double *pres_grad, double *pp_grad, double *pp_bnd) {
  if (!ff) { set_error("null argument"); return FVK_ERR_ARG; }
  FlFlow &F = ff->f;
  FlCtx &c = *F.c;
  int rc = 0;
  if (val) {
    FVK_HIP(hipStreamSynchronize(c.stream));
    FVK_HIP(hipMemcpy(val, F.val, sizeof(double) * c.nnz, hipMemcpyDeviceToHost));
  }
  if (rhs && (rc = fl_to_host(c.stream, F.rhs, c.nd, c.n, rhs))) return rc;
  if (diag && (rc = fl_to_host(c.stream, F.diag, 1, c.n, diag))) return rc;
  if (vel_grad && (rc = fl_to_host(c.stream, F.ug, c.nd * c.nd, c.n, vel_grad))) return rc;
  if (pres_grad && (rc = fl_to_host(c.stream, F.pg, c.nd, c.n, pres_grad))) return rc;
  if (pp_grad && (rc = fl_to_host(c.stream, F.qg, c.nd, c.n, pp_grad))) return rc;
  if (pp_bnd && c.nb && (rc = fl_to_host(c.stream, F.qb, 1, c.nb, pp_bnd))) return rc;
  return FVK_OK;
}

// One step of navierstokes' outer loop, by name of the reference routine
// it replaces: 0 zero (rhs and matrix), 1 ls_gradient(pres, neu 1),
// 2 ls_gradient(vel, neu 0), 3 ns_source, 4 construct_pde, 5 momentum
// solves (res0[0..n_dim)), 6 construct_pp, 7 pressure-correction solve
// (res0[0]), 8 ls_gradient(pp, neu 1), 9 ns_correction
int fvk_fl_flow_step(fvk_fl_flow *ff, int step, double *res0) {
  if (!ff || step < 0 || step > 9 || ((step == 5 || step == 7) && !res0)) { set_error("bad argument"); return FVK_ERR_ARG; }
  FlFlow &F = ff->f;
  FlCtx &c = *F.c;
  int rc = 0;
  switch (step) {
    case 0: rc = ff_zero(F); break;
    case 1: rc = ff_gradient(F, FL_PRES, 1); break;
    case 2: rc = ff_gradient(F, FL_VEL, 0); break;
    case 3: rc = ff_source(F); break;
    case 4: rc = ff_construct_pde(F); break;
    case 5: rc = ff_solve(F, 2, c.nd, F.u, F.rhs, F.cfg.n_iter_u, F.cfg.reduc_u, res0); break;
    case 6: rc = ff_construct_pp(F); break;
    case 7:
      k_fl_fill<<<1024, BS, 0, c.stream>>>((size_t)c.n, F.q, 0.0);
      rc = ff_solve(F, 1, 1, F.q, F.rhs, F.cfg.n_iter_p, F.cfg.reduc_p, res0);
      break;
    case 8: rc = ff_gradient(F, FL_PP, 1); break;
    case 9: rc = ff_correct(F); break;
  }
  if (!rc) FVK_HIP(hipStreamSynchronize(c.stream));
  return rc;
}

// n outer iterations of navierstokes.f90:515-641 with fp_iter_ctrl's exit
// rule (:806-845) for n_iter and res_drop; res0: 4 norms per iteration,
// norm_gl: the normalised residual per iteration; *n_done: iterations run
int fvk_fl_flow_run(fvk_fl_flow *ff, int n_iter, double res_drop, double *res0, double *norm_gl, int *n_done) {
  if (!ff || n_iter < 1 || !n_done) { set_error("bad argument"); return FVK_ERR_ARG; }
  FlFlow &F = ff->f;
  double norm_mx = 0.0;
  int fp_exit = 0, it;
  for (it = 1; it <= n_iter; ++it) {
    double r[4];
    int rc = ff_outer(F, r);
    if (rc) return rc;
    if (res0)
      for (int k = 0; k < 4; ++k) res0[4 * (it - 1) + k] = r[k];
    double gl = r[0];
    for (int k = 1; k < 4; ++k) gl = std::max(gl, r[k]);
    gl = std::max(gl, 0.0);  // res0(5:10) are zero
    norm_mx = std::max(norm_mx, gl);
    gl = gl / (norm_mx + SMALL);
    if (norm_gl) norm_gl[it - 1] = gl;
    if (gl < res_drop || it > n_iter - 2) ++fp_exit;
    if (fp_exit > 1) break;
  }
  *n_done = std::min(it, n_iter);
  return FVK_OK;
}

}  // extern "C"
