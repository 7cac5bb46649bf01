// Krylov layer on gfx950: SpMV over the slot-major ELL table, the Jacobi /
// DIC / DILU factor and its level-scheduled triangular sweeps, and the
// reference's PCG and BiCGStab recurrences with every scalar kept on the
// device.
//
//   solv_res      spray-simulation/solv_res.f90:5-30
//   solv_precond  spray-simulation/solv_precond.f90:5-30
//   solv_direct   spray-simulation/solv_direct.f90:5-43
//   solv_cg       spray-simulation/solv_cg.f90:7-87
//   solv_bicg     spray-simulation/solv_bicg.f90:7-121
//
// Row sums run over a row's faces in ascending face index, the order in
// which the reference's face-scatter loops reach that row, so SpMV,
// residual and the sweeps reproduce the reference bit for bit (compiled
// without FMA contraction).  Dot products and L1 norms are deterministic
// tree reductions (block partials, then the last block adds them in block
// order); they differ from the reference's sequential sums by rounding only.
#include <cstring>

#include "ctx.h"

namespace fvk {

// ---------------------------------------------------------------- kernels

__global__ __launch_bounds__(BS) void k_precond_jacobi(int n, const double *__restrict__ diag, double *__restrict__ c,
                                 double *__restrict__ cr) {
  int i = blockIdx.x * BS + threadIdx.x;
  if (i >= n) return;
  double d = diag[i];
  cr[i] = d;
  c[i] = 1.0 / (d + SMALL);
}

// One level of the DIC (imat 1) / DILU (imat 2) factorisation.
__global__ __launch_bounds__(BS) void k_precond_level(DMesh M, const int *__restrict__ rows, int cnt, int imat,
                                const double *__restrict__ diag, const double *__restrict__ ngp,
                                const double *__restrict__ ngn, double *__restrict__ c,
                                double *__restrict__ cr) {
  int t = blockIdx.x * BS + threadIdx.x;
  if (t >= cnt) return;
  const int r = rows[t], n = M.n, rr = M.rank[r];
  double d = diag[r];
  for (int k = 0; k < M.Ws; ++k) {
    const int j = M.spc[k * n + r];
    if (j < 0) break;
    if (M.rank[j] < rr) {
      const int f = M.spf[k * n + r];
      if (imat == 1) d = d - ngn[f] * ngn[f] * c[j];
      else d = d - ngp[f] * c[j] * ngn[f];
    }
  }
  cr[r] = d;
  c[r] = 1.0 / (d + SMALL);
}

__device__ __forceinline__ bool all_done(const KS *ks, int nvc) {
  if (!ks) return false;
  for (int q = 0; q < nvc; ++q)
    if (!ks[q].done) return false;
  return true;
}

__global__ __launch_bounds__(BS) void k_jacobi_apply(int n, int nvc, const double *__restrict__ c, const double *__restrict__ b,
                               double *__restrict__ x, const KS *ks) {
  int i = blockIdx.x * BS + threadIdx.x;
  if (i >= n || all_done(ks, nvc)) return;
  for (int q = 0; q < nvc; ++q) x[q * n + i] = b[q * n + i] * c[i];
}

// Forward sweep over one level: x = (b - L x) * c   (solv_direct.f90:12-24)
__global__ __launch_bounds__(BS) void k_sweep_fwd(DMesh M, const int *__restrict__ rows, int cnt, int nvc,
                            const double *__restrict__ spv, const double *__restrict__ c,
                            const double *__restrict__ b, double *__restrict__ x, const KS *ks) {
  int t = blockIdx.x * BS + threadIdx.x;
  if (t >= cnt || all_done(ks, nvc)) return;
  const int r = rows[t], n = M.n, rr = M.rank[r];
  double acc[3];
  for (int q = 0; q < nvc; ++q) acc[q] = b[q * n + r];
  for (int k = 0; k < M.Ws; ++k) {
    const int j = M.spc[k * n + r];
    if (j < 0) break;
    if (M.rank[j] < rr) {
      const double a = spv[k * n + r];
      for (int q = 0; q < nvc; ++q) acc[q] = acc[q] - a * x[q * n + j];
    }
  }
  const double cc = c[r];
  for (int q = 0; q < nvc; ++q) x[q * n + r] = acc[q] * cc;
}

// Backward sweep over one level: x = (x*c_r - U x) * c   (solv_direct.f90:26-38)
__global__ __launch_bounds__(BS) void k_sweep_bwd(DMesh M, const int *__restrict__ rows, int cnt, int nvc,
                            const double *__restrict__ spv, const double *__restrict__ c,
                            const double *__restrict__ cr, double *__restrict__ x, const KS *ks) {
  int t = blockIdx.x * BS + threadIdx.x;
  if (t >= cnt || all_done(ks, nvc)) return;
  const int r = rows[t], n = M.n, rr = M.rank[r];
  double acc[3];
  const double crr = cr[r];
  for (int q = 0; q < nvc; ++q) acc[q] = x[q * n + r] * crr;
  for (int k = 0; k < M.Ws; ++k) {
    const int j = M.spc[k * n + r];
    if (j < 0) break;
    if (M.rank[j] > rr) {
      const double a = spv[k * n + r];
      for (int q = 0; q < nvc; ++q) acc[q] = acc[q] - a * x[q * n + j];
    }
  }
  const double cc = c[r];
  for (int q = 0; q < nvc; ++q) x[q * n + r] = acc[q] * cc;
}

// y = diag*x + sum over the row's interior faces (solv_cg.f90:56-64)
__device__ __forceinline__ double row_spmv(const DMesh &M, const double *__restrict__ diag,
                                           const double *__restrict__ spv, const double *__restrict__ x,
                                           int i) {
  const int n = M.n;
  double y = diag[i] * x[i];
  for (int k = 0; k < M.Ws; ++k) {
    const int j = M.spc[k * n + i];
    if (j < 0) break;
    y = y + spv[k * n + i] * x[j];
  }
  return y;
}

__global__ __launch_bounds__(BS) void k_spmv(DMesh M, int nvc, const double *__restrict__ diag, const double *__restrict__ spv,
                       const double *__restrict__ x, double *__restrict__ y) {
  int i = blockIdx.x * BS + threadIdx.x;
  if (i >= M.n) return;
  for (int q = 0; q < nvc; ++q) y[q * M.n + i] = row_spmv(M, diag, spv, x + q * M.n, i);
}

// r = b - A x (or r = b), L1 norm per component, Krylov scalar set-up.
// kind: 0 = CG, 1 = BiCGStab.  (solv_res.f90, solv_cg.f90:19-33,
// solv_bicg.f90:19-38)
__global__ __launch_bounds__(BS) void k_residual(DMesh M, int nvc, const double *__restrict__ diag,
                           const double *__restrict__ spv, const double *__restrict__ x,
                           double *__restrict__ r, int not_r0, Red R, KS *ks) {
  const int n = M.n;
  double l1[3] = {0.0, 0.0, 0.0};
  for (int i = blockIdx.x * BS + threadIdx.x; i < n; i += gridDim.x * BS) {
    for (int q = 0; q < nvc; ++q) {
      double v = r[q * n + i];
      if (!not_r0) {
        const double *xq = x + q * n;
        v = v - diag[i] * xq[i];
        for (int k = 0; k < M.Ws; ++k) {
          const int j = M.spc[k * n + i];
          if (j < 0) break;
          v = v - spv[k * n + i] * xq[j];
        }
        r[q * n + i] = v;
      }
      l1[q] += fabs(v);
    }
  }
  if (grid_sum<3>(l1, R)) {
    double mx = 0.0;
    for (int q = 0; q < nvc; ++q) {
      ks[q].res0_n = l1[q];
      mx = q == 0 ? l1[q] : fmax(mx, l1[q]);
    }
    const int stop = mx < SMALL;  // solv_cg.f90:21
    for (int q = 0; q < nvc; ++q) ks[q].done = stop;
  }
}

// ---- PCG ----
// z = M^-1 r (Jacobi: computed here) and s = r.z; beta = (1/s_old)*s
__global__ __launch_bounds__(BS) void k_cg_rz(int n, int nvc, const double *__restrict__ r, const double *__restrict__ c,
                        int jacobi, double *__restrict__ z, Red R, KS *ks) {
  if (all_done(ks, nvc)) return;
  double rz[3] = {0.0, 0.0, 0.0};
  for (int i = blockIdx.x * BS + threadIdx.x; i < n; i += gridDim.x * BS)
    for (int q = 0; q < nvc; ++q) {
      if (ks[q].done) continue;
      double ri = r[q * n + i], zi;
      if (jacobi) {
        zi = ri * c[i];
        z[q * n + i] = zi;
      } else {
        zi = z[q * n + i];
      }
      rz[q] += ri * zi;
    }
  if (grid_sum<3>(rz, R))
    for (int q = 0; q < nvc; ++q) {
      KS &k = ks[q];
      if (k.done) continue;
      k.beta = 1.0 / k.s;
      k.s = rz[q];
      k.beta = k.beta * k.s;
    }
}

// p = z + beta p ; q = A p ; alpha = s / (p.q + small).  p of the
// neighbours is formed on the fly from (z, p_old), so p_new is written once
// and no extra pass over p is needed.
__global__ __launch_bounds__(BS) void k_cg_pq(DMesh M, int nvc, const double *__restrict__ diag, const double *__restrict__ spv,
                        const double *__restrict__ z, const double *__restrict__ pold, double *__restrict__ pnew,
                        double *__restrict__ qv, Red R, KS *ks) {
  if (all_done(ks, nvc)) return;
  const int n = M.n;
  double pq[3] = {0.0, 0.0, 0.0};
  double beta[3];
  for (int q = 0; q < nvc; ++q) beta[q] = ks[q].beta;
  for (int i = blockIdx.x * BS + threadIdx.x; i < n; i += gridDim.x * BS)
    for (int q = 0; q < nvc; ++q) {
      if (ks[q].done) continue;
      const double *zq = z + q * n, *po = pold + q * n;
      const double pi = zq[i] + beta[q] * po[i];
      double y = diag[i] * pi;
      for (int k = 0; k < M.Ws; ++k) {
        const int j = M.spc[k * n + i];
        if (j < 0) break;
        y = y + spv[k * n + i] * (zq[j] + beta[q] * po[j]);
      }
      pnew[q * n + i] = pi;
      qv[q * n + i] = y;
      pq[q] += pi * y;
    }
  if (grid_sum<3>(pq, R))
    for (int q = 0; q < nvc; ++q) {
      KS &k = ks[q];
      if (k.done) continue;
      k.alpha = k.s / (pq[q] + SMALL);
    }
}

// x += alpha p ; r -= alpha q ; res_n = |r|_1 / (res0_n + small) ; exit test.
// With Jacobi the next z = c r and s = r.z are formed here too.
__global__ __launch_bounds__(BS) void k_cg_upd(int n, int nvc, double *__restrict__ x, double *__restrict__ r,
                         const double *__restrict__ p, const double *__restrict__ qv,
                         const double *__restrict__ c, int jacobi, double *__restrict__ z, Red R, KS *ks) {
  if (all_done(ks, nvc)) return;
  double acc[6] = {0, 0, 0, 0, 0, 0};
  double alpha[3];
  for (int q = 0; q < nvc; ++q) alpha[q] = ks[q].alpha;
  for (int i = blockIdx.x * BS + threadIdx.x; i < n; i += gridDim.x * BS)
    for (int q = 0; q < nvc; ++q) {
      if (ks[q].done) continue;
      const int o = q * n + i;
      x[o] = x[o] + alpha[q] * p[o];
      const double ri = r[o] - alpha[q] * qv[o];
      r[o] = ri;
      acc[q] += fabs(ri);
      if (jacobi) {
        const double zi = ri * c[i];
        z[o] = zi;
        acc[3 + q] += ri * zi;
      }
    }
  if (grid_sum<6>(acc, R))
    for (int q = 0; q < nvc; ++q) {
      KS &k = ks[q];
      if (k.done) continue;
      k.res_n = acc[q] / (k.res0_n + SMALL);
      k.iter += 1;
      if (k.res_n < k.reduc || k.iter >= k.n_iter) {
        k.done = 1;
      } else if (jacobi) {
        k.beta = 1.0 / k.s;
        k.s = acc[3 + q];
        k.beta = k.beta * k.s;
      }
    }
}

// ---- BiCGStab (w_sc: res_n=1 alpha=2 beta0=3 beta=4 gam=5 omeg=6) ----
__global__ __launch_bounds__(BS) void k_bi_rr0(int n, int nvc, const double *__restrict__ r, const double *__restrict__ r0, Red R,
                         KS *ks) {
  if (all_done(ks, nvc)) return;
  double a[3] = {0, 0, 0};
  for (int i = blockIdx.x * BS + threadIdx.x; i < n; i += gridDim.x * BS)
    for (int q = 0; q < nvc; ++q)
      if (!ks[q].done) a[q] += r[q * n + i] * r0[q * n + i];
  if (grid_sum<3>(a, R))
    for (int q = 0; q < nvc; ++q) {
      KS &k = ks[q];
      if (k.done) continue;
      k.w4 = a[q];
      k.w6 = k.w4 * k.w5 / (k.w2 * k.w3 + SMALL);
      k.w3 = k.w4;
    }
}

// p = r + w6 (p - w2 u) ; Jacobi: z = c p
__global__ __launch_bounds__(BS) void k_bi_p(int n, int nvc, const double *__restrict__ r, double *__restrict__ p,
                       const double *__restrict__ u, const double *__restrict__ c, int jacobi,
                       double *__restrict__ z, const KS *ks) {
  int i = blockIdx.x * BS + threadIdx.x;
  if (i >= n) return;
  for (int q = 0; q < nvc; ++q) {
    if (ks[q].done) continue;
    const int o = q * n + i;
    const double pi = r[o] + ks[q].w6 * (p[o] - ks[q].w2 * u[o]);
    p[o] = pi;
    if (jacobi) z[o] = pi * c[i];
  }
}

// y = A z with one dot (mode 0: y.r0 -> w5) or two (mode 1: y.r, y.y -> w2)
__global__ __launch_bounds__(BS) void k_bi_spmv(DMesh M, int nvc, int mode, const double *__restrict__ diag,
                          const double *__restrict__ spv, const double *__restrict__ z, double *__restrict__ y,
                          const double *__restrict__ other, Red R, KS *ks) {
  if (all_done(ks, nvc)) return;
  const int n = M.n;
  double a[6] = {0, 0, 0, 0, 0, 0};
  for (int i = blockIdx.x * BS + threadIdx.x; i < n; i += gridDim.x * BS)
    for (int q = 0; q < nvc; ++q) {
      if (ks[q].done) continue;
      const double yi = row_spmv(M, diag, spv, z + q * n, i);
      y[q * n + i] = yi;
      a[q] += yi * other[q * n + i];
      if (mode) a[3 + q] += yi * yi;
    }
  if (grid_sum<6>(a, R))
    for (int q = 0; q < nvc; ++q) {
      KS &k = ks[q];
      if (k.done) continue;
      if (mode == 0) {
        k.w5 = a[q];
        k.w5 = k.w4 / (k.w5 + SMALL);
      } else {
        k.w1 = a[q];
        k.w2 = a[3 + q];
        k.w2 = k.w1 / (k.w2 + SMALL);
      }
    }
}

// x += w5 z ; r -= w5 u ; Jacobi: z = c r
__global__ __launch_bounds__(BS) void k_bi_upd1(int n, int nvc, double *__restrict__ x, double *__restrict__ r,
                          double *__restrict__ z, const double *__restrict__ u, const double *__restrict__ c,
                          int jacobi, const KS *ks) {
  int i = blockIdx.x * BS + threadIdx.x;
  if (i >= n) return;
  for (int q = 0; q < nvc; ++q) {
    if (ks[q].done) continue;
    const int o = q * n + i;
    const double w5 = ks[q].w5;
    x[o] = x[o] + w5 * z[o];
    const double ri = r[o] - w5 * u[o];
    r[o] = ri;
    if (jacobi) z[o] = ri * c[i];
  }
}

// x += w2 z ; r -= w2 v ; w1 = |r|_1/(res0_n+small), exit test ; next w4, w6, w3
__global__ __launch_bounds__(BS) void k_bi_upd2(int n, int nvc, double *__restrict__ x, double *__restrict__ r,
                          const double *__restrict__ z, const double *__restrict__ v,
                          const double *__restrict__ r0, Red R, KS *ks) {
  if (all_done(ks, nvc)) return;
  double a[6] = {0, 0, 0, 0, 0, 0};
  for (int i = blockIdx.x * BS + threadIdx.x; i < n; i += gridDim.x * BS)
    for (int q = 0; q < nvc; ++q) {
      if (ks[q].done) continue;
      const int o = q * n + i;
      const double w2 = ks[q].w2;
      x[o] = x[o] + w2 * z[o];
      const double ri = r[o] - w2 * v[o];
      r[o] = ri;
      a[q] += fabs(ri);
      a[3 + q] += ri * r0[o];
    }
  if (grid_sum<6>(a, R))
    for (int q = 0; q < nvc; ++q) {
      KS &k = ks[q];
      if (k.done) continue;
      k.w1 = a[q] / (k.res0_n + SMALL);
      k.res_n = k.w1;
      k.iter += 1;
      if (k.w1 < k.reduc || k.iter >= k.n_iter) {
        k.done = 1;
      } else {
        k.w4 = a[3 + q];
        k.w6 = k.w4 * k.w5 / (k.w2 * k.w3 + SMALL);
        k.w3 = k.w4;
      }
    }
}

// ------------------------------------------------------------ host side

int precond_build(Ctx &c, int imat) {
  const int n = c.n;
  if (imat == 0) {
    k_precond_jacobi<<<grid_for(n), BS, 0, c.stream>>>(n, c.diag, c.pc_c, c.pc_cr);
  } else {
    for (int l = 0; l < c.lf.count(); ++l) {
      int cnt = c.lf.ptr[l + 1] - c.lf.ptr[l];
      k_precond_level<<<grid_for(cnt), BS, 0, c.stream>>>(c.M, c.lf.rows + c.lf.ptr[l], cnt, imat, c.diag,
                                                            c.ngp, c.ngn, c.pc_c, c.pc_cr);
    }
  }
  FVK_HIP(hipGetLastError());
  c.pc_imat = imat;
  c.pc_valid = true;
  return 0;
}

int precond_apply(Ctx &c, int imat, int nvc, const double *b, double *x, const KS *gate) {
  const int n = c.n;
  if (imat == 0) {
    k_jacobi_apply<<<grid_for(n), BS, 0, c.stream>>>(n, nvc, c.pc_c, b, x, gate);
  } else {
    for (int l = 0; l < c.lf.count(); ++l) {
      int cnt = c.lf.ptr[l + 1] - c.lf.ptr[l];
      k_sweep_fwd<<<grid_for(cnt), BS, 0, c.stream>>>(c.M, c.lf.rows + c.lf.ptr[l], cnt, nvc, c.spv, c.pc_c,
                                                        b, x, gate);
    }
    for (int l = 0; l < c.lb.count(); ++l) {
      int cnt = c.lb.ptr[l + 1] - c.lb.ptr[l];
      k_sweep_bwd<<<grid_for(cnt), BS, 0, c.stream>>>(c.M, c.lb.rows + c.lb.ptr[l], cnt, nvc, c.spv, c.pc_c,
                                                        c.pc_cr, x, gate);
    }
  }
  FVK_HIP(hipGetLastError());
  return 0;
}

int spmv(Ctx &c, int nvc, const double *x, double *y) {
  k_spmv<<<grid_for(c.n), BS, 0, c.stream>>>(c.M, nvc, c.diag, c.spv, x, y);
  FVK_HIP(hipGetLastError());
  return 0;
}

static int ks_init(Ctx &c, int nvc, double reduc, int n_iter, bool bicg) {
  KS h[3];
  std::memset(h, 0, sizeof h);
  for (int q = 0; q < nvc; ++q) {
    h[q].reduc = reduc;
    h[q].n_iter = n_iter;
    if (bicg) {
      h[q].w2 = h[q].w3 = h[q].w5 = 1.0;
    } else {
      h[q].s = 1.0 / SMALL;  // large (precision.f90:9)
    }
    h[q].done = n_iter <= 0;
  }
  FVK_HIP(hipMemcpyAsync(c.ks, h, sizeof h, hipMemcpyHostToDevice, c.stream));
  return 0;
}

int residual(Ctx &c, int nvc, const double *x, double *r, bool not_r0) {
  k_residual<<<red_grid(c.n), BS, 0, c.stream>>>(c.M, nvc, c.diag, c.spv, x, r, not_r0 ? 1 : 0, c.red, c.ks);
  FVK_HIP(hipGetLastError());
  return 0;
}

// Workspace vectors (3 components each)
static inline double *wv(Ctx &c, int k) { return c.wk + (size_t)k * 3 * c.n; }

// Host check of the done flags (after a chunk of queued iterations).
static int all_done_host(Ctx &c, int nvc, bool &done) {
  KS h[3];
  FVK_HIP(hipMemcpyAsync(h, c.ks, sizeof(KS) * nvc, hipMemcpyDeviceToHost, c.stream));
  FVK_HIP(hipStreamSynchronize(c.stream));
  done = true;
  for (int q = 0; q < nvc; ++q) done = done && h[q].done;
  return 0;
}

// The full PCG: r0 already holds the residual and ks[] its norms.  With
// chunk > 0 the host stops queueing iterations once every component is done
// (checked every chunk iterations); T (optional) times each k_cg_pq launch.
static int cg_iterate(Ctx &c, int imat, int n_iter, int nvc, double *x, const double *r0, int chunk = 0,
                      KTimer *T = nullptr) {
  const int n = c.n, G = red_grid(n);
  double *r = wv(c, 0), *z = wv(c, 1), *P = wv(c, 2), *q = wv(c, 4);
  FVK_HIP(hipMemcpyAsync(r, r0, sizeof(double) * nvc * n, hipMemcpyDeviceToDevice, c.stream));
  FVK_HIP(hipMemsetAsync(P, 0, sizeof(double) * 2 * 3 * n, c.stream));
  const int jac = imat == 0;
  if (!c.pc_valid || c.pc_imat != imat) {
    int rc = precond_build(c, imat);
    if (rc) return rc;
  }
  if (!jac) {
    int rc = precond_apply(c, imat, nvc, r, z, c.ks);
    if (rc) return rc;
  }
  k_cg_rz<<<G, BS, 0, c.stream>>>(n, nvc, r, c.pc_c, jac, z, c.red, c.ks);
  for (int it = 0; it < n_iter; ++it) {
    double *pold = P + (size_t)(it & 1) * 3 * n, *pnew = P + (size_t)((it + 1) & 1) * 3 * n;
    if (T) T->begin(c.stream);
    k_cg_pq<<<G, BS, 0, c.stream>>>(c.M, nvc, c.diag, c.spv, z, pold, pnew, q, c.red, c.ks);
    if (T) T->end(c.stream);
    k_cg_upd<<<G, BS, 0, c.stream>>>(n, nvc, x, r, pnew, q, c.pc_c, jac, z, c.red, c.ks);
    if (!jac && it + 1 < n_iter) {
      int rc = precond_apply(c, imat, nvc, r, z, c.ks);
      if (rc) return rc;
      k_cg_rz<<<G, BS, 0, c.stream>>>(n, nvc, r, c.pc_c, 0, z, c.red, c.ks);
    }
    if (chunk > 0 && (it + 1) % chunk == 0 && it + 1 < n_iter) {
      bool done;
      if (T && T->flush()) return FVK_ERR_HIP;
      if (all_done_host(c, nvc, done)) return FVK_ERR_HIP;
      if (done) break;
    }
  }
  if (T && T->flush()) return FVK_ERR_HIP;
  FVK_HIP(hipGetLastError());
  return 0;
}

static int bicg_iterate(Ctx &c, int imat, int n_iter, int nvc, double *x, const double *r0) {
  const int n = c.n, G = red_grid(n), g1 = grid_for(n);
  double *r = wv(c, 0), *p = wv(c, 1), *u = wv(c, 2), *v = wv(c, 3), *z = wv(c, 4);
  FVK_HIP(hipMemcpyAsync(r, r0, sizeof(double) * nvc * n, hipMemcpyDeviceToDevice, c.stream));
  FVK_HIP(hipMemsetAsync(p, 0, sizeof(double) * 4 * 3 * n, c.stream));
  const int jac = imat == 0;
  if (!c.pc_valid || c.pc_imat != imat) {
    int rc = precond_build(c, imat);
    if (rc) return rc;
  }
  k_bi_rr0<<<G, BS, 0, c.stream>>>(n, nvc, r, r0, c.red, c.ks);
  for (int it = 0; it < n_iter; ++it) {
    k_bi_p<<<g1, BS, 0, c.stream>>>(n, nvc, r, p, u, c.pc_c, jac, z, c.ks);
    if (!jac && precond_apply(c, imat, nvc, p, z, c.ks)) return FVK_ERR_HIP;
    k_bi_spmv<<<G, BS, 0, c.stream>>>(c.M, nvc, 0, c.diag, c.spv, z, u, r0, c.red, c.ks);
    k_bi_upd1<<<g1, BS, 0, c.stream>>>(n, nvc, x, r, z, u, c.pc_c, jac, c.ks);
    if (!jac && precond_apply(c, imat, nvc, r, z, c.ks)) return FVK_ERR_HIP;
    k_bi_spmv<<<G, BS, 0, c.stream>>>(c.M, nvc, 1, c.diag, c.spv, z, v, r, c.red, c.ks);
    k_bi_upd2<<<G, BS, 0, c.stream>>>(n, nvc, x, r, z, v, r0, c.red, c.ks);
  }
  FVK_HIP(hipGetLastError());
  return 0;
}

int krylov(Ctx &c, int kind, int imat, double reduc, int n_iter, int nvc, double *x, double *r0, bool not_r0,
           int chunk, KTimer *T) {
  int rc = ks_init(c, nvc, reduc, n_iter, kind == 1);
  if (rc) return rc;
  if ((rc = residual(c, nvc, x, r0, not_r0))) return rc;
  if (kind == 0) return cg_iterate(c, imat, n_iter, nvc, x, r0, chunk, T);
  return bicg_iterate(c, imat, n_iter, nvc, x, r0);
}

// FP64 true residual of the current matrix: out = {|b-Ax|_1, |b-Ax|_2^2,
// |b|_1, |b|_2^2}
__global__ __launch_bounds__(BS) void k_true_res(DMesh M, const double *__restrict__ diag,
                                                 const double *__restrict__ spv, const double *__restrict__ x,
                                                 const double *__restrict__ b, Red R, double *out) {
  double a[4] = {0, 0, 0, 0};
  for (int i = blockIdx.x * BS + threadIdx.x; i < M.n; i += gridDim.x * BS) {
    const double bi = b[i], ri = bi - row_spmv(M, diag, spv, x, i);
    a[0] += fabs(ri);
    a[1] += ri * ri;
    a[2] += fabs(bi);
    a[3] += bi * bi;
  }
  if (grid_sum<4>(a, R))
    for (int j = 0; j < 4; ++j) out[j] = a[j];
}

int true_residual(Ctx &c, const double *x, const double *b, double *dev_out4, double *host_out4) {
  k_true_res<<<red_grid(c.n), BS, 0, c.stream>>>(c.M, c.diag, c.spv, x, b, c.red, dev_out4);
  FVK_HIP(hipGetLastError());
  FVK_HIP(hipMemcpyAsync(host_out4, dev_out4, 4 * sizeof(double), hipMemcpyDeviceToHost, c.stream));
  FVK_HIP(hipStreamSynchronize(c.stream));
  return 0;
}

}  // namespace fvk
