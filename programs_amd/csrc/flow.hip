// SIMPLE layer on gfx950: Green-Gauss gradients, momentum and
// pressure-correction assembly, and the outer iteration.
//
//   solv_grad      spray-simulation/cfd_setup.f90:765-866
//   cfd_setup_eqn  spray-simulation/cfd_setup.f90:22-544 (iph=1, n_tl=1,
//                  single phase: set_bnd/calc_rhs/calc_srhs are no-ops)
//   cfd_setup_pp   spray-simulation/cfd_setup.f90:549-760 (n_pp=1)
//   outer loop     spray-simulation/cfd_solve.f90:134-178
//
// The reference scatters face contributions into its two cells in face
// order.  Here a face kernel computes each face's quantities once, and a
// cell kernel gathers them over the cell's faces in ascending face index,
// which is the order the scatter reaches that cell: the accumulations are
// the same floating-point operations in the same order, without atomics.
#include <cstring>

#include "flow.h"

namespace fvk {


// ---------------------------------------------------------------- gradient
// One Green-Gauss sweep.  phi: nvc x n (cells), phib: nvc x nbf (boundary
// slots), g0: previous sweep (nvc*3 x n), g: result.  it is 1-based.
template <int nvc, int nd>
__global__ __launch_bounds__(BS) void k_grad(DMesh M, const double *__restrict__ phi, const double *__restrict__ phib,
                                             const double *__restrict__ g0, double *__restrict__ g, int it, int neu,
                                             const double *__restrict__ on_flag) {
  const int i = xcd_block() * BS + threadIdx.x, n = M.n;
  if (i >= n) return;
  double acc[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
  if (on_flag && !(*on_flag > SMALL)) {
    for (int q = 0; q < nvc; ++q)
      for (int d = 0; d < 3; ++d) g[(q * 3 + d) * n + i] = 0.0;
    return;
  }
  const double vri = M.vr[i];
  // the faces' geometry comes from the per-slot tables (the same values as
  // the face arrays, read by adjacent rows at adjacent addresses)
  for (int k = 0; k < M.W; ++k) {
    const size_t o = (size_t)k * n + i;
    const int e = M.cf[o];
    if (e == PAD) break;
    const bool own = e >= 0;
    const int other = M.go[o];
    // an interior face's geometry from its owner's entry (the same values)
    const int ko = M.gk[o];
    const size_t og = ko < 0 ? o : (size_t)ko * n + other;
    const double s = M.gs[og];
    const double nv[3] = {M.gnx[og], M.gny[og], M.gnz[og]};
    double pf[3];
    if (other >= 0) {
      const int ip = own ? i : other, in = own ? other : i;
      const double w = M.gw[og];
      if (it > 1) {
        const double tp[3] = {M.gpx[og], M.gpy[og], M.gpz[og]}, tn[3] = {M.gqx[og], M.gqy[og], M.gqz[og]};
        for (int q = 0; q < nvc; ++q) {
          double gp[3] = {g0[(q * 3 + 0) * n + ip], g0[(q * 3 + 1) * n + ip], g0[(q * 3 + 2) * n + ip]};
          double gn[3] = {g0[(q * 3 + 0) * n + in], g0[(q * 3 + 1) * n + in], g0[(q * 3 + 2) * n + in]};
          const double pp = phi[q * n + ip] + dotn(gp, tp, nd);
          const double pn = phi[q * n + in] + dotn(gn, tn, nd);
          pf[q] = w * pn + (1.0 - w) * pp;
        }
      } else {
        for (int q = 0; q < nvc; ++q) pf[q] = w * phi[q * n + in] + (1.0 - w) * phi[q * n + ip];
      }
      for (int q = 0; q < nvc; ++q)
        for (int d = 0; d < nd; ++d) {
          const double t = pf[q] * s * nv[d] * vri;
          acc[q][d] = own ? acc[q][d] + t : acc[q][d] - t;
        }
    } else {
      if (neu > 0) {
        for (int q = 0; q < nvc; ++q) pf[q] = phi[q * n + i];
        if (it > 1 && neu == 2) {
          const double t[3] = {M.gpx[o], M.gpy[o], nd > 2 ? M.gpz[o] : 0.0};
          for (int q = 0; q < nvc; ++q) {
            double gp[3] = {g0[(q * 3 + 0) * n + i], g0[(q * 3 + 1) * n + i], g0[(q * 3 + 2) * n + i]};
            pf[q] = pf[q] + dotn(gp, t, nd);
          }
        }
      } else {
        const int b = ~other;
        for (int q = 0; q < nvc; ++q) pf[q] = phib[q * M.nbf + b];
      }
      for (int q = 0; q < nvc; ++q)
        for (int d = 0; d < nd; ++d) acc[q][d] = acc[q][d] + pf[q] * s * nv[d] * vri;
    }
  }
  for (int q = 0; q < nvc; ++q)
    for (int d = 0; d < 3; ++d) g[(q * 3 + d) * n + i] = acc[q][d];
}

// solv_grad with n_iter sweeps; the last sweep lands in g.
int grad(Ctx &c, int nvc, const double *phi, const double *phib, double *g, int n_iter, int neu,
         const double *on_flag, double *g0) {
  // sweeps alternate between g0 and g so that the last one lands in g
  for (int it = 1; it <= n_iter; ++it) {
    double *out = ((n_iter - it) % 2 == 0) ? g : g0;
    const double *prev = ((n_iter - it) % 2 == 0) ? g0 : g;
    const dim3 G(grid_for(c.n)), B(BS);
    if (c.ndx == 3) {
      if (nvc == 1) k_grad<1, 3><<<G, B, 0, c.stream>>>(c.M, phi, phib, prev, out, it, neu, on_flag);
      else if (nvc == 2) k_grad<2, 3><<<G, B, 0, c.stream>>>(c.M, phi, phib, prev, out, it, neu, on_flag);
      else k_grad<3, 3><<<G, B, 0, c.stream>>>(c.M, phi, phib, prev, out, it, neu, on_flag);
    } else {
      if (nvc == 1) k_grad<1, 2><<<G, B, 0, c.stream>>>(c.M, phi, phib, prev, out, it, neu, on_flag);
      else if (nvc == 2) k_grad<2, 2><<<G, B, 0, c.stream>>>(c.M, phi, phib, prev, out, it, neu, on_flag);
      else k_grad<3, 2><<<G, B, 0, c.stream>>>(c.M, phi, phib, prev, out, it, neu, on_flag);
    }
  }
  FVK_HIP(hipGetLastError());
  return 0;
}

// ---------------------------------------------------------------- reductions
__global__ __launch_bounds__(BS) void k_absmax(int n, int nvc, const double *__restrict__ a, int nb, const double *__restrict__ b,
                         double *out_cell, double *out_all) {
  double m[2] = {0.0, 0.0};
  for (int i = blockIdx.x * BS + threadIdx.x; i < n; i += gridDim.x * BS)
    for (int q = 0; q < nvc; ++q) m[0] = fmax(m[0], fabs(a[q * n + i]));
  for (int i = blockIdx.x * BS + threadIdx.x; i < nb; i += gridDim.x * BS)
    for (int q = 0; q < nvc; ++q) m[1] = fmax(m[1], fabs(b[q * nb + i]));
  block_max<2>(m);
  if (threadIdx.x == 0) {
    // max is exact and order-free: combine with atomics on the bit pattern
    // of non-negative doubles (monotone as unsigned integers)
    atomicMax((unsigned long long *)out_cell, (unsigned long long)__double_as_longlong(m[0]));
    atomicMax((unsigned long long *)out_all, (unsigned long long)__double_as_longlong(fmax(m[0], m[1])));
  }
}

__global__ __launch_bounds__(RB) void k_sum(int n, const double *__restrict__ a, double *part) {
  double s[1] = {0.0};
  int r0, r1;
  block_rows(n, r0, r1);
  for (int i = r0 + threadIdx.x; i < r1; i += RB) s[0] += a[i];
  store_partials<1>(s, part);
}

__global__ __launch_bounds__(RB) void k_sum_total(const double *part, int P, double *out) {
  double s[1];
  load_totals<1>(part, P, s);
  if (threadIdx.x == 0) *out = s[0];
}

static int sum_to(Ctx &c, const double *a, double *out) {
  const int G = red_grid(c.n);
  double *part = c.part + (size_t)3 * RG * 8;
  k_sum<<<G, RB, 0, c.stream>>>(c.n, a, part);
  k_sum_total<<<1, RB, 0, c.stream>>>(part, G, out);
  FVK_HIP(hipGetLastError());
  return 0;
}

// ---------------------------------------------------------------- momentum
struct MomPar {
  double rho, mu, c_bl, urf;
  int c_sch;
};

__device__ __forceinline__ void load3(const double *a, int n, int i, int nvc, double *o) {
  for (int q = 0; q < nvc; ++q) o[q] = a[q * n + i];
}

// Velocity of cell i as seen by face g: the outlet back-flow projection of
// an earlier (lower index) outlet face of the cell has already replaced it.
__device__ __forceinline__ void cell_u_for_face(const double *u, int n, int i, int g, int nvc, const int *cell_oc,
                                                const int *oc_face, const double *oc_u, double *o) {
  load3(u, n, i, nvc, o);
  if (!cell_oc) return;
  const int m = cell_oc[i];
  if (m < 0) return;
  int k = 0;
  while (k < 4 && oc_face[m * 4 + k] >= 0 && oc_face[m * 4 + k] < g) ++k;
  if (k > 0)
    for (int q = 0; q < nvc; ++q) o[q] = oc_u[(m * 4 + k - 1) * 3 + q];
}

// Outlet back-flow projection, applied per cell in face order
// (cfd_setup.f90:349-356): versions of the cell velocity after each outlet face.
__global__ __launch_bounds__(BS) void k_outlet_pre(DMesh M, int n_oc, int nvc, const int *oc_cell, const int *oc_face,
                             const double *u, double *oc_u) {
  const int m = blockIdx.x * BS + threadIdx.x;
  if (m >= n_oc) return;
  const int i = oc_cell[m], n = M.n, nd = M.ndx;
  double pp[3] = {0, 0, 0};
  load3(u, n, i, nvc, pp);
  for (int k = 0; k < 4; ++k) {
    const int f = oc_face[m * 4 + k];
    if (f < 0) break;
    double nv[3];
    face_n(M, f, nv);
    if (dotn(pp, nv, nd) < -SMALL) {
      double t[3];
      for (int q = 0; q < nvc; ++q) t[q] = pp[q];
      double dd = dotn(t, nv, nd);
      for (int q = 0; q < nvc; ++q) t[q] = t[q] - nv[q] * dd;
      dd = magn(t, nvc) + SMALL;
      for (int q = 0; q < nvc; ++q) t[q] = t[q] / dd;
      dd = dotn(pp, t, nvc);
      for (int q = 0; q < nvc; ++q) pp[q] = t[q] * dd;
    }
    for (int q = 0; q < 3; ++q) oc_u[(m * 4 + k) * 3 + q] = pp[q];
  }
}

// Face part of cfd_setup_eqn's face loop (cfd_setup.f90:178-431).
template <int nd>
__global__ __launch_bounds__(BS) void k_mom_face(DMesh M, MomPar P, const double *__restrict__ u, double *__restrict__ ub,
                           const double *__restrict__ ddx, double *__restrict__ spd, double *__restrict__ frc_dns,
                           double *__restrict__ ngp, double *__restrict__ ngn, double *__restrict__ cor,
                           double *__restrict__ cfl, double *__restrict__ cinout, const int *cell_oc,
                           const int *oc_face, const double *oc_u) {
  constexpr int nvc = nd;
  const int f = xcd_block() * BS + threadIdx.x;
  if (f >= M.nf) return;
  const int n = M.n, nf = M.nf;
  const int ip = M.fip[f], in = M.fin[f];
  const double s = M.fs[f];
  double nv[3], tp[3];
  face_n(M, f, nv);
  face_dxp(M, f, tp);
  double up[3] = {0, 0, 0}, pa[3] = {0, 0, 0}, cr[3] = {0, 0, 0};
  cell_u_for_face(u, n, ip, f, nvc, cell_oc, oc_face, oc_u, up);
  for (int q = 0; q < nvc; ++q) {
    double g[3] = {ddx[(q * 3 + 0) * n + ip], ddx[(q * 3 + 1) * n + ip], ddx[(q * 3 + 2) * n + ip]};
    pa[q] = up[q] + dotn(g, tp, nd);
  }
  double cc = 0.0, d = 0.0;
  if (in >= 0) {
    const double w = M.fwfp[f];
    double tn[3], un[3] = {0, 0, 0}, na[3] = {0, 0, 0}, pfce[3] = {0, 0, 0};
    face_dxn(M, f, tn);
    cell_u_for_face(u, n, in, f, nvc, cell_oc, oc_face, oc_u, un);
    for (int q = 0; q < nvc; ++q) {
      double g[3] = {ddx[(q * 3 + 0) * n + in], ddx[(q * 3 + 1) * n + in], ddx[(q * 3 + 2) * n + in]};
      na[q] = un[q] + dotn(g, tn, nd);
    }
    for (int q = 0; q < nvc; ++q) pfce[q] = w * na[q] + (1.0 - w) * pa[q];
    const double sp = dotn(pfce, nv, nd);
    spd[f] = sp;
    const double frc = w * 1.0 + (1.0 - w) * 1.0;
    const double dns = w * P.rho + (1.0 - w) * P.rho;
    const double vis = w * P.mu + (1.0 - w) * P.mu;
    const double fd = frc * dns, fv = frc * vis;
    frc_dns[f] = fd;
    if (fd > SMALL && fabs(sp) > SMALL) {
      cc = fd * s * sp;
      double cimp[3], cexp[3];
      for (int q = 0; q < nvc; ++q) cimp[q] = cc > 0.0 ? up[q] : un[q];
      if (P.c_sch == 0) {
        for (int q = 0; q < nvc; ++q) cexp[q] = pfce[q];
      } else {
        int ic;
        const double *phc, *phd;
        double dxpn[3];
        const double xp[3] = {M.cx[ip], M.cy[ip], nd > 2 ? M.cz[ip] : 0.0};
        const double xn[3] = {M.cx[in], M.cy[in], nd > 2 ? M.cz[in] : 0.0};
        if (cc > 0.0) {
          ic = ip; phc = pa; phd = na;
          for (int e = 0; e < nd; ++e) dxpn[e] = (xn[e] + tn[e]) - (xp[e] + tp[e]);
        } else {
          ic = in; phc = na; phd = pa;
          for (int e = 0; e < nd; ++e) dxpn[e] = (xp[e] + tp[e]) - (xn[e] + tn[e]);
        }
        double phu[3];
        for (int q = 0; q < nvc; ++q) {
          double g[3] = {ddx[(q * 3 + 0) * n + ic], ddx[(q * 3 + 1) * n + ic], ddx[(q * 3 + 2) * n + ic]};
          phu[q] = phd[q] - 2.0 * dotn(g, dxpn, nd);
        }
        if (P.c_sch <= 3) {
          for (int q = 0; q < nvc; ++q) {
            double r = (phc[q] - phu[q]) / (phd[q] - phc[q] + SMALL);
            if (P.c_sch == 1) {
              r = fmax(0.0, fmin(1.0, r));
            } else if (P.c_sch == 2) {
              r = (r + fabs(r)) / (1.0 + fabs(r) + SMALL);
            } else if (P.c_sch == 3) {
              r = fmax(fmax(0.0, fmin(1.0, 2.0 * r)), fmin(2.0, r));
            }
            cexp[q] = phc[q] + 0.5 * r * (phd[q] - phc[q]);
          }
        } else {
          for (int q = 0; q < nvc; ++q) {  // HRIC with co_num = 0 (steady)
            const double pcn = (phc[q] - phu[q]) / (phd[q] - phu[q] + SMALL);
            double pfn = pcn;
            if (pcn > 0.0 && pcn < 0.5) pfn = 2.0 * pcn;
            else if (pcn > 0.5 && pcn < 1.0) pfn = 1.0;
            double g[3] = {ddx[(q * 3 + 0) * n + ic], ddx[(q * 3 + 1) * n + ic], ddx[(q * 3 + 2) * n + ic]};
            double th = dotn(g, nv, nd) / (magn(g, nd) + SMALL);
            th = sqrt(fabs(th));
            pfn = pfn * th + pcn * (1.0 - th);
            double gm = (phd[q] - phu[q]) / (phd[q] - phc[q] + SMALL);
            gm = gm * (1.0 - pfn);
            gm = fmax(fmin(gm, 1.0), 0.0);
            cexp[q] = gm * phc[q] + (1.0 - gm) * phd[q];
          }
        }
      }
      for (int q = 0; q < nvc; ++q) cr[q] = P.c_bl * cc * (cimp[q] - cexp[q]);
    }
    if (fv > SMALL) {
      d = fv * s / M.flpn[f];
      for (int q = 0; q < nvc; ++q) {
        const double dimp = un[q] - up[q];
        const double dexp = na[q] - pa[q];
        cr[q] = cr[q] + d * (dexp - dimp);
      }
    }
    ngp[f] = -d + fmin(cc, 0.0);
    ngn[f] = -d - fmax(cc, 0.0);
  } else {
    const int b = M.fib[f];
    const int bty = M.fbty[f] - M.fbty[f] % 10;
    double pb[3] = {0, 0, 0};
    for (int q = 0; q < nvc; ++q) pb[q] = ub[q * M.nbf + b];
    double sp = dotn(pb, nv, nd);
    if (bty == 20 && dotn(up, nv, nd) < -SMALL) {
      // the cell value after this face's projection (k_outlet_pre)
      const int m = cell_oc[ip];
      int k = 0;
      while (oc_face[m * 4 + k] != f) ++k;
      for (int q = 0; q < nvc; ++q) up[q] = oc_u[(m * 4 + k) * 3 + q];
      sp = dotn(up, nv, nd);
    }
    spd[f] = sp;
    const double frc = 1.0, dns = P.rho, vis = P.mu;
    const double fd = frc * dns, fv = frc * vis;
    frc_dns[f] = fd;
    double cin = 0.0;
    if (fd > SMALL && fabs(sp) > SMALL) {
      cc = fd * s * sp;
      if (bty == 10 || bty == 20) cin = frc * s * sp;
    }
    double dimp[3] = {0, 0, 0}, dexp[3] = {0, 0, 0};
    if (fv > SMALL) {
      d = fv * s / M.flpn[f];
      for (int q = 0; q < nvc; ++q) {
        dimp[q] = pb[q] - up[q];
        dexp[q] = pb[q] - pa[q];
      }
    }
    double gp = -d + fmin(cc, 0.0);
    if (bty == 40) {
      double dd = dotn(up, nv, nd);
      for (int q = 0; q < nvc; ++q) pb[q] = up[q] - nv[q] * dd;
      if (d > SMALL) {
        dd = dotn(up, nv, nd);
        for (int q = 0; q < nvc; ++q) dexp[q] = -(2.0 * nv[q] * dd);
      }
    }
    if (bty == 30 && d > SMALL) {
      double t[3], dd;
      for (int q = 0; q < nvc; ++q) t[q] = up[q] + SMALL;
      if (magn(pb, nvc) > SMALL)
        for (int q = 0; q < nvc; ++q) t[q] = pb[q];
      dd = dotn(t, nv, nd);
      for (int q = 0; q < nvc; ++q) t[q] = t[q] - nv[q] * dd;
      dd = magn(t, nvc) + SMALL;
      for (int q = 0; q < nvc; ++q) t[q] = t[q] / dd;
      dd = dotn(pb, t, nvc);
      for (int q = 0; q < nvc; ++q) pb[q] = t[q] * dd;
      dd = dotn(up, t, nvc);
      for (int q = 0; q < nvc; ++q) dexp[q] = pb[q] - t[q] * dd;
    }
    if (bty == 20) {
      for (int q = 0; q < nvc; ++q) pb[q] = up[q];
      gp = 0.0;
    }
    if (d > SMALL)
      for (int q = 0; q < nvc; ++q) cr[q] = d * (dexp[q] - dimp[q]);
    for (int q = 0; q < nvc; ++q) ub[q * M.nbf + b] = pb[q];
    ngp[f] = gp;
    cinout[f] = cin;
  }
  for (int q = 0; q < nvc; ++q) cor[q * nf + f] = cr[q];
  cfl[f] = cc;
}

// c_in / c_out: sequential sums in face order (cfd_setup.f90:372-379).
__global__ __launch_bounds__(BS) void k_cinout(const int *faces, int n_in, int n_out, const double *cinout, FlowSc *sc) {
  if (blockIdx.x != 0 || threadIdx.x != 0) return;
  double ci = sc->c_in, co = sc->c_out;
  for (int k = 0; k < n_in; ++k) ci = ci - cinout[faces[k]];
  for (int k = 0; k < n_out; ++k) co = co + cinout[faces[n_in + k]];
  sc->c_in = ci;
  sc->c_out = co;
}

// Cell part: gather the face contributions in ascending face order, then the
// volume loop (cfd_setup.f90:437-519).
__global__ __launch_bounds__(BS) void k_mom_cell(DMesh M, MomPar P, int nvc, double *__restrict__ u, const double *__restrict__ ub,
                           const double *__restrict__ ngp, const double *__restrict__ ngn,
                           const double *__restrict__ cor, const double *__restrict__ cfl,
                           const double *__restrict__ dpdx, double *__restrict__ diag, double *__restrict__ spv,
                           double *__restrict__ rhs, double *__restrict__ cs, double *__restrict__ frc_vol) {
  const int i = xcd_block() * BS + threadIdx.x;
  if (i >= M.n) return;
  const int n = M.n, nf = M.nf;
  double pl = 0.0, c_s = 0.0, r[3] = {0.0, 0.0, 0.0};
  for (int q = 0; q < nvc; ++q) r[q] = r[q] + 0.0;  // eqn%rhs, zero in single-phase flow
  int ks = 0;
  for (int k = 0; k < M.W; ++k) {
    const size_t o = (size_t)k * n + i;
    const int e = M.cf[o];
    if (e == PAD) break;
    const bool own = e >= 0;
    const int f = own ? e : ~e;
    const int other = M.go[o];  // the other row, or ~fib on a boundary face
    if (other >= 0) {
      if (own) {
        const double a = ngp[f];
        pl = pl - a;
        for (int q = 0; q < nvc; ++q) r[q] = r[q] + cor[q * nf + f];
        c_s = c_s + cfl[f];
        spv[(size_t)ks * M.ld + i] = a;
      } else {
        const double a = ngn[f];
        pl = pl - a;
        for (int q = 0; q < nvc; ++q) r[q] = r[q] - cor[q * nf + f];
        c_s = c_s - cfl[f];
        spv[(size_t)ks * M.ld + i] = a;
      }
      ++ks;
    } else {
      const double a = ngp[f];
      const int b = ~other;
      pl = pl - a;
      for (int q = 0; q < nvc; ++q) r[q] = r[q] + cor[q * nf + f];
      for (int q = 0; q < nvc; ++q) r[q] = r[q] - a * ub[q * M.nbf + b];
      c_s = c_s + cfl[f];
    }
  }
  double ui[3] = {0, 0, 0};
  load3(u, n, i, nvc, ui);
  const double fvol = 1.0 * M.v[i];
  for (int q = 0; q < nvc; ++q) r[q] = r[q] - dpdx[q * n + i] * fvol;
  pl = (1.0 / P.urf) * pl;
  for (int q = 0; q < nvc; ++q) r[q] = r[q] + (1.0 - P.urf) * pl * ui[q];
  diag[i] = pl;
  for (int q = 0; q < nvc; ++q) rhs[q * n + i] = r[q];
  cs[i] = c_s;
  frc_vol[i] = fvol;
}

// Final outlet-projected cell velocity replaces the cell value.
__global__ __launch_bounds__(BS) void k_outlet_post(int n_oc, int n, int nvc, const int *oc_cell, const int *oc_face, const double *oc_u,
                              double *u) {
  const int m = blockIdx.x * BS + threadIdx.x;
  if (m >= n_oc) return;
  int k = 0;
  while (k < 4 && oc_face[m * 4 + k] >= 0) ++k;
  if (k == 0) return;
  for (int q = 0; q < nvc; ++q) u[q * n + oc_cell[m]] = oc_u[(m * 4 + k - 1) * 3 + q];
}

// Residual normalisation of the momentum equation (cfd_setup.f90:530-543).
__global__ __launch_bounds__(BS) void k_mom_norm(int it, int nvc, const KS *ks, FlowSc *sc) {
  if (blockIdx.x != 0 || threadIdx.x != 0) return;
  double mx0 = ks[0].res0_n;
  for (int q = 0; q < nvc; ++q) {
    sc->res0_n_u[q] = ks[q].res0_n;
    sc->it_u[q] = ks[q].iter;
    if (q) mx0 = fmax(mx0, ks[q].res0_n);
  }
  if (it == 1) sc->res_n_u = 0.0;
  if (mx0 > SMALL) {
    sc->res_n_u = fmax(sc->res_n_u, mx0);
    mx0 = mx0 / sc->res_n_u;
    sc->mx_res_n = fmax(sc->mx_res_n, mx0);
  }
  sc->h_reduc = mx0;
  if (fabs(mx0) > 1.0 / SMALL) sc->diverged = 1;
}

// ------------------------------------------------------- pressure correction
__global__ __launch_bounds__(BS) void k_plr(int n, const double *diag, double *pl_r) {
  const int i = blockIdx.x * BS + threadIdx.x;
  if (i < n) pl_r[i] = 1.0 / (diag[i] + SMALL);
}

// c_rat (cfd_setup.f90:590-594)
__global__ __launch_bounds__(BS) void k_crat(FlowSc *sc) {
  if (blockIdx.x != 0 || threadIdx.x != 0) return;
  const double delta = 1.0 / 1e10;
  double c_rat = sc->c_out / (sc->c_in + delta);
  if (c_rat < delta) c_rat = 1.0;
  c_rat = 1.0 / c_rat;
  if (fabs(sc->c_in) < delta) c_rat = 0.0;
  sc->c_rat = c_rat;
}

template <int nd>
__global__ __launch_bounds__(BS) void k_pp_face(DMesh M, const double *__restrict__ p, const double *__restrict__ dpdx,
                          const double *__restrict__ pl_r, const double *__restrict__ frc_dns,
                          const double *__restrict__ spd, double *__restrict__ ngp, double *__restrict__ ngn,
                          double *__restrict__ cpp, const FlowSc *sc) {
  const int f = xcd_block() * BS + threadIdx.x;
  if (f >= M.nf) return;
  const int n = M.n, ip = M.fip[f], in = M.fin[f];
  if (in >= 0) {
    const double w = M.fwfp[f];
    double tp[3], tn[3];
    face_dxp(M, f, tp);
    face_dxn(M, f, tn);
    const double gp[3] = {dpdx[0 * n + ip], dpdx[1 * n + ip], nd > 2 ? dpdx[2 * n + ip] : 0.0};
    const double gn[3] = {dpdx[0 * n + in], dpdx[1 * n + in], nd > 2 ? dpdx[2 * n + in] : 0.0};
    const double pa = p[ip] + dotn(gp, tp, nd);
    const double na = p[in] + dotn(gn, tn, nd);
    double dp = na - pa;
    double gf[3], dxc[3];
    for (int e = 0; e < nd; ++e) gf[e] = w * gn[e] + (1.0 - w) * gp[e];
    dxc[0] = M.cx[in] - M.cx[ip];
    dxc[1] = M.cy[in] - M.cy[ip];
    dxc[2] = nd > 2 ? M.cz[in] - M.cz[ip] : 0.0;
    dp = dp - dotn(gf, dxc, nd);
    const double p_r = w * pl_r[in] + (1.0 - w) * pl_r[ip];
    const double v = w * M.v[in] + (1.0 - w) * M.v[ip];
    const double d = frc_dns[f] * p_r * M.fs[f] * v / M.flpn[f];
    const double c = -(d * dp);
    ngp[f] = -d;
    ngn[f] = -d;
    cpp[f] = c;
  } else {
    const int bty = M.fbty[f] - M.fbty[f] % 10;
    double c = 0.0;
    if (bty == 20) {
      c = frc_dns[f] * M.fs[f] * spd[f];
      c = c * (sc->c_rat - 1.0);
    }
    cpp[f] = c;
  }
}

__global__ __launch_bounds__(BS) void k_pp_cell(DMesh M, const double *__restrict__ ngp, const double *__restrict__ ngn,
                          const double *__restrict__ cpp, const double *__restrict__ cs, double *__restrict__ diag,
                          double *__restrict__ spv, double *__restrict__ rhs) {
  const int i = xcd_block() * BS + threadIdx.x;
  if (i >= M.n) return;
  const int n = M.n;
  double pl = 0.0, r = -cs[i];
  int ks = 0;
  for (int k = 0; k < M.W; ++k) {
    const size_t o = (size_t)k * n + i;
    const int e = M.cf[o];
    if (e == PAD) break;
    const bool own = e >= 0;
    const int f = own ? e : ~e;
    if (M.go[o] >= 0) {
      const double a = own ? ngp[f] : ngn[f];
      pl = pl - a;
      r = own ? r - cpp[f] : r + cpp[f];
      spv[(size_t)ks * M.ld + i] = a;
      ++ks;
    } else if (M.fbty[f] - M.fbty[f] % 10 == 20) {
      r = r - cpp[f];
    }
  }
  diag[i] = pl;
  rhs[i] = r;
}

__global__ __launch_bounds__(BS) void k_pp_norm(int it, int which, const KS *ks, const double *sum_rhs, FlowSc *sc) {
  if (blockIdx.x != 0 || threadIdx.x != 0) return;
  double mx0 = ks[0].res0_n + SMALL;
  if (it == 1) sc->pp_res_n = 0.0;
  sc->pp_res_n = fmax(sc->pp_res_n, mx0);
  mx0 = mx0 / sc->pp_res_n;
  sc->mx_res_n = fmax(sc->mx_res_n, mx0);
  if (which == 0) {
    sc->h_redpp = mx0;
    sc->h_redpp_sum = *sum_rhs;
    sc->res0_n_pp = ks[0].res0_n;
    sc->it_pp = ks[0].iter;
  } else {
    sc->h_red_p = mx0;
    sc->h_red_p_sum = *sum_rhs;
    sc->res0_n_p = ks[0].res0_n;
    sc->it_p = ks[0].iter;
  }
}

// relaxed pressure update p += urf (p' - p'_1) (cfd_setup.f90:675-683)
__global__ __launch_bounds__(BS) void k_p_relax(int n, double urf, const double *pc, int row1, double *p) {
  const int i = blockIdx.x * BS + threadIdx.x;
  if (i < n) p[i] = p[i] + urf * (pc[i] - pc[row1]);
}

// u -= grad(p') * pl_r * frc_vol (cfd_setup.f90:695-700)
__global__ __launch_bounds__(BS) void k_u_corr(int n, int nvc, const double *ddxp, const double *pl_r, const double *frc_vol,
                         double *u) {
  const int i = blockIdx.x * BS + threadIdx.x;
  if (i >= n) return;
  for (int q = 0; q < nvc; ++q) u[q * n + i] = u[q * n + i] - ddxp[q * n + i] * pl_r[i] * frc_vol[i];
}

// flux correction (cfd_setup.f90:708-728), face part
template <int nd>
__global__ __launch_bounds__(BS) void k_flux_face(DMesh M, const double *__restrict__ pc, const double *__restrict__ ddxp,
                            const double *__restrict__ ngp, const double *__restrict__ frc_dns,
                            double *__restrict__ spd, double *__restrict__ fc1) {
  const int f = xcd_block() * BS + threadIdx.x;
  if (f >= M.nf) return;
  const int in = M.fin[f];
  if (in < 0) return;
  const int n = M.n, ip = M.fip[f];
  double tp[3], tn[3];
  face_dxp(M, f, tp);
  face_dxn(M, f, tn);
  const double gp[3] = {ddxp[0 * n + ip], ddxp[1 * n + ip], nd > 2 ? ddxp[2 * n + ip] : 0.0};
  const double gn[3] = {ddxp[0 * n + in], ddxp[1 * n + in], nd > 2 ? ddxp[2 * n + in] : 0.0};
  const double pax = dotn(gp, tp, nd);
  const double pa = pc[ip] + pax;
  const double nax = dotn(gn, tn, nd);
  const double na = pc[in] + nax;
  fc1[f] = ngp[f] * (nax - pax);
  spd[f] = spd[f] + ngp[f] * (na - pa) / (frc_dns[f] * M.fs[f]);
}

// cell part: rhs_p = rhs + (accumulated correction), SIMPLER only
__global__ __launch_bounds__(BS) void k_flux_cell(DMesh M, const double *__restrict__ fc1, const double *__restrict__ rhs,
                            double *__restrict__ rhs_p) {
  const int i = xcd_block() * BS + threadIdx.x;
  if (i >= M.n) return;
  const int n = M.n;
  double r = 0.0;
  for (int k = 0; k < M.W; ++k) {
    const size_t o = (size_t)k * n + i;
    const int e = M.cf[o];
    if (e == PAD) break;
    const bool own = e >= 0;
    const int f = own ? e : ~e;
    if (M.go[o] < 0) continue;
    r = own ? r - fc1[f] : r + fc1[f];
  }
  rhs_p[i] = rhs[i] + r;
}

__global__ __launch_bounds__(BS) void k_outer_begin(FlowSc *sc) {
  if (blockIdx.x != 0 || threadIdx.x != 0) return;
  sc->mx_res_n = 0.0;
  sc->c_in = 0.0;
  sc->c_out = 0.0;
  sc->h_reduc = sc->h_redpp = sc->h_red_p = 0.0;
  sc->h_redpp_sum = sc->h_red_p_sum = 0.0;
  sc->phi_mx_cell = sc->phi_mx_all = 0.0;
}

// ------------------------------------------------------------- host driver


// The flow's own stores of the Krylov scalars a pressure solve to tolerance
// leaves for k_pp_norm: res0_n = |b|_1 (x0 = 0) and the iteration count.
__global__ void k_ks_store(KS *ks, double res0_n, int iters, double res_n) {
  if (threadIdx.x != 0) return;
  ks[0].res0_n = res0_n;
  ks[0].iter = iters;
  ks[0].res_n = res_n;
  ks[0].done = 1;
}

__global__ void k_hist_push(FlowSc *sc, FlowSc *hist, int cap) {
  if (threadIdx.x != 0) return;
  const int k = sc->hist_k;
  if (k < cap) hist[k] = *sc;
  sc->hist_k = k + 1;
}

__global__ void k_hist_reset(FlowSc *sc) {
  if (threadIdx.x == 0) sc->hist_k = 0;
}

int mom_setup(Flow &F, int it, bool &skipped) {
  Ctx &c = *F.c;
  const int n = c.n, nf = c.nf, nvc = F.nvc;
  MomPar P{F.cfg.rho, F.cfg.mu, F.cfg.c_bl, F.cfg.urf_u, F.cfg.c_sch};
  skipped = false;
  if (fill(c, &F.sc->phi_mx_cell, 2, 0.0)) return FVK_ERR_HIP;
  k_absmax<<<grid_for(n > c.nbf ? n : c.nbf) < 2048 ? grid_for(n > c.nbf ? n : c.nbf) : 2048, BS, 0, c.stream>>>(n, nvc, F.u, c.nbf, F.ub, &F.sc->phi_mx_cell,
                                                                 &F.sc->phi_mx_all);
  // the early return when every velocity is zero (cfd_setup.f90:170-171);
  // a captured iteration follows one that did not take it, and with fixed
  // non-zero boundary values it is never taken again
  if (!F.capture) {
    double mx[2];
    FVK_HIP(hipMemcpyAsync(mx, &F.sc->phi_mx_cell, 2 * sizeof(double), hipMemcpyDeviceToHost, c.stream));
    FVK_HIP(hipStreamSynchronize(c.stream));
    if (mx[1] < SMALL) {
      if (fill(c, c.diag, n, 0.0) || fill(c, F.cs, n, 0.0)) return FVK_ERR_HIP;
      skipped = true;
      return 0;
    }
  }
  int rc = grad(c, nvc, F.u, F.ub, F.ddx, F.cfg.n_grad, 0, &F.sc->phi_mx_cell, F.g0);
  if (rc) return rc;
  if (F.n_oc)
    k_outlet_pre<<<grid_for(F.n_oc), BS, 0, c.stream>>>(c.M, F.n_oc, nvc, F.oc_cell, F.oc_face, F.u, F.oc_u);
  if (c.ndx == 3)
    k_mom_face<3><<<grid_for(nf), BS, 0, c.stream>>>(c.M, P, F.u, F.ub, F.ddx, F.spd, F.frc_dns, c.ngp, c.ngn, F.cor,
                                                     F.cfl, F.cinout, F.n_oc ? F.cell_oc : nullptr, F.oc_face, F.oc_u);
  else
    k_mom_face<2><<<grid_for(nf), BS, 0, c.stream>>>(c.M, P, F.u, F.ub, F.ddx, F.spd, F.frc_dns, c.ngp, c.ngn, F.cor,
                                                     F.cfl, F.cinout, F.n_oc ? F.cell_oc : nullptr, F.oc_face, F.oc_u);
  if (F.n_in + F.n_out) k_cinout<<<1, 1, 0, c.stream>>>(F.io_faces, F.n_in, F.n_out, F.cinout, F.sc);
  // the volume loop sees the outlet-projected cell velocities
  if (F.n_oc)
    k_outlet_post<<<grid_for(F.n_oc), BS, 0, c.stream>>>(F.n_oc, n, nvc, F.oc_cell, F.oc_face, F.oc_u, F.u);
  k_mom_cell<<<grid_for(n), BS, 0, c.stream>>>(c.M, P, nvc, F.u, F.ub, c.ngp, c.ngn, F.cor, F.cfl, F.dpdx, c.diag,
                                               c.spv, F.rhs, F.cs, F.frc_vol);
  FVK_HIP(hipGetLastError());
  c.pc_valid = false;
  rc = krylov(c, 1, F.cfg.imat_u, F.cfg.reduc_u, F.cfg.n_bicg, nvc, F.u, F.rhs, false);
  if (rc) return rc;
  k_mom_norm<<<1, 1, 0, c.stream>>>(it, nvc, c.ks, F.sc);
  FVK_HIP(hipGetLastError());
  return 0;
}

// ---- pressure correction (cfd_setup_pp), in phases split at its solves

static int pp_assemble(Flow &F) {
  Ctx &c = *F.c;
  const int n = c.n, nf = c.nf;
  k_plr<<<grid_for(n), BS, 0, c.stream>>>(n, c.diag, F.pl_r);
  k_crat<<<1, 1, 0, c.stream>>>(F.sc);
  if (c.ndx == 3)
    k_pp_face<3><<<grid_for(nf), BS, 0, c.stream>>>(c.M, F.p, F.dpdx, F.pl_r, F.frc_dns, F.spd, c.ngp, c.ngn, F.cpp, F.sc);
  else
    k_pp_face<2><<<grid_for(nf), BS, 0, c.stream>>>(c.M, F.p, F.dpdx, F.pl_r, F.frc_dns, F.spd, c.ngp, c.ngn, F.cpp, F.sc);
  k_pp_cell<<<grid_for(n), BS, 0, c.stream>>>(c.M, c.ngp, c.ngn, F.cpp, F.cs, c.diag, c.spv, F.rhs);
  if (fill(c, F.pc, n, 0.0)) return FVK_ERR_HIP;
  FVK_HIP(hipGetLastError());
  c.pc_valid = false;
  return 0;
}

// Mean relative row defect sum_i |sum_j a_ij| / a_ii: zero for the
// pure-Neumann (singular) pressure systems of closed domains.
template <int WS>
__global__ __launch_bounds__(RB) void k_row_defect(DMesh M, const double *__restrict__ diag,
                                                   const double *__restrict__ spv, double *__restrict__ part) {
  double a[1] = {0.0};
  int r0, r1;
  block_rows(M.n, r0, r1);
  for (int i = r0 + threadIdx.x; i < r1; i += RB) {
    double s = diag[i];
#pragma unroll
    for (int k = 0; k < nslots<WS>(M); ++k) {
      const int j = M.spc[(size_t)k * M.ld + i];
      const double v = spv[(size_t)k * M.ld + i];
      if (j >= 0) s += v;
    }
    a[0] += fabs(s) / fabs(diag[i]);
  }
  store_partials<1>(a, part);
}

__global__ __launch_bounds__(BS) void k_shift(int n, double *__restrict__ x, const double *__restrict__ sum) {
  const int i = blockIdx.x * BS + threadIdx.x;
  if (i < n) x[i] -= *sum / n;
}

// Whether the current matrix is singular with the constants as its null
// space (a pure-Neumann system of a closed domain or of one whose outlets
// fix no pressure): its rows sum to zero.
static int is_singular(Flow &F, bool &singular) {
  Ctx &c = *F.c;
  const int G = red_grid(c.n);
  double *part = c.part + (size_t)3 * RG * 8;
#define L(W) k_row_defect<W><<<G, RB, 0, c.stream>>>(c.M, c.diag, c.spv, part)
  switch (c.M.Ws) {
    case 3: L(3); break;
    case 4: L(4); break;
    case 6: L(6); break;
    default: L(0); break;
  }
#undef L
  k_sum_total<<<1, RB, 0, c.stream>>>(part, G, F.red_scratch);
  double d;
  FVK_HIP(hipMemcpyAsync(&d, F.red_scratch, sizeof d, hipMemcpyDeviceToHost, c.stream));
  FVK_HIP(hipStreamSynchronize(c.stream));
  singular = d / c.n <= 1e-12;
  return 0;
}

// x -= mean(x)
static int zero_mean(Flow &F, double *x) {
  Ctx &c = *F.c;
  int rc;
  if ((rc = sum_to(c, x, F.red_scratch))) return rc;
  k_shift<<<grid_for(c.n), BS, 0, c.stream>>>(c.n, x, F.red_scratch);
  FVK_HIP(hipGetLastError());
  return 0;
}

// A singular system is solved to tolerance on its consistent part: the
// mean of the right-hand side (the net mass imbalance, which no pressure
// correction can remove) is taken out of a copy of b, or the residual could
// never fall below it.  The solution is then defined up to a constant that
// depends on the solver's path, and the Green-Gauss gradient of a constant
// is not zero on warped cells (their face vectors do not close), so the
// fields would depend on the solver: such solutions are shifted to mean
// zero, whatever solver produced them.
struct NullSpace {
  bool singular = false;
  const double *b = nullptr;  // the right-hand side to solve with
};

static int null_begin(Flow &F, const double *b, NullSpace &N) {
  int rc;
  if ((rc = is_singular(F, N.singular))) return rc;
  N.b = b;
  if (!N.singular) return 0;
  // F.g0 (the gradient sweeps' scratch) is free during the pressure solve
  if ((rc = copy_d(*F.c, b, F.g0, F.c->n))) return rc;
  if ((rc = zero_mean(F, F.g0))) return rc;
  N.b = F.g0;
  return 0;
}

static int null_end(Flow &F, double *x, const NullSpace &N) { return N.singular ? zero_mean(F, x) : 0; }

// The pressure-correction (or SIMPLER pressure) solve: the reference's PCG
// (cfd_setup.f90:662,742) with p_solver 0, else a solve from x = 0 to an
// FP64 true relative L1 residual p_tol: 1 FP32-inner defect correction,
// 2 the same with the learned factor (plain FP64 PCG when it fails),
// 3 plain FP64 Jacobi PCG.
// solves to tolerance check convergence from the host: graphs stop there
static bool pp_host_driven(const Flow &F) { return F.cfg.p_solver != 0; }

static int pp_solve(Flow &F, double *x, double *b) {
  Ctx &c = *F.c;
  const int chunk = F.capture ? 0 : 64;
  // iteration cap of a solve to tolerance: far above what a solve that can
  // reach p_tol needs, low enough to end one that cannot
  const int cap = 20000;
  switch (F.cfg.p_solver) {
    case 0:
      return krylov(c, 0, F.cfg.imat_p, F.cfg.reduc_p, F.cfg.n_cg, 1, x, b, true,
                    F.cfg.n_cg > 64 ? chunk : 0);
    case 3: {
      KS h;
      NullSpace N;
      int rc = null_begin(F, b, N);
      if (!rc) rc = pcg_tol(c, 0, N.b, x, F.cfg.p_tol, cap, chunk, nullptr, h);
      return rc ? rc : null_end(F, x, N);
    }
    default: {
      int rc;
      NullSpace N;
      if ((rc = null_begin(F, b, N))) return rc;
      const bool fac = F.cfg.p_solver == 2;
      if (fac) {
        if (!F.model) { set_error("p_solver 2 needs a model (fvk_flow_set_model)"); return FVK_ERR_STATE; }
        if (F.train_steps > 0 && (rc = learned_train(c, F.model, F.train_steps, F.train_lr, nullptr))) return rc;
        if ((rc = learned_factor(c, F.model))) return rc;
      }
      DcResult R;
      if ((rc = dc_solve(c, fac, N.b, x, F.cfg.p_tol, cap, R, nullptr))) return rc;
      if (!R.converged) {
        // the plain FP64 Jacobi PCG from x = 0 takes over (with a cap of its
        // own)
        KS h;
        if ((rc = pcg_tol(c, 0, N.b, x, F.cfg.p_tol, cap, chunk, nullptr, h))) return rc;
        k_ks_store<<<1, 1, 0, c.stream>>>(c.ks, R.b_l1, R.iters + h.iter, h.res_n);
        FlowSc s;
        FVK_HIP(hipMemcpy(&s, F.sc, sizeof s, hipMemcpyDeviceToHost));
        s.fallbacks += 1;
        FVK_HIP(hipMemcpy(F.sc, &s, sizeof s, hipMemcpyHostToDevice));
      } else {
        k_ks_store<<<1, 1, 0, c.stream>>>(c.ks, R.b_l1, R.iters, R.res_l1);
      }
      FVK_HIP(hipGetLastError());
      return null_end(F, x, N);
    }
  }
}

__global__ __launch_bounds__(BS) void k_add_to(int n, double *__restrict__ p, const double *__restrict__ d) {
  const int i = blockIdx.x * BS + threadIdx.x;
  if (i < n) p[i] = p[i] + d[i];
}

// The SIMPLER pressure solve.  The reference's call solves A dp = rhs_p
// with phi => vol%p and not_r0, i.e. it adds its correction to the current
// pressure (cfd_setup.f90:736-742).  The reference recurrence (p_solver 0)
// does that in place; a solve to tolerance starts from x = 0, so it solves
// for the correction in F.pc (free after pp_post1) and adds it.
static int p_solve_simpler(Flow &F) {
  if (F.cfg.p_solver == 0) return pp_solve(F, F.p, F.rhs_p);
  int rc = pp_solve(F, F.pc, F.rhs_p);
  if (rc) return rc;
  k_add_to<<<grid_for(F.c->n), BS, 0, F.c->stream>>>(F.c->n, F.p, F.pc);
  FVK_HIP(hipGetLastError());
  return 0;
}

// after the p' solve: norms, relaxed update, p' gradient, velocity and flux
// corrections (and the SIMPLER pressure right-hand side)
static int pp_post1(Flow &F, int it) {
  Ctx &c = *F.c;
  const int n = c.n, nf = c.nf, nvc = F.nvc;
  int rc;
  if ((rc = sum_to(c, F.rhs, F.red_scratch))) return rc;
  k_pp_norm<<<1, 1, 0, c.stream>>>(it, 0, c.ks, F.red_scratch, F.sc);
  if (!F.cfg.simpler) k_p_relax<<<grid_for(n), BS, 0, c.stream>>>(n, F.cfg.urf_p, F.pc, F.row_of_cell1, F.p);
  rc = grad(c, 1, F.pc, nullptr, F.ddxp, F.cfg.n_grad, 2, nullptr, F.g0);
  if (rc) return rc;
  k_u_corr<<<grid_for(n), BS, 0, c.stream>>>(n, nvc, F.ddxp, F.pl_r, F.frc_vol, F.u);
  if (c.ndx == 3)
    k_flux_face<3><<<grid_for(nf), BS, 0, c.stream>>>(c.M, F.pc, F.ddxp, c.ngp, F.frc_dns, F.spd, F.cpp);
  else
    k_flux_face<2><<<grid_for(nf), BS, 0, c.stream>>>(c.M, F.pc, F.ddxp, c.ngp, F.frc_dns, F.spd, F.cpp);
  if (F.cfg.simpler) k_flux_cell<<<grid_for(n), BS, 0, c.stream>>>(c.M, F.cpp, F.rhs, F.rhs_p);
  FVK_HIP(hipGetLastError());
  return 0;
}

// after the SIMPLER pressure solve: its norms
static int pp_post2(Flow &F, int it) {
  Ctx &c = *F.c;
  int rc;
  if ((rc = sum_to(c, F.rhs_p, F.red_scratch))) return rc;
  k_pp_norm<<<1, 1, 0, c.stream>>>(it, 1, c.ks, F.red_scratch, F.sc);
  FVK_HIP(hipGetLastError());
  return 0;
}

static int pp_final(Flow &F) {
  Ctx &c = *F.c;
  return grad(c, 1, F.p, nullptr, F.dpdx, F.cfg.n_grad, 2, nullptr, F.g0);
}

int pp_setup(Flow &F, int it) {
  int rc;
  if ((rc = pp_assemble(F))) return rc;
  if ((rc = pp_solve(F, F.pc, F.rhs))) return rc;
  if ((rc = pp_post1(F, it))) return rc;
  if (F.cfg.simpler) {
    if ((rc = p_solve_simpler(F))) return rc;
    if ((rc = pp_post2(F, it))) return rc;
  }
  return pp_final(F);
}

int outer_begin(Flow &F) {
  k_outer_begin<<<1, 1, 0, F.c->stream>>>(F.sc);
  FVK_HIP(hipGetLastError());
  return 0;
}

// ---- HIP graphs of one outer iteration

// Segments of an outer iteration between host-driven solves (p_solver > 0):
// s = 0: outer_begin, momentum, p' assembly [, p' solve, post1[, p solve,
// post2], final, history] as far as no host-driven solve intervenes.
static int record_segment(Flow &F, int it, int seg) {
  int rc;
  bool skipped;
  const bool host = pp_host_driven(F);
  if (seg == 0) {
    if ((rc = outer_begin(F))) return rc;
    if ((rc = mom_setup(F, it, skipped))) return rc;
    if ((rc = pp_assemble(F))) return rc;
    if (host) return 0;
    if ((rc = pp_solve(F, F.pc, F.rhs))) return rc;
  }
  if (seg <= 1) {
    if ((rc = pp_post1(F, it))) return rc;
    if (F.cfg.simpler) {
      if (host) return 0;
      if ((rc = p_solve_simpler(F))) return rc;
    }
  }
  if (F.cfg.simpler && (seg <= 1 || seg == 2)) {
    if ((rc = pp_post2(F, it))) return rc;
  }
  if ((rc = pp_final(F))) return rc;
  k_hist_push<<<1, 1, 0, F.c->stream>>>(F.sc, F.hist, F.hist_cap);
  FVK_HIP(hipGetLastError());
  return 0;
}

static int capture(Flow &F, int it, int seg, hipGraphExec_t &exec) {
  Ctx &c = *F.c;
  hipGraph_t g = nullptr;
  F.capture = true;
  FVK_HIP(hipStreamBeginCapture(c.stream, hipStreamCaptureModeThreadLocal));
  int rc = record_segment(F, it, seg);
  hipError_t e = hipStreamEndCapture(c.stream, &g);
  F.capture = false;
  if (rc) {
    if (g) hipGraphDestroy(g);
    return rc;
  }
  if (e != hipSuccess) { set_error(std::string("graph capture: ") + hipGetErrorString(e)); return FVK_ERR_HIP; }
  e = hipGraphInstantiate(&exec, g, nullptr, nullptr, 0);
  hipGraphDestroy(g);
  if (e != hipSuccess) { set_error(std::string("graph instantiate: ") + hipGetErrorString(e)); return FVK_ERR_HIP; }
  return 0;
}

int flow_run_graph(Flow &F, int it0, int n, fvk_outer_hist *h, int &replayed) {
  Ctx &c = *F.c;
  replayed = 0;
  if (n <= 0) return 0;
  // history slots on the device
  if (F.hist_cap < n) {
    FlowSc *hp = nullptr;
    int rc = dalloc(c, &hp, (size_t)n);
    if (rc) return rc;
    F.hist = hp;
    F.hist_cap = n;
  }
  for (auto g : F.graphs) hipGraphExecDestroy(g);
  F.graphs.clear();
  k_hist_reset<<<1, 1, 0, c.stream>>>(F.sc);
  // iteration it0 eagerly: it can take the all-zero early return, and it
  // sets the first-iteration normalisations
  int rc;
  bool skipped = false;
  if ((rc = outer_begin(F))) return rc;
  if ((rc = mom_setup(F, it0, skipped))) return rc;
  if ((rc = pp_setup(F, it0))) return rc;
  k_hist_push<<<1, 1, 0, c.stream>>>(F.sc, F.hist, F.hist_cap);
  FVK_HIP(hipGetLastError());
  int it = it0 + 1;
  if (skipped) {
    // keep iterating eagerly until the flow moves
    for (; it < it0 + n && skipped; ++it) {
      if ((rc = outer_begin(F)) || (rc = mom_setup(F, it, skipped)) || (rc = pp_setup(F, it))) return rc;
      k_hist_push<<<1, 1, 0, c.stream>>>(F.sc, F.hist, F.hist_cap);
    }
  }
  if (it < it0 + n) {
    // captured once with an iteration number > 1 (only it == 1 changes what
    // the kernels do), replayed for the remaining iterations
    const bool host = pp_host_driven(F);
    const int nseg = host ? (F.cfg.simpler ? 3 : 2) : 1;
    F.graphs.resize(nseg, nullptr);
    for (int s = 0; s < nseg; ++s)
      if ((rc = capture(F, it, s, F.graphs[s]))) return rc;
    // A replayed segment assembles matrices (and rebuilds the momentum
    // solve's preconditioner) without the host code that marks the
    // preconditioner vector stale, so the host does that after each launch.
    for (; it < it0 + n; ++it) {
      FVK_HIP(hipGraphLaunch(F.graphs[0], c.stream));
      c.pc_valid = false;
      if (host) {
        if ((rc = pp_solve(F, F.pc, F.rhs))) return rc;
        FVK_HIP(hipGraphLaunch(F.graphs[1], c.stream));
        c.pc_valid = false;
        if (F.cfg.simpler) {
          if ((rc = p_solve_simpler(F))) return rc;
          FVK_HIP(hipGraphLaunch(F.graphs[2], c.stream));
          c.pc_valid = false;
        }
      }
      ++replayed;
    }
  }
  // history
  std::vector<FlowSc> hs(n);
  FVK_HIP(hipStreamSynchronize(c.stream));
  FVK_HIP(hipMemcpy(hs.data(), F.hist, sizeof(FlowSc) * n, hipMemcpyDeviceToHost));
  for (int k = 0; k < n; ++k) {
    const FlowSc &s = hs[k];
    if (h) {
      fvk_outer_hist &o = h[k];
      o.reduc = s.h_reduc; o.redpp = s.h_redpp; o.redpp_sum = s.h_redpp_sum;
      o.red_p = s.h_red_p; o.red_p_sum = s.h_red_p_sum; o.mx_res_n = s.mx_res_n;
      for (int q = 0; q < 3; ++q) { o.it_u[q] = s.it_u[q]; o.res0_n_u[q] = s.res0_n_u[q]; }
      o.it_pp = s.it_pp; o.it_p = s.it_p;
      o.res0_n_pp = s.res0_n_pp;
      o.res0_n_p = s.res0_n_p;
    }
    if (s.diverged) { set_error("momentum residual exceeded large"); return FVK_ERR_DIVERGED; }
  }
  return sweep_status(c);
}

}  // namespace fvk
