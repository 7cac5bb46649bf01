/*
 * fvo_cfd.c — oracle restatement of the SIMPLE layer (TEST INFRASTRUCTURE).
 *
 *   solv_grad      spray-simulation/cfd_setup.f90:765-866
 *   cfd_setup_eqn  spray-simulation/cfd_setup.f90:22-544   (iph=1, n_tl=1,
 *                  turb=spr=.false.: set_bnd/calc_rhs/calc_srhs are no-ops,
 *                  cfd_define.f90:118-337)
 *   cfd_setup_pp   spray-simulation/cfd_setup.f90:549-760
 *   outer loop     spray-simulation/cfd_solve.f90:134-178
 *
 * pinned: golden dumps of the flang-built reference (fields and print-line
 * histories after N outer iterations, shipped SIMPLER/DIC/DILU settings and
 * the relaxed/Jacobi variant), GG exactness for linear fields, the pure
 * diffusion stencil at zero velocity, and symmetry of the pressure matrix.
 */
#include "fvo.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define V1(n) ((double *)calloc((size_t)(n) + 1, sizeof(double)))
#define I1(n) ((int *)calloc((size_t)(n) + 1, sizeof(int)))

static double dotn(const double *a, const double *b, int n) {
  double s = 0.0;
  for (int d = 0; d < n; ++d) s = s + a[d] * b[d];
  return s;
}
static double magn(const double *a, int n) { return sqrt(dotn(a, a, n)); }
static double dmax(double a, double b) { return a > b ? a : b; }
static double dmin(double a, double b) { return a < b ? a : b; }

fvo_flow *fvo_flow_new(fvo_mesh *m, const fvo_cfg *cfg) {
  fvo_flow *f = (fvo_flow *)calloc(1, sizeof(fvo_flow));
  f->m = m;
  f->cfg = *cfg;
  f->n_vc = m->n_dx;
  int nv = m->n_vol, nb = m->n_vbf, nf = m->n_fce;
  f->u = V1((nb + 1) * f->n_vc);
  f->p = V1(nv);
  f->dpdx = V1(3 * (nv + 1));
  f->spd = V1(nf);
  f->frc_dns = V1(nf);
  f->cs = V1(nv);
  f->frc_vol = V1(nv);
  f->frc_c = V1(nv); f->dns_c = V1(nv); f->vis_c = V1(nv);
  f->frc_b = V1(nb); f->dns_b = V1(nb); f->vis_b = V1(nb);
  f->co_num = V1(nv);
  for (int i = 1; i <= nv; ++i) { f->frc_c[i] = 1.0; f->dns_c[i] = cfg->rho; f->vis_c[i] = cfg->mu; }
  for (int i = nv + 1; i <= nb; ++i) { f->frc_b[i] = 1.0; f->dns_b[i] = cfg->rho; f->vis_b[i] = cfg->mu; }
  f->A.m = m;
  f->A.pl = V1(nv);
  f->A.ng_p = V1(nf);
  f->A.ng_n = V1(nf);
  f->pc_last = V1(nv);
  f->rhs_last = V1(nv);
  return f;
}

void fvo_flow_free(fvo_flow *f) {
  if (!f) return;
  void *ptrs[] = {f->u, f->p, f->dpdx, f->spd, f->frc_dns, f->cs, f->frc_vol, f->frc_c, f->dns_c,
                  f->vis_c, f->frc_b, f->dns_b, f->vis_b, f->co_num, f->A.pl, f->A.ng_p, f->A.ng_n, f->pc_last, f->rhs_last};
  for (size_t k = 0; k < sizeof(ptrs) / sizeof(ptrs[0]); ++k) free(ptrs[k]);
  free(f);
}

void fvo_flow_set_bnd_u(fvo_flow *f, int code, const double *u3) {
  const fvo_mesh *m = f->m;
  for (int fc = 1; fc <= m->n_fce; ++fc)
    if (m->f_in[fc] == 0 && m->f_bty[fc] == code)
      for (int c = 0; c < f->n_vc; ++c) f->u[m->f_ib[fc] * f->n_vc + c] = u3[c];
}

/* solv_grad.  grad(ivol, ivc, idx) at grad[(ivol*n_vc + ivc)*3 + idx]. */
void fvo_grad(const fvo_mesh *m, int n_vc, const double *phi, double *grad, int n_iter, int neu) {
  int nd = m->n_dx, nv = m->n_vol;
  size_t gsz = (size_t)(nv + 1) * n_vc * 3;
  double *g0 = (double *)calloc(gsz, sizeof(double));
  for (int it = 1; it <= n_iter; ++it) {
    memset(grad, 0, gsz * sizeof(double));
    for (int fc = 1; fc <= m->n_fce; ++fc) {
      int ip = m->f_ip[fc], in = m->f_in[fc];
      const double *n = &m->f_n[3 * fc];
      double s = m->f_s[fc], w = m->f_wfp[fc];
      double pf[3];
      if (in > 0) {
        if (it > 1) {
          for (int c = 0; c < n_vc; ++c) {
            double pp = phi[ip * n_vc + c] + dotn(&g0[(ip * n_vc + c) * 3], &m->f_dxp[3 * fc], nd);
            double pn = phi[in * n_vc + c] + dotn(&g0[(in * n_vc + c) * 3], &m->f_dxn[3 * fc], nd);
            pf[c] = w * pn + (1.0 - w) * pp;
          }
        } else {
          for (int c = 0; c < n_vc; ++c) pf[c] = w * phi[in * n_vc + c] + (1.0 - w) * phi[ip * n_vc + c];
        }
        for (int c = 0; c < n_vc; ++c)
          for (int d = 0; d < nd; ++d) {
            grad[(ip * n_vc + c) * 3 + d] = grad[(ip * n_vc + c) * 3 + d] + pf[c] * s * n[d] * m->v_vr[ip];
            grad[(in * n_vc + c) * 3 + d] = grad[(in * n_vc + c) * 3 + d] - pf[c] * s * n[d] * m->v_vr[in];
          }
      } else {
        int ib = m->f_ib[fc];
        if (neu > 0) {
          for (int c = 0; c < n_vc; ++c) pf[c] = phi[ip * n_vc + c];
          if (it > 1 && neu == 2) {
            double t[3];
            for (int d = 0; d < nd; ++d) t[d] = m->f_x[3 * fc + d] - m->v_x[3 * ip + d];
            for (int c = 0; c < n_vc; ++c) pf[c] = pf[c] + dotn(&g0[(ip * n_vc + c) * 3], t, nd);
          }
        } else {
          for (int c = 0; c < n_vc; ++c) pf[c] = phi[ib * n_vc + c];
        }
        for (int c = 0; c < n_vc; ++c)
          for (int d = 0; d < nd; ++d)
            grad[(ip * n_vc + c) * 3 + d] = grad[(ip * n_vc + c) * 3 + d] + pf[c] * s * n[d] * m->v_vr[ip];
      }
    }
    if (it != n_iter) memcpy(g0, grad, gsz * sizeof(double));
  }
  free(g0);
}

/* cfd_setup_eqn for the momentum equation (iph=1, ieq=1, isc=1, vof=.true.) */
int fvo_setup_eqn(fvo_flow *F, int it) {
  fvo_mesh *m = F->m;
  const fvo_cfg *cf = &F->cfg;
  int nd = m->n_dx, nv = m->n_vol, nvc = F->n_vc;
  double sm = fvo_small(), lg = fvo_large();
  double urf = cf->urf_u, urf_r = 1.0 / urf, c_bl = cf->c_bl;
  double *pl = F->A.pl, *ngp = F->A.ng_p, *ngn = F->A.ng_n, *u = F->u;
  double *rhs = V1(nv * nvc); /* rhs(ivol, ivc) at [ivc*nv + ivol-1] */
  double *ddx = (double *)calloc((size_t)(nv + 1) * nvc * 3, sizeof(double));
  int *set = I1(nv);
  int *neu = I1(m->n_fce);
#define RHS(i, c) rhs[(c) * nv + (i) - 1]
#define DDX(i, c) (&ddx[((i) * nvc + (c)) * 3])
  for (int i = 1; i <= nv; ++i) pl[i] = 0.0;
  for (int i = 1; i <= nv; ++i) F->cs[i] = 0.0;
  double tmp = 0.0, phi_mx = 0.0;
  for (int i = 1; i <= nv; ++i) {
    for (int c = 0; c < nvc; ++c) tmp = dmax(tmp, fabs(u[i * nvc + c]));
    phi_mx = dmax(phi_mx, tmp);
  }
  if (phi_mx > sm) fvo_grad(m, nvc, u, ddx, cf->n_grad, 0);
  for (int fc = 1; fc <= m->n_fce; ++fc)
    if (m->f_in[fc] == 0) {
      int ib = m->f_ib[fc];
      for (int c = 0; c < nvc; ++c) tmp = dmax(tmp, fabs(u[ib * nvc + c]));
      phi_mx = dmax(phi_mx, tmp);
    }
  /* source loop: eqn%rhs (zero in the plain flow) is added, cfd_setup.f90:165-168 */
  for (int i = 1; i <= nv; ++i) {
    F->frc_vol[i] = F->frc_c[i] * m->v_v[i];
    for (int c = 0; c < nvc; ++c) RHS(i, c) = RHS(i, c) + 0.0;
  }
  double rhs_mx = 0.0;
  for (int k = 0; k < nv * nvc; ++k) rhs_mx = dmax(rhs_mx, fabs(rhs[k]));
  if (phi_mx < sm && rhs_mx < sm) {
    free(rhs); free(ddx); free(set); free(neu);
    return 1;
  }
  double pa[3], na[3], pfce[3], cimp[3] = {0, 0, 0}, cexp[3], cor[3], dimp[3], dexp[3], phu[3];
  for (int fc = 1; fc <= m->n_fce; ++fc) {
    int ip = m->f_ip[fc], in = m->f_in[fc];
    const double *n = &m->f_n[3 * fc];
    double s = m->f_s[fc], w = m->f_wfp[fc];
    for (int c = 0; c < nvc; ++c) pa[c] = u[ip * nvc + c] + dotn(DDX(ip, c), &m->f_dxp[3 * fc], nd);
    if (in > 0) {
      for (int c = 0; c < nvc; ++c) na[c] = u[in * nvc + c] + dotn(DDX(in, c), &m->f_dxn[3 * fc], nd);
      for (int c = 0; c < nvc; ++c) pfce[c] = w * na[c] + (1.0 - w) * pa[c];
      if (nvc > 1) F->spd[fc] = dotn(pfce, n, nd);
      double frc = w * F->frc_c[in] + (1.0 - w) * F->frc_c[ip];
      double dns = w * F->dns_c[in] + (1.0 - w) * F->dns_c[ip];
      double vis = w * F->vis_c[in] + (1.0 - w) * F->vis_c[ip];
      F->frc_dns[fc] = frc * dns;
      double frc_vis = frc * vis;
      double cc = 0.0, d = 0.0;
      for (int c = 0; c < nvc; ++c) cor[c] = 0.0;
      if (F->frc_dns[fc] > sm && fabs(F->spd[fc]) > sm) {
        cc = F->frc_dns[fc] * s * F->spd[fc];
        F->cs[ip] = F->cs[ip] + cc;
        F->cs[in] = F->cs[in] - cc;
        if (cc > 0.0) for (int c = 0; c < nvc; ++c) cimp[c] = u[ip * nvc + c];
        if (cc < 0.0) for (int c = 0; c < nvc; ++c) cimp[c] = u[in * nvc + c];
        if (cf->c_sch == 0) {
          for (int c = 0; c < nvc; ++c) cexp[c] = pfce[c];
        } else if (cf->c_sch > 0) {
          int ic;
          const double *phc, *phd;
          double dxpn[3];
          if (cc > 0.0) {
            ic = ip; phc = pa; phd = na;
            for (int d2 = 0; d2 < nd; ++d2)
              dxpn[d2] = (m->v_x[3 * in + d2] + m->f_dxn[3 * fc + d2]) - (m->v_x[3 * ip + d2] + m->f_dxp[3 * fc + d2]);
          } else {
            ic = in; phc = na; phd = pa;
            for (int d2 = 0; d2 < nd; ++d2)
              dxpn[d2] = (m->v_x[3 * ip + d2] + m->f_dxp[3 * fc + d2]) - (m->v_x[3 * in + d2] + m->f_dxn[3 * fc + d2]);
          }
          for (int c = 0; c < nvc; ++c) phu[c] = phd[c] - 2.0 * dotn(DDX(ic, c), dxpn, nd);
          if (cf->c_sch <= 3) {
            for (int c = 0; c < nvc; ++c) cexp[c] = (phc[c] - phu[c]) / (phd[c] - phc[c] + sm);
            if (cf->c_sch == 1) {
              for (int c = 0; c < nvc; ++c) cexp[c] = dmax(0.0, dmin(1.0, cexp[c]));
            } else if (cf->c_sch == 2) {
              for (int c = 0; c < nvc; ++c) cexp[c] = (cexp[c] + fabs(cexp[c])) / (1.0 + fabs(cexp[c]) + sm);
            } else if (cf->c_sch == 3) {
              for (int c = 0; c < nvc; ++c) {
                double a = dmax(0.0, dmin(1.0, 2.0 * cexp[c]));
                cexp[c] = dmax(a, dmin(2.0, cexp[c]));
              }
            }
            for (int c = 0; c < nvc; ++c) cexp[c] = phc[c] + 0.5 * cexp[c] * (phd[c] - phc[c]);
          } else {
            /* HRIC, :277-311 (co_num is zero in steady runs) */
            for (int c = 0; c < nvc; ++c) {
              double pcn = (phc[c] - phu[c]) / (phd[c] - phu[c] + sm);
              double pfn = pcn;
              if (pcn > 0.0 && pcn < 0.5) pfn = 2.0 * pcn;
              else if (pcn > 0.5 && pcn < 1.0) pfn = 1.0;
              double co = F->co_num[ic];
              if (co > 0.3 && co < 0.7) pfn = pcn + (pfn - pcn) * (0.7 - co) / (0.7 - 0.3);
              else if (co > 0.7) pfn = pcn;
              double th = dotn(DDX(ic, c), n, nd) / (magn(DDX(ic, c), nd) + sm);
              th = sqrt(fabs(th));
              pfn = pfn * th + pcn * (1.0 - th);
              double gm = (phd[c] - phu[c]) / (phd[c] - phc[c] + sm);
              gm = gm * (1.0 - pfn);
              gm = dmax(dmin(gm, 1.0), 0.0);
              cexp[c] = gm * phc[c] + (1.0 - gm) * phd[c];
            }
          }
        }
        for (int c = 0; c < nvc; ++c) cor[c] = c_bl * cc * (cimp[c] - cexp[c]);
      }
      if (frc_vis > sm) {
        d = frc_vis * s / m->f_lpn[fc];
        for (int c = 0; c < nvc; ++c) {
          dimp[c] = u[in * nvc + c] - u[ip * nvc + c];
          dexp[c] = na[c] - pa[c];
          cor[c] = cor[c] + d * (dexp[c] - dimp[c]);
        }
      }
      ngp[fc] = -d + dmin(cc, 0.0);
      ngn[fc] = -d - dmax(cc, 0.0);
      if (set[ip]) ngp[fc] = 0.0;
      if (set[in]) ngn[fc] = 0.0;
      pl[ip] = pl[ip] - ngp[fc];
      pl[in] = pl[in] - ngn[fc];
      for (int c = 0; c < nvc; ++c) {
        RHS(ip, c) = RHS(ip, c) + cor[c];
        RHS(in, c) = RHS(in, c) - cor[c];
      }
    } else {
      int ib = m->f_ib[fc];
      int bty = m->f_bty[fc] - m->f_bty[fc] % 10;
      double *pb = &u[ib * nvc], *pp = &u[ip * nvc];
      if (nvc > 1) {
        F->spd[fc] = dotn(pb, n, nd);
        if (bty == 20 && dotn(pp, n, nd) < -sm) {
          double t[3], dd;
          for (int c = 0; c < nvc; ++c) t[c] = pp[c];
          dd = dotn(t, n, nd);
          for (int c = 0; c < nvc; ++c) t[c] = t[c] - n[c] * dd;
          dd = magn(t, nvc) + sm;
          for (int c = 0; c < nvc; ++c) t[c] = t[c] / dd;
          dd = dotn(pp, t, nvc);
          for (int c = 0; c < nvc; ++c) pp[c] = t[c] * dd;
          F->spd[fc] = dotn(pp, n, nd);
        }
      }
      double frc = F->frc_b[ib], dns = F->dns_b[ib], vis = F->vis_b[ib];
      F->frc_dns[fc] = frc * dns;
      double frc_vis = frc * vis;
      double cc = 0.0, d = 0.0;
      for (int c = 0; c < nvc; ++c) cor[c] = 0.0;
      if (F->frc_dns[fc] > sm && fabs(F->spd[fc]) > sm) {
        cc = F->frc_dns[fc] * s * F->spd[fc];
        F->cs[ip] = F->cs[ip] + cc;
        if (bty == 10) F->c_in = F->c_in - frc * s * F->spd[fc];
        if (bty == 20) F->c_out = F->c_out + frc * s * F->spd[fc];
      }
      if (frc_vis > sm) {
        d = frc_vis * s / m->f_lpn[fc];
        for (int c = 0; c < nvc; ++c) {
          dimp[c] = pb[c] - pp[c];
          dexp[c] = pb[c] - pa[c];
        }
      }
      ngp[fc] = -d + dmin(cc, 0.0);
      if (set[ip]) ngp[fc] = 0.0;
      if (bty == 40) {
        if (nvc == 1) {
          neu[fc] = 1;
        } else {
          double dd = dotn(pp, n, nd);
          for (int c = 0; c < nvc; ++c) pb[c] = pp[c] - n[c] * dd;
          if (d > sm) {
            dd = dotn(pp, n, nd);
            for (int c = 0; c < nvc; ++c) dexp[c] = -(2.0 * n[c] * dd);
          }
        }
      }
      if (bty == 30 && nvc > 1 && d > sm) {
        double t[3], dd;
        for (int c = 0; c < nvc; ++c) t[c] = pp[c] + sm;
        if (magn(pb, nvc) > sm) for (int c = 0; c < nvc; ++c) t[c] = pb[c];
        dd = dotn(t, n, nd);
        for (int c = 0; c < nvc; ++c) t[c] = t[c] - n[c] * dd;
        dd = magn(t, nvc) + sm;
        for (int c = 0; c < nvc; ++c) t[c] = t[c] / dd;
        dd = dotn(pb, t, nvc);
        for (int c = 0; c < nvc; ++c) pb[c] = t[c] * dd;
        dd = dotn(pp, t, nvc);
        for (int c = 0; c < nvc; ++c) dexp[c] = pb[c] - t[c] * dd;
      }
      if (bty == 20 || neu[fc]) {
        for (int c = 0; c < nvc; ++c) pb[c] = pp[c];
        ngp[fc] = 0.0;
      }
      if (d > sm)
        for (int c = 0; c < nvc; ++c) cor[c] = d * (dexp[c] - dimp[c]);
      pl[ip] = pl[ip] - ngp[fc];
      for (int c = 0; c < nvc; ++c) RHS(ip, c) = RHS(ip, c) + cor[c];
      for (int c = 0; c < nvc; ++c) RHS(ip, c) = RHS(ip, c) - ngp[fc] * pb[c];
    }
  }
  /* volume loop, :437-519 */
  for (int i = 1; i <= nv; ++i) {
    F->frc_vol[i] = F->frc_c[i] * m->v_v[i];
    if (nvc > 1)
      for (int c = 0; c < nvc; ++c) RHS(i, c) = RHS(i, c) - F->dpdx[3 * i + c] * F->frc_vol[i];
    if (set[i]) {
      for (int c = 0; c < nvc; ++c) RHS(i, c) = u[i * nvc + c];
      pl[i] = 1.0;
    }
    pl[i] = urf_r * pl[i];
    for (int c = 0; c < nvc; ++c) RHS(i, c) = RHS(i, c) + (1.0 - urf) * pl[i] * u[i * nvc + c];
  }
  double res0_n[3] = {0, 0, 0};
  int iters[3] = {0, 0, 0};
  fvo_solv_bicg(&F->A, cf->imat_u, cf->reduc_u, cf->n_bicg, nvc, u, nvc, rhs, res0_n, 0, iters, NULL);
  for (int c = 0; c < nvc; ++c) { F->res0_n_u[c] = res0_n[c]; F->it_u[c] = iters[c]; }
  if (it == 1) F->res_n_u = 0.0;
  double mx0 = res0_n[0];
  for (int c = 1; c < nvc; ++c) mx0 = dmax(mx0, res0_n[c]);
  if (mx0 > sm) {
    F->res_n_u = dmax(F->res_n_u, mx0);
    mx0 = mx0 / F->res_n_u;
    F->mx_res_n = dmax(F->mx_res_n, mx0);
  }
  F->h_reduc = mx0;
  free(rhs); free(ddx); free(set); free(neu);
#undef RHS
#undef DDX
  if (fabs(mx0) > lg) return -1; /* the reference stops here, :543 */
  return 0;
}

static void pp_norm(fvo_flow *F, int it, double r0n, double *out) {
  double sm = fvo_small();
  double mx0 = r0n + sm;
  if (it == 1) F->pp_res_n = 0.0;
  F->pp_res_n = dmax(F->pp_res_n, mx0);
  mx0 = mx0 / F->pp_res_n;
  F->mx_res_n = dmax(F->mx_res_n, mx0);
  *out = mx0;
}

/* cfd_setup_pp with n_pp=1 */
int fvo_setup_pp(fvo_flow *F, int it) {
  fvo_mesh *m = F->m;
  const fvo_cfg *cf = &F->cfg;
  int nd = m->n_dx, nv = m->n_vol, nvc = F->n_vc;
  double sm = fvo_small();
  double *pl = F->A.pl, *ngp = F->A.ng_p, *ngn = F->A.ng_n;
  double *pl_r = V1(nv), *rhs = V1(nv), *rhs_p = V1(nv), *pc = V1(nv);
  double *ddx = (double *)calloc((size_t)(nv + 1) * 3, sizeof(double));
  for (int i = 1; i <= nv; ++i) { pl_r[i] = 1.0 / (pl[i] + sm); pl[i] = 0.0; }
  for (int i = 1; i <= nv; ++i) rhs[i] = -F->cs[i];
  double delta = 1.0 / 1e10;
  double c_rat = F->c_out / (F->c_in + delta);
  if (c_rat < delta) c_rat = 1.0;
  c_rat = 1.0 / c_rat;
  if (fabs(F->c_in) < delta) c_rat = 0.0;
  for (int fc = 1; fc <= m->n_fce; ++fc) {
    int ip = m->f_ip[fc], in = m->f_in[fc];
    if (in > 0) {
      double w = m->f_wfp[fc];
      double pa = F->p[ip] + dotn(&F->dpdx[3 * ip], &m->f_dxp[3 * fc], nd);
      double na = F->p[in] + dotn(&F->dpdx[3 * in], &m->f_dxn[3 * fc], nd);
      double dp = na - pa;
      double gf[3], dxc[3];
      for (int d = 0; d < nd; ++d) gf[d] = w * F->dpdx[3 * in + d] + (1.0 - w) * F->dpdx[3 * ip + d];
      for (int d = 0; d < nd; ++d) dxc[d] = m->v_x[3 * in + d] - m->v_x[3 * ip + d];
      dp = dp - dotn(gf, dxc, nd);
      double p_r = w * pl_r[in] + (1.0 - w) * pl_r[ip];
      double v = w * m->v_v[in] + (1.0 - w) * m->v_v[ip];
      double d = F->frc_dns[fc] * p_r * m->f_s[fc] * v / m->f_lpn[fc];
      double c = -(d * dp);
      ngp[fc] = -d;
      ngn[fc] = -d;
      pl[ip] = pl[ip] - ngp[fc];
      pl[in] = pl[in] - ngn[fc];
      rhs[ip] = rhs[ip] - c;
      rhs[in] = rhs[in] + c;
    } else {
      int bty = m->f_bty[fc] - m->f_bty[fc] % 10;
      if (bty == 20) {
        double c = F->frc_dns[fc] * m->f_s[fc] * F->spd[fc];
        c = c * (c_rat - 1.0);
        rhs[ip] = rhs[ip] - c;
      }
    }
  }
  /* p' solve, :656-672 */
  double r0n[1];
  int its[1];
  for (int i = 1; i <= nv; ++i) pc[i] = 0.0;
  if (cf->p_solver == 3)
    fvo_pcg_tol(&F->A, rhs, pc, cf->p_tol, cf->p_cap, its, r0n);
  else
    fvo_solv_cg(&F->A, cf->imat_p, cf->reduc_p, cf->n_cg, 1, pc, 1, rhs + 1, r0n, 1, its, NULL);
  F->res0_n_pp = r0n[0];
  F->it_pp = its[0];
  for (int i = 1; i <= nv; ++i) { F->pc_last[i] = pc[i]; F->rhs_last[i] = rhs[i]; }
  pp_norm(F, it, r0n[0], &F->h_redpp);
  double sr = 0.0;
  for (int i = 1; i <= nv; ++i) sr = sr + rhs[i];
  F->h_redpp_sum = sr;
  if (!cf->simpler)
    for (int i = 1; i <= nv; ++i) F->p[i] = F->p[i] + cf->urf_p * (pc[i] - pc[1]);
  fvo_grad(m, 1, pc, ddx, cf->n_grad, 2);
  for (int i = 1; i <= nv; ++i)
    for (int c = 0; c < nvc; ++c)
      F->u[i * nvc + c] = F->u[i * nvc + c] - ddx[3 * i + c] * pl_r[i] * F->frc_vol[i];
  if (cf->simpler) for (int i = 1; i <= nv; ++i) rhs_p[i] = rhs[i];
  for (int i = 1; i <= nv; ++i) rhs[i] = 0.0;
  for (int fc = 1; fc <= m->n_fce; ++fc) {
    int ip = m->f_ip[fc], in = m->f_in[fc];
    if (in <= 0) continue;
    double pax = dotn(&ddx[3 * ip], &m->f_dxp[3 * fc], nd);
    double pa = pc[ip] + pax;
    double nax = dotn(&ddx[3 * in], &m->f_dxn[3 * fc], nd);
    double na = pc[in] + nax;
    double dp = ngp[fc] * (nax - pax);
    rhs[ip] = rhs[ip] - dp;
    rhs[in] = rhs[in] + dp;
    dp = ngp[fc] * (na - pa) / (F->frc_dns[fc] * m->f_s[fc]);
    F->spd[fc] = F->spd[fc] + dp;
  }
  if (cf->simpler) for (int i = 1; i <= nv; ++i) rhs_p[i] = rhs_p[i] + rhs[i];
  /* pressure, :736-759 */
  if (cf->simpler) {
    if (cf->p_solver == 3) {
      /* the reference's call adds the solution of A dp = rhs_p to p
         (phi => vol%p with not_r0, cfd_setup.f90:736-742) */
      fvo_pcg_tol(&F->A, rhs_p, pc, cf->p_tol, cf->p_cap, its, r0n);
      for (int i = 1; i <= nv; ++i) F->p[i] = F->p[i] + pc[i];
    } else {
      fvo_solv_cg(&F->A, cf->imat_p, cf->reduc_p, cf->n_cg, 1, F->p, 1, rhs_p + 1, r0n, 1, its, NULL);
    }
    F->res0_n_p = r0n[0];
    F->it_p = its[0];
    pp_norm(F, it, r0n[0], &F->h_red_p);
    sr = 0.0;
    for (int i = 1; i <= nv; ++i) sr = sr + rhs_p[i];
    F->h_red_p_sum = sr;
  }
  fvo_grad(m, 1, F->p, ddx, cf->n_grad, 2);
  for (int i = 1; i <= nv; ++i)
    for (int d = 0; d < nd; ++d) F->dpdx[3 * i + d] = ddx[3 * i + d];
  free(pl_r); free(rhs); free(rhs_p); free(pc); free(ddx);
  return 0;
}

int fvo_outer(fvo_flow *F, int it, double *hist5) {
  F->mx_res_n = 0.0;
  F->c_in = 0.0;
  F->c_out = 0.0;
  F->h_reduc = F->h_redpp = F->h_red_p = 0.0;
  int rc = fvo_setup_eqn(F, it);
  if (rc < 0) return rc;
  fvo_setup_pp(F, it);
  if (hist5) {
    hist5[0] = F->h_reduc;
    hist5[1] = F->h_redpp;
    hist5[2] = F->h_red_p;
    hist5[3] = F->mx_res_n;
    hist5[4] = F->h_redpp_sum;
  }
  return 0;
}
