// FP32-inner CG inside FP64 defect correction, and the factorised
// preconditioner z = G G^T r with G = D^-1/2 (I + L) Delta.
//
// No counterpart in the reference: spray-simulation solves its pressure
// systems in FP64 with solv_cg.f90 / solv_precond.f90.  This path keeps the
// reference's PCG recurrence (the same scalar decisions, cgcore.h) but runs
// the inner SpMV, preconditioner and vector updates in FP32, with FP64 dot
// products, FP64 alpha/beta and an FP64 true residual b - A x on the FP64
// matrix between inner solves:
//
//   x = 0
//   repeat: r = b - A x (FP64); stop if |r|_1 / |b|_1 < tol
//           solve Ah eh = D^-1/2 r in FP32 to relative L1 reduction
//             eta = max(inner_reduc, 0.3 tol / (|r|_1 / |b|_1))
//           x += D^-1/2 eh (FP64)
//
// The inner system is the Jacobi-scaled one, Ah = D^-1/2 A D^-1/2 (unit
// diagonal, FP32 off-diagonals, 16-bit column offsets where they fit): its
// SpMV reads no diagonal, Jacobi PCG on A is plain CG on Ah (z = r), and the
// factor becomes Gh = (I + L) Delta with zh = Gh Gh^T rh (L in FP16).
// Per inner iteration three kernels (factor) or two (Jacobi):
//   k_mx_pq   decide beta; p = z + beta p; q = Ah p; partial p.q
//   k_mx_upd  alpha; e += alpha p; r' = r - alpha q; partial |r'|_1 and
//             Jacobi: partial r'.r'
//             factor: t = Delta^2 (r' + L^T r'), formed on the fly for the
//             higher neighbours (r is double-buffered)
//   k_mx_gz   factor: z = t + L t, partial r'.z
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "cgcore.h"

namespace fvk {

typedef _Float16 lval_t;

// Column of slot k of row i: 32-bit rows, or 16-bit offsets from the row.
__device__ __forceinline__ int colat(const int *c, size_t k, int i) { return c[k]; }
__device__ __forceinline__ int colat(const short *c, size_t k, int i) { return i + (int)c[k]; }

template <int W>
__device__ __forceinline__ int nw(int w) { return W > 0 ? W : w; }

__device__ __forceinline__ bool gate_done(const KS *g) { return g && g->done; }

// Jacobi-scaled FP32 copy of the matrix, w = a_ii^-1/2
__global__ __launch_bounds__(BS) void k_mx_conv(DMesh M, const double *__restrict__ diag,
                                                const double *__restrict__ spv, float *__restrict__ s32,
                                                float *__restrict__ w) {
  const int i = blockIdx.x * BS + threadIdx.x;
  const int n = M.n;
  if (i >= n) return;
  const double wi = 1.0 / sqrt(diag[i]);
  w[i] = (float)wi;
  for (int k = 0; k < M.Ws; ++k) {
    const size_t o = (size_t)k * n + i;
    const int j = M.spc[o];
    s32[o] = j >= 0 ? (float)(spv[o] * wi / sqrt(diag[j])) : 0.f;
  }
}

// FP64 true residual r = b - A x, its scaled FP32 copy D^-1/2 r, partials
// {|r|_1, |r|_2^2, |b|_1, |b|_2^2}
template <int WS>
__global__ __launch_bounds__(RB) void k_mx_res(DMesh M, const double *__restrict__ diag,
                                               const double *__restrict__ spv, const double *__restrict__ x,
                                               const double *__restrict__ b, float *__restrict__ r32,
                                               double *__restrict__ part) {
  double a[4] = {0, 0, 0, 0};
  int r0, r1;
  block_rows(M.n, r0, r1);
  for (int i = r0 + threadIdx.x; i < r1; i += RB) {
    const double bi = b[i], ri = bi - row_spmv<WS>(M, diag, spv, x, i);
    r32[i] = (float)(ri / sqrt(diag[i]));
    a[0] += fabs(ri);
    a[1] += ri * ri;
    a[2] += fabs(bi);
    a[3] += bi * bi;
  }
  store_partials<4>(a, part);
}

// Start of an inner solve: e = 0, p_old = 0, partials |r|_1 [0] and
// Jacobi: r.r [3]; factor: t = Delta^2 (r + L^T r).
template <int WU, class I>
__global__ __launch_bounds__(RB) void k_mx_start(int n, int fac, const float *__restrict__ r,
                                                 float *__restrict__ ev, float *__restrict__ p0, int wu,
                                                 const I *__restrict__ ucol, const lval_t *__restrict__ uval,
                                                 const float *__restrict__ ee, float *__restrict__ t,
                                                 double *__restrict__ part) {
  double a[4] = {0, 0, 0, 0};
  int r0, r1;
  block_rows(n, r0, r1);
  for (int i = r0 + threadIdx.x; i < r1; i += RB) {
    const float ri = r[i];
    ev[i] = 0.f;
    p0[i] = 0.f;
    a[0] += fabs((double)ri);
    if (!fac) {
      a[3] += (double)ri * (double)ri;
    } else {
      float acc = ri;
#pragma unroll
      for (int k = 0; k < nw<WU>(wu); ++k) {
        const size_t o = (size_t)k * n + i;
        acc += (float)uval[o] * r[colat(ucol, o, i)];
      }
      t[i] = ee[i] * acc;
    }
  }
  store_partials<4>(a, part);
}

// Decide iteration it of the inner PCG; p = z + beta p ; q = Ah p ; p.q
template <int WS, class I>
__device__ __forceinline__ void mx_pq_row(int n, int i, const I *__restrict__ col, const float *__restrict__ spv,
                                          const float *__restrict__ z, const float *__restrict__ po, float bf,
                                          float *__restrict__ pn, float *__restrict__ q, int ws, double &pq) {
  const float pi = z[i] + bf * po[i];
  float y = pi;
#pragma unroll
  for (int k = 0; k < nw<WS>(ws); ++k) {
    const size_t o = (size_t)k * n + i;
    const int j = colat(col, o, i);
    y += spv[o] * (z[j] + bf * po[j]);
  }
  pn[i] = pi;
  q[i] = y;
  pq += (double)pi * (double)y;
}

template <int WS, class I>
__global__ __launch_bounds__(RB) void k_mx_pq(int n, int ws, int it, const I *__restrict__ col,
                                              const float *__restrict__ spv, const float *__restrict__ z,
                                              const float *__restrict__ pold, float *__restrict__ pnew,
                                              float *__restrict__ q, const double *__restrict__ partU, int PU,
                                              const double *__restrict__ partR, int PR, double *__restrict__ partQ,
                                              KS *ring, KS *out) {
  double L1[1], rz[1];
  if (partR) {
    load_totals<1>(partU, PU, L1);
    load_totals<1>(partR, PR, rz);
  } else {
    const int cs[2] = {0, 3};
    double t[2];
    load_sel<2>(partU, PU, cs, t);
    L1[0] = t[0];
    rz[0] = t[1];
  }
  __shared__ KS sk[1];
  if (threadIdx.x == 0) {
    const KS *prev = ring + ((it + 1) & 1) * 3;
    cg_decide(it, 1, L1, rz, prev, sk);
    if (prev->pad) sk[0].done = 1;  // breakdown flagged by the last update
    if (blockIdx.x == 0) ring[(it & 1) * 3] = out[0] = sk[0];
  }
  __syncthreads();
  if (sk[0].done) return;
  const float bf = (float)sk[0].beta;
  double pq = 0.0;
  int r0, r1;
  block_rows(n, r0, r1);
  int i = r0 + threadIdx.x;
  for (; i + RB < r1; i += 2 * RB) {
    mx_pq_row<WS, I>(n, i, col, spv, z, pold, bf, pnew, q, ws, pq);
    mx_pq_row<WS, I>(n, i + RB, col, spv, z, pold, bf, pnew, q, ws, pq);
  }
  if (i < r1) mx_pq_row<WS, I>(n, i, col, spv, z, pold, bf, pnew, q, ws, pq);
  double v[1] = {pq};
  store_partials<1>(v, partQ);
}

// alpha = s/(p.q + small) from the partials; false on breakdown (a
// non-positive or non-finite p.q), flagged for the next decision.
__device__ __forceinline__ bool mx_alpha(KS *cur, const double *partQ, int PQ, double &alpha) {
  double tq[1];
  load_totals<1>(partQ, PQ, tq);
  alpha = cur->s / (tq[0] + SMALL);
  if (!(tq[0] > 0.0) || !isfinite(alpha)) {
    if (blockIdx.x == 0 && threadIdx.x == 0) cur->pad = 1;
    return false;
  }
  return true;
}

// Jacobi: e += alpha p ; r -= alpha q (in place) ; partials |r|_1, r.r
__global__ __launch_bounds__(RB) void k_mx_upd_jac(int n, int it, float *__restrict__ ev, float *__restrict__ r,
                                                   const float *__restrict__ p, const float *__restrict__ q,
                                                   const double *__restrict__ partQ, int PQ,
                                                   double *__restrict__ partU, KS *ring) {
  KS *cur = ring + (it & 1) * 3;
  if (cur->done) return;
  double alpha;
  if (!mx_alpha(cur, partQ, PQ, alpha)) return;
  const float af = (float)alpha;
  double a[4] = {0, 0, 0, 0};
  int r0, r1;
  block_rows(n, r0, r1);
  auto row = [&](int i) {
    ev[i] += af * p[i];
    const float ri = r[i] - af * q[i];
    r[i] = ri;
    a[0] += fabs((double)ri);
    a[3] += (double)ri * (double)ri;
  };
  int i = r0 + threadIdx.x;
  for (; i + RB < r1; i += 2 * RB) {
    row(i);
    row(i + RB);
  }
  if (i < r1) row(i);
  store_partials<4>(a, partU);
  if (blockIdx.x == 0 && threadIdx.x == 0) cur->alpha = alpha;
}

// Factor: e += alpha p ; r' = r - alpha q ; t = Delta^2 (r' + L^T r') with
// r' of the higher neighbours formed from (r, q) ; partial |r'|_1
template <int WU, class I>
__global__ __launch_bounds__(RB) void k_mx_upd_fac(int n, int it, float *__restrict__ ev,
                                                   const float *__restrict__ rin, float *__restrict__ rout,
                                                   const float *__restrict__ p, const float *__restrict__ q, int wu,
                                                   const I *__restrict__ ucol, const lval_t *__restrict__ uval,
                                                   const float *__restrict__ ee, float *__restrict__ t,
                                                   const double *__restrict__ partQ, int PQ,
                                                   double *__restrict__ partU, KS *ring) {
  KS *cur = ring + (it & 1) * 3;
  if (cur->done) return;
  double alpha;
  if (!mx_alpha(cur, partQ, PQ, alpha)) return;
  const float af = (float)alpha;
  double a[4] = {0, 0, 0, 0};
  int r0, r1;
  block_rows(n, r0, r1);
  auto row = [&](int i) {
    ev[i] += af * p[i];
    const float ri = rin[i] - af * q[i];
    rout[i] = ri;
    a[0] += fabs((double)ri);
    float acc = ri;
#pragma unroll
    for (int k = 0; k < nw<WU>(wu); ++k) {
      const size_t o = (size_t)k * n + i;
      const int j = colat(ucol, o, i);
      acc += (float)uval[o] * (rin[j] - af * q[j]);
    }
    t[i] = ee[i] * acc;
  };
  int i = r0 + threadIdx.x;
  for (; i + RB < r1; i += 2 * RB) {
    row(i);
    row(i + RB);
  }
  if (i < r1) row(i);
  store_partials<4>(a, partU);
  if (blockIdx.x == 0 && threadIdx.x == 0) cur->alpha = alpha;
}

// z = t + L t, partial r.z
template <int WL, class I>
__global__ __launch_bounds__(RB) void k_mx_gz(int n, int wl, const I *__restrict__ lcol,
                                              const lval_t *__restrict__ lval, const float *__restrict__ t,
                                              const float *__restrict__ r, float *__restrict__ z,
                                              double *__restrict__ partR, const KS *gate) {
  if (gate_done(gate)) return;
  double a[1] = {0.0};
  int r0, r1;
  block_rows(n, r0, r1);
  auto row = [&](int i) {
    float acc = t[i];
#pragma unroll
    for (int k = 0; k < nw<WL>(wl); ++k) {
      const size_t o = (size_t)k * n + i;
      acc += (float)lval[o] * t[colat(lcol, o, i)];
    }
    z[i] = acc;
    a[0] += (double)r[i] * (double)acc;
  };
  int i = r0 + threadIdx.x;
  for (; i + RB < r1; i += 2 * RB) {
    row(i);
    row(i + RB);
  }
  if (i < r1) row(i);
  store_partials<1>(a, partR);
}

// Final decision after the last queued inner iteration (one block).
__global__ __launch_bounds__(RB) void k_mx_final(int it, const double *partU, int PU, const double *partR, int PR,
                                                 KS *ring, KS *out) {
  double L1[1], rz[1];
  if (partR) {
    load_totals<1>(partU, PU, L1);
    load_totals<1>(partR, PR, rz);
  } else {
    const int cs[2] = {0, 3};
    double t[2];
    load_sel<2>(partU, PU, cs, t);
    L1[0] = t[0];
    rz[0] = t[1];
  }
  if (threadIdx.x == 0) {
    KS cur[1];
    const KS *prev = ring + ((it + 1) & 1) * 3;
    cg_decide(it, 1, L1, rz, prev, cur);
    if (prev->pad) cur[0].done = 1;
    out[0] = cur[0];
  }
}

__global__ __launch_bounds__(RB) void k_mx_tot4(const double *part, int P, double *out) {
  double s[4];
  load_totals<4>(part, P, s);
  if (threadIdx.x == 0)
    for (int j = 0; j < 4; ++j) out[j] = s[j];
}

// a *= s (the inner right-hand side normalised to mean |r| = 1)
__global__ __launch_bounds__(BS) void k_scale32(int n, float *__restrict__ a, float s) {
  const int i = blockIdx.x * BS + threadIdx.x;
  if (i < n) a[i] *= s;
}

// x += sigma D^-1/2 eh
__global__ __launch_bounds__(BS) void k_mx_add(int n, double *__restrict__ x, const float *__restrict__ ev,
                                               const double *__restrict__ diag, double sigma) {
  const int i = blockIdx.x * BS + threadIdx.x;
  if (i < n) x[i] += sigma * (double)ev[i] / sqrt(diag[i]);
}

// Factor from per-interior-face (l, sig): e = exp(2 s) with s the sum of sig
// over the faces the row owns, slot values of L (FP16).
__global__ __launch_bounds__(BS) void k_fac_rows(int n, int WL, int WU, const int *__restrict__ lfi,
                                                 const int *__restrict__ ufi, const float *__restrict__ fl,
                                                 const float *__restrict__ fsig, lval_t *__restrict__ lval,
                                                 lval_t *__restrict__ uval, float *__restrict__ ee) {
  const int i = blockIdx.x * BS + threadIdx.x;
  if (i >= n) return;
  float s = 0.f;
  for (int k = 0; k < WU; ++k) {
    const int f = ufi[(size_t)k * n + i];
    uval[(size_t)k * n + i] = (lval_t)(f >= 0 ? fl[f] : 0.f);
    if (f >= 0) s += fsig[f];
  }
  for (int k = 0; k < WL; ++k) {
    const int f = lfi[(size_t)k * n + i];
    lval[(size_t)k * n + i] = (lval_t)(f >= 0 ? fl[f] : 0.f);
  }
  ee[i] = expf(2.f * s);
}

// ------------------------------------------------------------ host side

static inline double *pb(Ctx &c, int k) { return c.part + (size_t)k * RG * 8; }

#define FVK_W(ws, F) \
  switch (ws) {      \
    case 3: F(3); break; \
    case 4: F(4); break; \
    case 6: F(6); break; \
    default: F(0); break; \
  }

int mix_build(Ctx &c) {
  Mix &X = c.mx;
  if (X.built) return 0;
  const int n = c.n, nf = c.nf, Ws = c.Ws;
  // A's interior slots back on the host
  std::vector<int> spc((size_t)Ws * n);
  FVK_HIP(hipStreamSynchronize(c.stream));
  FVK_HIP(hipMemcpy(spc.data(), c.M.spc, spc.size() * sizeof(int), hipMemcpyDeviceToHost));
  // interior faces in face order; per-row lower / upper lists
  std::vector<int> face_fi(nf, -1), fi_face, fi_p, fi_n;
  std::vector<int> nl(n, 0), nu(n, 0);
  for (int f = 0; f < nf; ++f) {
    if (c.h_fin[f] < 0) continue;
    face_fi[f] = (int)fi_face.size();
    fi_face.push_back(f);
    fi_p.push_back(c.h_fip[f]);
    fi_n.push_back(c.h_fin[f]);
    nu[c.h_fip[f]]++;
    nl[c.h_fin[f]]++;
  }
  const int nfi = (int)fi_face.size();
  int WL = 0, WU = 0;
  for (int i = 0; i < n; ++i) { WL = std::max(WL, nl[i]); WU = std::max(WU, nu[i]); }
  WL = std::max(WL, 1);
  WU = std::max(WU, 1);
  std::vector<int> lc((size_t)WL * n), uc((size_t)WU * n), lfi((size_t)WL * n, -1), ufi((size_t)WU * n, -1);
  for (int i = 0; i < n; ++i) {
    for (int k = 0; k < WL; ++k) lc[(size_t)k * n + i] = i;
    for (int k = 0; k < WU; ++k) uc[(size_t)k * n + i] = i;
  }
  std::fill(nl.begin(), nl.end(), 0);
  std::fill(nu.begin(), nu.end(), 0);
  for (int q = 0; q < nfi; ++q) {
    const int p = fi_p[q], m = fi_n[q];
    const size_t ou = (size_t)(nu[p]++) * n + p, ol = (size_t)(nl[m]++) * n + m;
    uc[ou] = m; ufi[ou] = q;
    lc[ol] = p; lfi[ol] = q;
  }
  // A's columns, self-padded
  std::vector<int> ac(spc);
  for (int i = 0; i < n; ++i)
    for (int k = 0; k < Ws; ++k)
      if (ac[(size_t)k * n + i] < 0) ac[(size_t)k * n + i] = i;
  long span = 0;
  for (size_t o = 0; o < ac.size(); ++o) span = std::max(span, std::labs((long)ac[o] - (long)(o % n)));
  X.i16 = span <= 32767;
  auto up_cols = [&](const std::vector<int> &v, void **dst) -> int {
    if (X.i16) {
      std::vector<short> s(v.size());
      for (size_t o = 0; o < v.size(); ++o) s[o] = (short)(v[o] - (int)(o % n));
      short *d;
      int rc = dalloc(c, &d, s.size());
      if (rc) return rc;
      FVK_HIP(hipMemcpy(d, s.data(), s.size() * sizeof(short), hipMemcpyHostToDevice));
      *dst = d;
    } else {
      int *d;
      int rc = dalloc(c, &d, v.size());
      if (rc) return rc;
      FVK_HIP(hipMemcpy(d, v.data(), v.size() * sizeof(int), hipMemcpyHostToDevice));
      *dst = d;
    }
    return 0;
  };
  auto up_int = [&](const std::vector<int> &v, int **dst) -> int {
    int rc = dalloc(c, dst, v.size());
    if (rc) return rc;
    if (!v.empty()) FVK_HIP(hipMemcpy(*dst, v.data(), v.size() * sizeof(int), hipMemcpyHostToDevice));
    return 0;
  };
  int rc;
  if ((rc = up_cols(ac, &X.col))) return rc;
  if ((rc = up_cols(lc, &X.lcol))) return rc;
  if ((rc = up_cols(uc, &X.ucol))) return rc;
  if ((rc = up_int(lfi, &X.lfi))) return rc;
  if ((rc = up_int(ufi, &X.ufi))) return rc;
  if ((rc = up_int(fi_face, &X.fi_face))) return rc;
  if ((rc = up_int(face_fi, &X.face_fi))) return rc;
  if ((rc = up_int(fi_p, &X.fi_p))) return rc;
  if ((rc = up_int(fi_n, &X.fi_n))) return rc;
  if ((rc = dalloc(c, &X.spv, (size_t)Ws * n))) return rc;
  lval_t *lv, *uv;
  if ((rc = dalloc(c, &lv, (size_t)WL * n))) return rc;
  if ((rc = dalloc(c, &uv, (size_t)WU * n))) return rc;
  X.lval = lv;
  X.uval = uv;
  if ((rc = dalloc(c, &X.w, n))) return rc;
  if ((rc = dalloc(c, &X.e, n))) return rc;
  if ((rc = dalloc(c, &X.fl, std::max(nfi, 1)))) return rc;
  if ((rc = dalloc(c, &X.fsig, std::max(nfi, 1)))) return rc;
  if ((rc = dalloc(c, &X.vec, (size_t)10 * n))) return rc;
  if ((rc = dalloc(c, &X.tot, 8))) return rc;
  X.WL = WL;
  X.WU = WU;
  X.nfi = nfi;
  X.built = true;
  return 0;
}

int factor_assemble(Ctx &c) {
  Mix &X = c.mx;
  k_fac_rows<<<grid_for(c.n), BS, 0, c.stream>>>(c.n, X.WL, X.WU, X.lfi, X.ufi, X.fl, X.fsig, (lval_t *)X.lval,
                                                  (lval_t *)X.uval, X.e);
  FVK_HIP(hipGetLastError());
  X.fac_valid = true;
  return 0;
}

// Work vectors (FP32, n each)
enum { V_E = 0, V_R0, V_R1, V_Z, V_P0, V_P1, V_Q, V_T, V_A, V_B };
static inline float *fv(Ctx &c, int k) { return c.mx.vec + (size_t)k * c.n; }

template <class I>
static int inner_solve(Ctx &c, bool fac, int n_iter, double reduc, int &iters, int &broken, KTimer *T) {
  Mix &X = c.mx;
  const int n = c.n, G = red_grid(n), ws = c.Ws, wl = X.WL, wu = X.WU;
  const I *col = (const I *)X.col, *lcol = (const I *)X.lcol, *ucol = (const I *)X.ucol;
  const lval_t *lval = (const lval_t *)X.lval, *uval = (const lval_t *)X.uval;
  // scalars: reduc, n_iter, s = large
  KS h[6];
  std::memset(h, 0, sizeof h);
  for (int s = 0; s < 2; ++s) {
    h[s * 3].reduc = reduc;
    h[s * 3].n_iter = n_iter;
    h[s * 3].s = 1.0 / SMALL;
  }
  FVK_HIP(hipMemcpyAsync(c.ksr, h, sizeof h, hipMemcpyHostToDevice, c.stream));
  float *r[2] = {fv(c, V_R0), fv(c, V_R1)}, *P[2] = {fv(c, V_P0), fv(c, V_P1)};
  // Jacobi: z is r itself, updated in place
  float *z = fac ? fv(c, V_Z) : r[0], *q = fv(c, V_Q), *t = fv(c, V_T), *ev = fv(c, V_E);
#define L(W) \
  k_mx_start<W, I><<<G, RB, 0, c.stream>>>(n, fac ? 1 : 0, r[0], ev, P[0], wu, ucol, uval, X.e, t, pb(c, 0))
  FVK_W(wu, L);
#undef L
  const double *pR = nullptr;
  if (fac) {
#define L(W) k_mx_gz<W, I><<<G, RB, 0, c.stream>>>(n, wl, lcol, lval, t, r[0], z, pb(c, 2), nullptr)
    FVK_W(wl, L);
#undef L
    pR = pb(c, 2);
  }
  const int chunk = 32;
  int it = 0;
  for (; it < n_iter; ++it) {
    float *pold = P[it & 1], *pnew = P[(it + 1) & 1], *rin = r[it & 1], *rout = r[(it + 1) & 1];
    hipEvent_t ea, eb;
    if (T && T->take(ea, eb)) {
      const float *zc = z, *pc = pold, *sp = X.spv;
#define L(W)                                                                                                  \
  hipExtLaunchKernelGGL(k_mx_pq<W, I>, dim3(G), dim3(RB), 0, c.stream, ea, eb, 0, n, ws, it, col, sp, zc, pc, pnew, \
                        q, (const double *)pb(c, 0), G, pR, G, pb(c, 1), c.ksr, c.ks)
      FVK_W(ws, L);
#undef L
    } else {
#define L(W)                                                                                                  \
  k_mx_pq<W, I><<<G, RB, 0, c.stream>>>(n, ws, it, col, X.spv, z, pold, pnew, q, pb(c, 0), G, pR, G, pb(c, 1), \
                                        c.ksr, c.ks)
      FVK_W(ws, L);
#undef L
    }
    if (!fac) {
      k_mx_upd_jac<<<G, RB, 0, c.stream>>>(n, it, ev, r[0], pnew, q, pb(c, 1), G, pb(c, 0), c.ksr);
    } else {
#define L(W)                                                                                                   \
  k_mx_upd_fac<W, I><<<G, RB, 0, c.stream>>>(n, it, ev, rin, rout, pnew, q, wu, ucol, uval, X.e, t, pb(c, 1), G, \
                                             pb(c, 0), c.ksr)
      FVK_W(wu, L);
#undef L
#define L(W) k_mx_gz<W, I><<<G, RB, 0, c.stream>>>(n, wl, lcol, lval, t, rout, z, pb(c, 2), c.ksr + (it & 1) * 3)
      FVK_W(wl, L);
#undef L
    }
    if ((it + 1) % chunk == 0 && it + 1 < n_iter) {
      KS k;
      if (T && T->flush()) return FVK_ERR_HIP;
      FVK_HIP(hipMemcpyAsync(&k, c.ks, sizeof k, hipMemcpyDeviceToHost, c.stream));
      FVK_HIP(hipStreamSynchronize(c.stream));
      if (k.done) { ++it; break; }
    }
  }
  k_mx_final<<<1, RB, 0, c.stream>>>(it, pb(c, 0), G, pR, G, c.ksr, c.ks);
  FVK_HIP(hipGetLastError());
  if (T && T->flush()) return FVK_ERR_HIP;
  KS k, kr[2];
  FVK_HIP(hipMemcpyAsync(&k, c.ks, sizeof k, hipMemcpyDeviceToHost, c.stream));
  FVK_HIP(hipMemcpyAsync(&kr[0], c.ksr, sizeof(KS), hipMemcpyDeviceToHost, c.stream));
  FVK_HIP(hipMemcpyAsync(&kr[1], c.ksr + 3, sizeof(KS), hipMemcpyDeviceToHost, c.stream));
  FVK_HIP(hipStreamSynchronize(c.stream));
  iters = k.iter;
  broken = kr[0].pad || kr[1].pad;
  // the solution lives in r[..] only through e; the last residual buffer is
  // not needed after the FP64 residual is recomputed
  return 0;
}

int dc_solve(Ctx &c, bool use_factor, const double *b, double *x, double tol, int max_iter, DcResult &res,
             KTimer *T) {
  int rc = mix_build(c);
  if (rc) return rc;
  Mix &X = c.mx;
  if (use_factor && !X.fac_valid) {
    set_error("no preconditioner factor assembled");
    return FVK_ERR_STATE;
  }
  const int n = c.n, G = red_grid(n);
  k_mx_conv<<<grid_for(n), BS, 0, c.stream>>>(c.M, c.diag, c.spv, X.spv, X.w);
  FVK_HIP(hipMemsetAsync(x, 0, sizeof(double) * n, c.stream));
  res = DcResult();
  for (int outer = 0;; ++outer) {
#define L(W) k_mx_res<W><<<G, RB, 0, c.stream>>>(c.M, c.diag, c.spv, x, b, fv(c, V_R0), pb(c, 3))
    FVK_W(c.Ws, L);
#undef L
    k_mx_tot4<<<1, RB, 0, c.stream>>>(pb(c, 3), G, X.tot);
    double h[4];
    FVK_HIP(hipMemcpyAsync(h, X.tot, sizeof h, hipMemcpyDeviceToHost, c.stream));
    FVK_HIP(hipStreamSynchronize(c.stream));
    res.b_l1 = h[2];
    res.res_l1 = h[2] > 0 ? h[0] / h[2] : h[0];
    res.res_l2 = h[3] > 0 ? std::sqrt(h[1] / h[3]) : std::sqrt(h[1]);
    res.outer = outer;
    if (!(res.res_l1 == res.res_l1)) { res.broken = 1; break; }
    if (res.res_l1 < tol) { res.converged = true; break; }
    if (outer >= X.max_outer || res.iters >= max_iter || res.broken) break;
    // inner target: what is left to reach tol (with a margin), but no
    // deeper than inner_reduc, where FP32 rounding takes over
    const double eta = std::min(0.5, std::max(X.inner_reduc, 0.3 * tol / res.res_l1));
    // The inner right-hand side is normalised to a mean |r| of 1: the PCG
    // recurrence's alpha = s/(p.q + small) (solv_cg.f90:69-70) carries an
    // absolute small = 1e-20 that stalls the solve once p.q is that tiny,
    // which the residuals of later correction steps reach.
    const double sigma = h[0] > 0 ? h[0] / n : 1.0;
    k_scale32<<<grid_for(n), BS, 0, c.stream>>>(n, fv(c, V_R0), (float)(1.0 / sigma));
    int its = 0, broken = 0;
    rc = X.i16 ? inner_solve<short>(c, use_factor, max_iter - res.iters, eta, its, broken, T)
               : inner_solve<int>(c, use_factor, max_iter - res.iters, eta, its, broken, T);
    if (rc) return rc;
    res.iters += its;
    res.broken |= broken;
    k_mx_add<<<grid_for(n), BS, 0, c.stream>>>(n, x, fv(c, V_E), c.diag, sigma);
    FVK_HIP(hipGetLastError());
  }
  return 0;
}

// z = G G^T r = D^-1/2 (I + L) Delta^2 (I + L^T) D^-1/2 r on device rows
// (FP32 inside), for tests
__global__ __launch_bounds__(BS) void k_scale_f32(int n, const double *a, const double *diag, float *b) {
  const int i = blockIdx.x * BS + threadIdx.x;
  if (i < n) b[i] = (float)(a[i] / sqrt(diag[i]));
}
__global__ __launch_bounds__(BS) void k_scale_f64(int n, const float *a, const double *diag, double *b) {
  const int i = blockIdx.x * BS + threadIdx.x;
  if (i < n) b[i] = (double)a[i] / sqrt(diag[i]);
}

template <class I>
static int apply_t(Ctx &c, const double *r, double *zout) {
  Mix &X = c.mx;
  const int n = c.n, G = red_grid(n), wl = X.WL, wu = X.WU;
  const I *lcol = (const I *)X.lcol, *ucol = (const I *)X.ucol;
  const lval_t *lval = (const lval_t *)X.lval, *uval = (const lval_t *)X.uval;
  float *r32 = fv(c, V_A), *z = fv(c, V_B), *t = fv(c, V_T), *ev = fv(c, V_E), *p0 = fv(c, V_P0);
  k_scale_f32<<<grid_for(n), BS, 0, c.stream>>>(n, r, c.diag, r32);
#define L(W) k_mx_start<W, I><<<G, RB, 0, c.stream>>>(n, 1, r32, ev, p0, wu, ucol, uval, X.e, t, pb(c, 0))
  FVK_W(wu, L);
#undef L
#define L(W) k_mx_gz<W, I><<<G, RB, 0, c.stream>>>(n, wl, lcol, lval, t, r32, z, pb(c, 2), nullptr)
  FVK_W(wl, L);
#undef L
  k_scale_f64<<<grid_for(n), BS, 0, c.stream>>>(n, z, c.diag, zout);
  FVK_HIP(hipGetLastError());
  return 0;
}

int factor_apply(Ctx &c, const double *r, double *z) {
  if (!c.mx.fac_valid) { set_error("no preconditioner factor assembled"); return FVK_ERR_STATE; }
  return c.mx.i16 ? apply_t<short>(c, r, z) : apply_t<int>(c, r, z);
}

}  // namespace fvk
