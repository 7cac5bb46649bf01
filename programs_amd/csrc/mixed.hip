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
//             eta = max(inner_reduc, eta_margin tol / (|r|_1 / |b|_1))
//           x += D^-1/2 eh (FP64)
//
// With the context setting dc_carry, a correction step after the first
// continues the Krylov recurrence of the one before (residual replacement,
// van der Vorst and Ye 2000): its first direction is p = z + beta p_last with
// the last direction and r.z of the previous inner solve rescaled to the new
// normalisation, instead of p = z.  It saves iterations when the inner
// solves stop early (inner_reduc 1e-3), but after deep FP32 inner solves the
// replaced residual is too far from the recursive one and the continued
// search can stall, so restarts are the default.
//
// The inner system is the Jacobi-scaled one, Ah = D^-1/2 A D^-1/2 (unit
// diagonal, FP32 off-diagonals, 16-bit column offsets where they fit): its
// SpMV reads no diagonal, Jacobi PCG on A is plain CG on Ah (z = r), and the
// factor becomes Gh = (I + L) Delta with zh = Gh Gh^T rh (L in FP16).
// Per inner iteration two kernels with Jacobi (z = r):
//   k_mx_pq      decide beta; p = r + beta p; q = Ah p; partial p.q
//   k_mx_upd_jac alpha; e += alpha p; r -= alpha q; partials |r|_1, r.r
// and three with the factor:
//   k_mx_zp      decide beta; e += alpha p (the last update's); z = t + L t;
//                p = z + beta p (in place)
//   k_mx_q       q = Ah p; partial p.q
//   k_mx_upd_fac alpha; r' = r - alpha q; u = r' + L^T r' (r' of the higher
//                neighbours formed on the fly from r and q, r
//                double-buffered); t = Delta^2 u; partials |r'|_1 and
//                r'.z = u^T Delta^2 u (z = (I + L) t never needs storing:
//                r^T (I + L) Delta^2 (I + L^T) r is that sum).
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "cgcore.h"

namespace fvk {

typedef _Float16 lval_t;
// t = Delta^2 (I + L^T) r as the update kernel hands it to the direction
// kernel: FP16, divided by a scale both kernels derive from the same
// committed scalars: the decision's r.z = sum Delta^2 u^2, i.e. that of the
// residual before the update, 64 sqrt(r.z / n).  With r'.z' the new
// residual's, |t| / scale <= Delta_max sqrt(n) / 64 * sqrt(r'.z' / r.z)
// whatever the residual's shape; PCG does not make r.z decrease
// monotonically, so a step where r.z grows enough can overflow FP16.
// The first direction of a solve reads the start's t in FP32.
typedef _Float16 t16_t;

// Column of slot k of row i: 32-bit rows, or 16-bit offsets from the row.
__device__ __forceinline__ int colat(const int *c, size_t k, int i) { return c[k]; }
__device__ __forceinline__ int colat(const short *c, size_t k, int i) { return i + (int)c[k]; }

template <int W>
__device__ __forceinline__ int nw(int w) { return W > 0 ? W : w; }

__device__ __forceinline__ bool gate_done(const KS *g) { return g && g->done; }

// Stride of the FP32 work vectors and of the slot-major tables of the
// inner system (Ah's columns and values, L's columns and values): n plus
// 64 KiB + 256 B, so that the same row of different vectors or slots does
// not sit at the same offset of its memory page.  With strides a power of
// two apart the streams of one kernel collided in the same HBM channels
// (5 % slower on the 2^21-row system for the vectors alone).
__host__ __device__ __forceinline__ size_t mx_ld(int n) { return (size_t)n + 16448; }

// Scale of the t stored after a decision with scalars k (1 when r.z is not
// usable); FP32 arithmetic, the same in both kernels (an FP64 square root
// and division here cost the update kernel 15 %)
__device__ __forceinline__ float t_scale(const KS &k, int n) {
  const float s = 64.f * sqrtf((float)k.s / (float)n);
  return s > 0.f && s < 1e30f ? s : 1.f;
}
// FP16, rounded to nearest.  Without growth of r.z the scale keeps |t| /
// scale below Delta_max sqrt(n) / 64 < 65504 for n < 2.8e8 rows even when
// all of r.z sits in one row.  An overflow makes the next p.q non-finite,
// which the update kernel's decision takes as a breakdown before applying
// the step: the inner solve ends with a finite correction, which is kept,
// and the defect correction restarts from that x (after repeated
// breakdowns the caller's fallback takes over).  (Rounding toward zero, which
// saturates, cost 13 % more iterations; a clamp before the conversion cost
// the update kernel registers and 15 % of its time.)
__device__ __forceinline__ t16_t to_t16(float v) { return (t16_t)v; }

// Jacobi-scaled FP32 copy of the matrix (the Jacobi path's SpMV)
__global__ __launch_bounds__(BS) void k_mx_conv(DMesh M, const double *__restrict__ diag,
                                                const double *__restrict__ spv, float *__restrict__ s32) {
  const int i = blockIdx.x * BS + threadIdx.x;
  const int n = M.n;
  if (i >= n) return;
  const double wi = 1.0 / sqrt(diag[i]);
  for (int k = 0; k < M.Ws; ++k) {
    const size_t o = (size_t)k * M.ld + i;
    const int j = M.spc[o];
    s32[(size_t)k * mx_ld(n) + i] = j >= 0 ? (float)(spv[o] * wi / sqrt(diag[j])) : 0.f;
  }
}

// FP64 true residual r = b - A x, its scaled FP32 copy scale D^-1/2 r, partials
// {|r|_1, |r|_2^2, |b|_1, |b|_2^2}.  gate (residual replacement inside an
// inner solve): nothing once that solve is done or broke down.
template <int WS>
__global__ __launch_bounds__(RB) void k_mx_res(DMesh M, const double *__restrict__ diag,
                                               const double *__restrict__ spv, const double *__restrict__ x,
                                               const double *__restrict__ b, float *__restrict__ r32,
                                               double *__restrict__ part, const KS *gate, double scale) {
  if (gate && (gate->done || gate->pad)) return;
  double a[4] = {0, 0, 0, 0};
  int r0, r1;
  block_rows(M.n, r0, r1);
  for (int i = r0 + threadIdx.x; i < r1; i += RB) {
    const double bi = b[i], ri = bi - row_spmv<WS>(M, diag, spv, x, i);
    r32[i] = (float)(ri / sqrt(diag[i]) * scale);
    a[0] += fabs(ri);
    a[1] += ri * ri;
    a[2] += fabs(bi);
    a[3] += bi * bi;
  }
  store_partials<4>(a, part);
}

// Start of an inner solve: e = 0, p_old = f p_last (p_last null: 0),
// partials |r|_1 [0] and Jacobi: r.r [3]; factor: t = Delta^2 (r + L^T r).
// p_last may be p0 itself (each row reads and writes its own element).
// After a residual replacement (ev and p0 null: both kept) gate is the last
// iteration's scalars: nothing once the solve is done or broke down, and the
// last update's alpha, already applied to x, is cleared for the next
// direction kernel.
template <int WU, class I, class TT>
__global__ __launch_bounds__(RB) void k_mx_start(int n, int fac, const float *__restrict__ r,
                                                 float *__restrict__ ev, float *p0, const float *p_last, float f,
                                                 int wu, const I *__restrict__ ucol,
                                                 const lval_t *__restrict__ uval, const lval_t *__restrict__ ee,
                                                 TT *__restrict__ t, double *__restrict__ part, KS *gate) {
  if (gate && (gate->done || gate->pad)) return;
  double a[4] = {0, 0, 0, 0};
  int r0, r1;
  block_rows(n, r0, r1);
  for (int i = r0 + threadIdx.x; i < r1; i += RB) {
    const float ri = r[i];
    if (ev) ev[i] = 0.f;
    if (p0) p0[i] = p_last ? f * p_last[i] : 0.f;
    a[0] += fabs((double)ri);
    if (!fac) {
      a[3] += (double)ri * (double)ri;
    } else {
      float acc = ri;
#pragma unroll
      for (int k = 0; k < nw<WU>(wu); ++k) {
        const size_t o = (size_t)k * mx_ld(n) + i;
        acc += (float)uval[o] * r[colat(ucol, o, i)];
      }
      const float ei = (float)ee[i];
      t[i] = (TT)(ei * acc);
      a[3] += (double)ei * (double)acc * (double)acc;
    }
  }
  store_partials<4>(a, part);
  if (gate && blockIdx.x == 0 && threadIdx.x == 0) gate->alpha = 0.0;
}

// Decide iteration it of the inner PCG; p = z + beta p ; q = Ah p ; p.q.
// A row's loads do not depend on beta: the first row of every thread is
// loaded before the partials of the previous kernel are reduced, so the
// reduction's latency hides behind it.
template <int WS, class I>
struct PqRow {
  float zi, pi_old, a[WS > 0 ? WS : 1], zj[WS > 0 ? WS : 1], pj[WS > 0 ? WS : 1];
  __device__ __forceinline__ void load(int n, int i, const I *__restrict__ col, const float *__restrict__ spv,
                                       const float *__restrict__ z, const float *__restrict__ po) {
    zi = z[i];
    pi_old = po[i];
#pragma unroll
    for (int k = 0; k < WS; ++k) {
      const size_t o = (size_t)k * mx_ld(n) + i;
      const int j = colat(col, o, i);
      a[k] = spv[o];
      zj[k] = z[j];
      pj[k] = po[j];
    }
  }
  __device__ __forceinline__ void finish(int i, float bf, float *__restrict__ pn, float *__restrict__ q,
                                         double &pq) const {
    const float p = zi + bf * pi_old;
    float y = p;
#pragma unroll
    for (int k = 0; k < WS; ++k) y += a[k] * (zj[k] + bf * pj[k]);
    pn[i] = p;
    q[i] = y;
    pq += (double)p * (double)y;
  }
};

template <int WS, class I>
__device__ __forceinline__ void mx_pq_row(int n, int i, const I *__restrict__ col, const float *__restrict__ spv,
                                          const float *__restrict__ z, const float *__restrict__ po, float bf,
                                          float *__restrict__ pn, float *__restrict__ q, int ws, double &pq) {
  const float pi = z[i] + bf * po[i];
  float y = pi;
#pragma unroll
  for (int k = 0; k < nw<WS>(ws); ++k) {
    const size_t o = (size_t)k * mx_ld(n) + i;
    const int j = colat(col, o, i);
    y += spv[o] * (z[j] + bf * po[j]);
  }
  pn[i] = pi;
  q[i] = y;
  pq += (double)pi * (double)y;
}

// The decision of iteration it from the partials |r|_1 [0] and r.z [3] of
// the last update (issued with mx_issue; thread 0 decides for the block).
__device__ __forceinline__ void mx_issue(SelLoads<2> &S, const double *partU, int PU) {
  const int cs[2] = {0, 3};
  S.issue(partU, PU, cs);
}
// take_alpha: the caller applies the last update's e += alpha p itself, so
// sk[0].alpha keeps that alpha for the block while the committed scalars say
// 0 (unless the solve is done, when the host applies it).
__device__ __forceinline__ void mx_decide(int it, SelLoads<2> &S, KS *ring, KS *out, KS *sk,
                                          bool take_alpha = false) {
  double t[2];
  S.finish(t);
  if (threadIdx.x == 0) {
    const KS *prev = ring + ((it + 1) & 1) * 3;
    cg_decide(it, 1, t, t + 1, prev, sk);
    if (prev->pad) sk[0].done = 1;  // breakdown flagged by the last update
    if (blockIdx.x == 0) {
      KS k = sk[0];
      if (take_alpha && !k.done) k.alpha = 0.0;
      ring[(it & 1) * 3] = out[0] = k;
    }
  }
  __syncthreads();
}

template <int WS, class I>
__global__ __launch_bounds__(RB, 8) void k_mx_pq(int n, int ws, int it, const I *__restrict__ col,
                                              const float *__restrict__ spv, const float *__restrict__ z,
                                              const float *__restrict__ pold, float *__restrict__ pnew,
                                              float *__restrict__ q, const double *__restrict__ partU, int PU,
                                              double *__restrict__ partQ, KS *ring, KS *out) {
  int r0, r1;
  block_rows(n, r0, r1);
  const int i0 = r0 + threadIdx.x;
  SelLoads<2> S;
  mx_issue(S, partU, PU);
  PqRow<WS, I> A0;
  if (WS > 0 && i0 < r1) A0.load(n, i0, col, spv, z, pold);
  __shared__ KS sk[1];
  mx_decide(it, S, ring, out, sk);
  if (sk[0].done) return;
  const float bf = (float)sk[0].beta;
  double pq = 0.0;
  int i = i0;
  if (WS > 0) {
    if (i0 < r1) A0.finish(i0, bf, pnew, q, pq);
    i = i0 + RB;
  }
  for (; i + RB < r1; i += 2 * RB) {
    mx_pq_row<WS, I>(n, i, col, spv, z, pold, bf, pnew, q, ws, pq);
    mx_pq_row<WS, I>(n, i + RB, col, spv, z, pold, bf, pnew, q, ws, pq);
  }
  if (i < r1) mx_pq_row<WS, I>(n, i, col, spv, z, pold, bf, pnew, q, ws, pq);
  double v[1] = {pq};
  store_partials<1>(v, partQ);
}

// Factor, iteration it: decide beta; e += alpha p with the last update's
// alpha and direction; z = t + L t; p = z + beta p in place.  The first row
// of a thread is loaded before the decision is known (more rows ahead
// measured slower).
template <int WL, class I, class TT>
__device__ __forceinline__ float mx_z_row(int n, int i, int wl, const I *__restrict__ lcol,
                                          const lval_t *__restrict__ lval, const TT *__restrict__ t) {
  float acc = (float)t[i];
#pragma unroll
  for (int k = 0; k < nw<WL>(wl); ++k) {
    const size_t o = (size_t)k * mx_ld(n) + i;
    acc += (float)lval[o] * (float)t[colat(lcol, o, i)];
  }
  return acc;
}

template <int WL, class I, class TT>
struct ZRow {
  float ti, pi, ei, l[WL > 0 ? WL : 1], tj[WL > 0 ? WL : 1];
  __device__ __forceinline__ void load(int n, int i, const I *__restrict__ lcol, const lval_t *__restrict__ lval,
                                       const TT *__restrict__ t, const float *__restrict__ p,
                                       const float *__restrict__ ev) {
    ti = (float)t[i];
    pi = p[i];
    ei = ev[i];
#pragma unroll
    for (int k = 0; k < WL; ++k) {
      const size_t o = (size_t)k * mx_ld(n) + i;
      l[k] = (float)lval[o];
      tj[k] = (float)t[colat(lcol, o, i)];
    }
  }
  __device__ __forceinline__ float z() const {
    float acc = ti;
#pragma unroll
    for (int k = 0; k < WL; ++k) acc += l[k] * tj[k];
    return acc;
  }
};

template <int WL, class I, class TT>
__global__ __launch_bounds__(RB, 8) void k_mx_zp(int n, int wl, int it, const I *__restrict__ lcol,
                                              const lval_t *__restrict__ lval, const TT *__restrict__ t,
                                              float *__restrict__ p, float *__restrict__ ev,
                                              const double *__restrict__ partU, int PU, KS *ring, KS *out) {
  int r0, r1;
  block_rows(n, r0, r1);
  const int i0 = r0 + threadIdx.x;
  SelLoads<2> S;
  mx_issue(S, partU, PU);
  ZRow<WL, I, TT> A0;
  if (WL > 0) {
    if (i0 < r1) A0.load(n, i0, lcol, lval, t, p, ev);
  }
  __shared__ KS sk[1];
  mx_decide(it, S, ring, out, sk, true);
  if (sk[0].done) return;
  const float bf = (float)sk[0].beta, af = (float)sk[0].alpha;
  // t was stored divided by the scale of the previous decision (FP32 and
  // unscaled from the start of the solve or a residual replacement)
  const float ts = std::is_same<TT, float>::value ? 1.f : t_scale(ring[((it + 1) & 1) * 3], n);
  int i = i0;
  if (WL > 0) {
    if (i0 < r1) {
      ev[i0] = A0.ei + af * A0.pi;
      p[i0] = ts * A0.z() + bf * A0.pi;
    }
    i = i0 + RB;
  }
  auto row = [&](int i) {
    const float z = ts * mx_z_row<WL, I, TT>(n, i, wl, lcol, lval, t), pi = p[i];
    ev[i] += af * pi;
    p[i] = z + bf * pi;
  };
  for (; i + RB < r1; i += 2 * RB) {
    row(i);
    row(i + RB);
  }
  if (i < r1) row(i);
}

// Factor, iteration it: q = Ah p; partial p.q (nothing once the solve is
// done).  Ah is symmetric, so only the entries towards the upper neighbours
// are stored (au, FP32, one slot per upper neighbour); the entry towards a
// lower neighbour j is j's upper entry, read from au[iau][j] (iau: the slot
// of this row in j's upper list, one byte; padding points at a zero slot).
// Those reads go to rows just streamed, so the kernel reads half the values
// from HBM (35n instead of 44n bytes per launch on a hexahedral mesh; 5 %
// faster, as the kernel is not bound by HBM alone).  The first row of a
// thread is loaded before the done flag is read, so that its latency hides
// behind it.
template <int WX, class I>
struct QRow {
  float pi, a[WX > 0 ? WX : 1], pu[WX > 0 ? WX : 1], b[WX > 0 ? WX : 1], pl[WX > 0 ? WX : 1];
  __device__ __forceinline__ void load(int n, int i, const I *__restrict__ ucol, const I *__restrict__ lcol,
                                       const float *__restrict__ au, const unsigned char *__restrict__ iau,
                                       const float *__restrict__ p) {
    const size_t ld = mx_ld(n);
    pi = p[i];
#pragma unroll
    for (int k = 0; k < WX; ++k) {
      const size_t o = (size_t)k * ld + i;
      a[k] = au[o];
      pu[k] = p[colat(ucol, o, i)];
      const int j = colat(lcol, o, i);
      b[k] = au[(size_t)iau[o] * ld + j];
      pl[k] = p[j];
    }
  }
  __device__ __forceinline__ float y() const {
    float acc = pi;
#pragma unroll
    for (int k = 0; k < WX; ++k) acc += a[k] * pu[k];
#pragma unroll
    for (int k = 0; k < WX; ++k) acc += b[k] * pl[k];
    return acc;
  }
};

template <int WX, class I>
__global__ __launch_bounds__(RB, 8) void k_mx_q(int n, int wx, int it, const I *__restrict__ ucol,
                                                const I *__restrict__ lcol, const float *__restrict__ au,
                                                const unsigned char *__restrict__ iau, const float *__restrict__ p,
                                                float *__restrict__ q, double *__restrict__ partQ, const KS *ring) {
  int r0, r1;
  block_rows(n, r0, r1);
  const int i0 = r0 + threadIdx.x;
  constexpr int NA = WX == 0 ? 0 : 1;
  QRow<WX, I> A[NA > 0 ? NA : 1];
#pragma unroll
  for (int a = 0; a < NA; ++a)
    if (i0 + a * RB < r1) A[a].load(n, i0 + a * RB, ucol, lcol, au, iau, p);
  if (ring[(it & 1) * 3].done) return;
  double pq = 0.0;
#pragma unroll
  for (int a = 0; a < NA; ++a) {
    const int i = i0 + a * RB;
    if (i < r1) {
      const float y = A[a].y();
      q[i] = y;
      pq += (double)A[a].pi * (double)y;
    }
  }
  const size_t ld = mx_ld(n);
  auto row = [&](int i) {
    const float pi = p[i];
    float y = pi;
#pragma unroll
    for (int k = 0; k < nw<WX>(wx); ++k) {
      const size_t o = (size_t)k * ld + i;
      y += au[o] * p[colat(ucol, o, i)];
    }
#pragma unroll
    for (int k = 0; k < nw<WX>(wx); ++k) {
      const size_t o = (size_t)k * ld + i;
      const int j = colat(lcol, o, i);
      y += au[(size_t)iau[o] * ld + j] * p[j];
    }
    q[i] = y;
    pq += (double)pi * (double)y;
  };
  int i = i0 + NA * RB;
  for (; i + RB < r1; i += 2 * RB) {
    row(i);
    row(i + RB);
  }
  if (i < r1) row(i);
  double v[1] = {pq};
  store_partials<1>(v, partQ);
}

// Factor path on meshes whose rows have few neighbours (triangles: three in
// all, so the upper and lower lists above are half padding): q = Ah p from
// the full FP32 rows of Ah (the Jacobi path's slots), partial p.q; nothing
// once the solve is done.
template <int WS, class I>
__global__ __launch_bounds__(RB, 8) void k_mx_qf(int n, int ws, int it, const I *__restrict__ col,
                                                 const float *__restrict__ spv, const float *__restrict__ p,
                                                 float *__restrict__ q, double *__restrict__ partQ, const KS *ring) {
  int r0, r1;
  block_rows(n, r0, r1);
  if (ring[(it & 1) * 3].done) return;
  const size_t ld = mx_ld(n);
  double pq = 0.0;
  auto row = [&](int i) {
    const float pi = p[i];
    float y = pi;
#pragma unroll
    for (int k = 0; k < nw<WS>(ws); ++k) {
      const size_t o = (size_t)k * ld + i;
      y += spv[o] * p[colat(col, o, i)];
    }
    q[i] = y;
    pq += (double)pi * (double)y;
  };
  int i = r0 + threadIdx.x;
  for (; i + RB < r1; i += 2 * RB) {
    row(i);
    row(i + RB);
  }
  if (i < r1) row(i);
  double v[1] = {pq};
  store_partials<1>(v, partQ);
}

// Upper entries of Ah: au[k][p] = a_pm / sqrt(a_pp a_mm) for the k-th upper
// neighbour m of row p (zero on padding; slot WX stays zero)
__global__ __launch_bounds__(BS) void k_mx_au(int n, int WU, const int *__restrict__ ufi,
                                              const int *__restrict__ fi_face, const int *__restrict__ fi_n,
                                              const double *__restrict__ ngp, const double *__restrict__ diag,
                                              float *__restrict__ au) {
  const int i = blockIdx.x * BS + threadIdx.x;
  if (i >= n) return;
  const double wi = 1.0 / sqrt(diag[i]);
  for (int k = 0; k < WU; ++k) {
    const int q = ufi[(size_t)k * n + i];
    au[(size_t)k * mx_ld(n) + i] = q >= 0 ? (float)(ngp[fi_face[q]] * wi / sqrt(diag[fi_n[q]])) : 0.f;
  }
}

// The partials of p.q, issued before an update kernel's row loads
__device__ __forceinline__ void mx_issue_q(SelLoads<1> &S, const double *partQ, int PQ) {
  const int cs[1] = {0};
  S.issue(partQ, PQ, cs);
}

// alpha = s/(p.q + small) from the partials; false on breakdown (a
// non-positive or non-finite p.q), flagged for the next decision.
__device__ __forceinline__ bool mx_alpha(KS *cur, SelLoads<1> &S, double &alpha) {
  double tq[1];
  S.finish(tq);
  alpha = cur->s / (tq[0] + SMALL);
  if (!(tq[0] > 0.0) || !isfinite(tq[0]) || !isfinite(alpha)) {
    if (blockIdx.x == 0 && threadIdx.x == 0) cur->pad = 1;
    return false;
  }
  return true;
}

// Jacobi: e += alpha p ; r -= alpha q (in place) ; partials |r|_1, r.r.
// The first two rows of a thread are loaded before alpha is known.
__global__ __launch_bounds__(RB) void k_mx_upd_jac(int n, int it, float *__restrict__ ev, float *__restrict__ r,
                                                   const float *__restrict__ p, const float *__restrict__ q,
                                                   const double *__restrict__ partQ, int PQ,
                                                   double *__restrict__ partU, KS *ring) {
  int r0, r1;
  block_rows(n, r0, r1);
  const int i0 = r0 + threadIdx.x, i1 = i0 + RB;
  SelLoads<1> S;
  mx_issue_q(S, partQ, PQ);
  float e0 = 0.f, e1 = 0.f, ra0 = 0.f, ra1 = 0.f, p0 = 0.f, p1 = 0.f, q0 = 0.f, q1 = 0.f;
  if (i0 < r1) { e0 = ev[i0]; ra0 = r[i0]; p0 = p[i0]; q0 = q[i0]; }
  if (i1 < r1) { e1 = ev[i1]; ra1 = r[i1]; p1 = p[i1]; q1 = q[i1]; }
  KS *cur = ring + (it & 1) * 3;
  if (cur->done) return;
  double alpha;
  if (!mx_alpha(cur, S, alpha)) return;
  const float af = (float)alpha;
  double a[4] = {0, 0, 0, 0};
  auto fin = [&](int i, float e, float ri_, float pi, float qi) {
    ev[i] = e + af * pi;
    const float ri = ri_ - af * qi;
    r[i] = ri;
    a[0] += fabs((double)ri);
    a[3] += (double)ri * (double)ri;
  };
  if (i0 < r1) fin(i0, e0, ra0, p0, q0);
  if (i1 < r1) fin(i1, e1, ra1, p1, q1);
  auto row = [&](int i) { fin(i, ev[i], r[i], p[i], q[i]); };
  int i = i0 + 2 * RB;
  for (; i + RB < r1; i += 2 * RB) {
    row(i);
    row(i + RB);
  }
  if (i < r1) row(i);
  store_partials<4>(a, partU);
  if (blockIdx.x == 0 && threadIdx.x == 0) cur->alpha = alpha;
}

// Factor: r' = r - alpha q ; u = r' + L^T r' with r' of the higher
// neighbours formed from (r, q) ; t = Delta^2 u ; partials |r'|_1 [0] and
// r'.z = u Delta^2 u [3] (e += alpha p is left to the next k_mx_zp, which
// reads p anyway).  The first rows of a thread are loaded before alpha is
// known.
template <int WU, class I>
struct UpdRow {
  float ri, qi, ee, u[WU > 0 ? WU : 1], rj[WU > 0 ? WU : 1], qj[WU > 0 ? WU : 1];
  __device__ __forceinline__ void load(int n, int i, const float *__restrict__ rin, const float *__restrict__ q,
                                       const I *__restrict__ ucol, const lval_t *__restrict__ uval,
                                       const lval_t *__restrict__ eev) {
    ri = rin[i];
    qi = q[i];
    ee = (float)eev[i];
#pragma unroll
    for (int k = 0; k < WU; ++k) {
      const size_t o = (size_t)k * mx_ld(n) + i;
      const int j = colat(ucol, o, i);
      u[k] = (float)uval[o];
      rj[k] = rin[j];
      qj[k] = q[j];
    }
  }
};

template <int WU, class I>
__global__ __launch_bounds__(RB, 8) void k_mx_upd_fac(int n, int it, const float *__restrict__ rin,
                                                      float *__restrict__ rout, const float *__restrict__ q, int wu,
                                                   const I *__restrict__ ucol, const lval_t *__restrict__ uval,
                                                   const lval_t *__restrict__ ee, t16_t *__restrict__ t,
                                                   const double *__restrict__ partQ, int PQ,
                                                   double *__restrict__ partU, KS *ring) {
  int r0, r1;
  block_rows(n, r0, r1);
  const int i0 = r0 + threadIdx.x, i1 = i0 + RB;
  SelLoads<1> S;
  mx_issue_q(S, partQ, PQ);
  // one row loaded ahead (two measured slower: register pressure)
  constexpr int NA = WU == 0 ? 0 : 1;
  UpdRow<WU, I> A0, A1;
  if (NA > 0 && i0 < r1) A0.load(n, i0, rin, q, ucol, uval, ee);
  if (NA > 1 && i1 < r1) A1.load(n, i1, rin, q, ucol, uval, ee);
  KS *cur = ring + (it & 1) * 3;
  if (cur->done) return;
  double alpha;
  if (!mx_alpha(cur, S, alpha)) return;
  const float af = (float)alpha, tsi = 1.f / t_scale(*cur, n);
  double a[4] = {0, 0, 0, 0};
  auto fin = [&](int i, const UpdRow<WU, I> &R) {
    const float ri = R.ri - af * R.qi;
    rout[i] = ri;
    a[0] += fabs((double)ri);
    float acc = ri;
#pragma unroll
    for (int k = 0; k < WU; ++k) acc += R.u[k] * (R.rj[k] - af * R.qj[k]);
    t[i] = to_t16(R.ee * acc * tsi);
    a[3] += (double)R.ee * (double)acc * (double)acc;
  };
  auto row = [&](int i) {
    const float ri = rin[i] - af * q[i];
    rout[i] = ri;
    a[0] += fabs((double)ri);
    float acc = ri;
#pragma unroll
    for (int k = 0; k < nw<WU>(wu); ++k) {
      const size_t o = (size_t)k * mx_ld(n) + i;
      const int j = colat(ucol, o, i);
      acc += (float)uval[o] * (rin[j] - af * q[j]);
    }
    const float ei = (float)ee[i];
    t[i] = to_t16(ei * acc * tsi);
    a[3] += (double)ei * (double)acc * (double)acc;
  };
  int i = i0;
  if (NA > 0 && i0 < r1) fin(i0, A0);
  if (NA > 1 && i1 < r1) fin(i1, A1);
  i = i0 + NA * RB;
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
      const size_t o = (size_t)k * mx_ld(n) + i;
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
__global__ __launch_bounds__(RB) void k_mx_final(int it, const double *partU, int PU, KS *ring, KS *out) {
  const int cs[2] = {0, 3};
  double tt[2];
  load_sel<2>(partU, PU, cs, tt);
  if (threadIdx.x == 0) {
    KS cur[1];
    const KS *prev = ring + ((it + 1) & 1) * 3;
    cg_decide(it, 1, tt, tt + 1, prev, cur);
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

// x += sigma D^-1/2 eh with eh = e + alpha p (the factor path's last update,
// not yet applied to e; p null: none)
__global__ __launch_bounds__(BS) void k_mx_add(int n, double *__restrict__ x, const float *__restrict__ ev,
                                               const float *__restrict__ p, float alpha,
                                               const double *__restrict__ diag, double sigma) {
  const int i = blockIdx.x * BS + threadIdx.x;
  if (i >= n) return;
  // no pending step: p is not read (after a breakdown it may hold inf)
  const float e = p && alpha != 0.f ? ev[i] + alpha * p[i] : ev[i];
  x[i] += sigma * (double)e / sqrt(diag[i]);
}

// Residual replacement inside an inner solve, first step: x += sigma D^-1/2
// (e + alpha p) with the last update's alpha (p null: none pending), e = 0;
// nothing once the solve is done or broke down.  The FP64 residual of the new
// x then replaces the recursive FP32 one (k_mx_res, k_mx_start), while the
// search direction and r.z carry on.
__global__ __launch_bounds__(BS) void k_mx_rep_x(int n, double *__restrict__ x, float *__restrict__ ev,
                                                 const float *__restrict__ p, const double *__restrict__ diag,
                                                 double sigma, const KS *prev) {
  if (prev->done || prev->pad) return;
  const int i = blockIdx.x * BS + threadIdx.x;
  if (i >= n) return;
  const float e = p ? ev[i] + (float)prev->alpha * p[i] : ev[i];
  x[i] += sigma * (double)e / sqrt(diag[i]);
  ev[i] = 0.f;
}

constexpr double FAC_E2_MIN = -9.7, FAC_E2_MAX = 11.0;

// Factor from per-interior-face (l, sig): e = exp(2 s) with s the sum of sig
// over the faces the row owns, slot values of L (FP16).
__global__ __launch_bounds__(BS) void k_fac_rows(int n, int WL, int WU, const int *__restrict__ lfi,
                                                 const int *__restrict__ ufi, const float *__restrict__ fl,
                                                 const float *__restrict__ fsig, lval_t *__restrict__ lval,
                                                 lval_t *__restrict__ uval, lval_t *__restrict__ ee) {
  const int i = blockIdx.x * BS + threadIdx.x;
  if (i >= n) return;
  double s = 0.0;
  for (int k = 0; k < WU; ++k) {
    const int f = ufi[(size_t)k * n + i];
    uval[(size_t)k * mx_ld(n) + i] = (lval_t)(f >= 0 ? fl[f] : 0.f);
    if (f >= 0) s += (double)fsig[f];
  }
  for (int k = 0; k < WL; ++k) {
    const int f = lfi[(size_t)k * n + i];
    lval[(size_t)k * mx_ld(n) + i] = (lval_t)(f >= 0 ? fl[f] : 0.f);
  }
  // Delta^2 rounded once to FP16 from FP64, the exponent held within FP16's
  // normal range
  ee[i] = (lval_t)exp(fmin(fmax(2.0 * s, FAC_E2_MIN), FAC_E2_MAX));
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
  std::vector<int> spc((size_t)Ws * c.M.ld);
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
  // Factor columns at the padded stride (self on padding), face ids at n.
  // Both lists get the same width WX = max(WL, WU) so that one kernel
  // template covers the symmetric SpMV; iau: for a lower slot of row m
  // (face owner p), the slot of m in p's upper list (WX, a zero slot of
  // au, on padding).
  const size_t ld = mx_ld(n);
  const int WX = std::max(WL, WU);
  WL = WU = WX;
  std::vector<int> lc((size_t)WL * ld), uc((size_t)WU * ld), lfi((size_t)WL * n, -1), ufi((size_t)WU * n, -1);
  std::vector<unsigned char> iau((size_t)WL * ld, (unsigned char)WX);
  if (WX > 254) { set_error("too many neighbours per row for the inner solve"); return FVK_ERR_ARG; }
  for (size_t o = 0; o < lc.size(); ++o) lc[o] = (int)(o % ld);
  for (size_t o = 0; o < uc.size(); ++o) uc[o] = (int)(o % ld);
  std::fill(nl.begin(), nl.end(), 0);
  std::fill(nu.begin(), nu.end(), 0);
  for (int q = 0; q < nfi; ++q) {
    const int p = fi_p[q], m = fi_n[q];
    const int ku = nu[p]++, kl = nl[m]++;
    uc[(size_t)ku * ld + p] = m; ufi[(size_t)ku * n + p] = q;
    lc[(size_t)kl * ld + m] = p; lfi[(size_t)kl * n + m] = q;
    iau[(size_t)kl * ld + m] = (unsigned char)ku;
  }
  // A's columns at the padded stride, self on padding
  std::vector<int> ac((size_t)Ws * ld);
  for (size_t o = 0; o < ac.size(); ++o) ac[o] = (int)(o % ld);
  for (int k = 0; k < Ws; ++k)
    for (int i = 0; i < n; ++i) {
      const int j = spc[(size_t)k * c.M.ld + i];
      if (j >= 0) ac[(size_t)k * ld + i] = j;
    }
  long span = 0;
  for (size_t o = 0; o < ac.size(); ++o) span = std::max(span, std::labs((long)ac[o] - (long)(o % ld)));
  X.i16 = span <= 32767;
  {
    // bytes per row of the factor path's SpMV: the full rows (value and
    // column per slot) against the upper entries plus the gathered lower
    // ones (value, two columns and a slot byte per padded slot pair)
    const int cb = X.i16 ? 2 : 4;
    X.qfull = Ws * (4 + cb) < WX * (5 + 2 * cb);
  }
  auto up_cols = [&](const std::vector<int> &v, void **dst) -> int {
    if (X.i16) {
      std::vector<short> s(v.size());
      for (size_t o = 0; o < v.size(); ++o) s[o] = (short)(v[o] - (int)(o % ld));
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
  if ((rc = dalloc(c, &X.spv, (size_t)Ws * ld))) return rc;
  if ((rc = dalloc(c, &X.au, (size_t)(WX + 1) * ld))) return rc;
  unsigned char *iau_d;
  if ((rc = dalloc(c, &iau_d, iau.size()))) return rc;
  FVK_HIP(hipMemcpy(iau_d, iau.data(), iau.size(), hipMemcpyHostToDevice));
  X.iau = iau_d;
  lval_t *lv, *uv;
  if ((rc = dalloc(c, &lv, (size_t)WL * ld))) return rc;
  if ((rc = dalloc(c, &uv, (size_t)WU * ld))) return rc;
  X.lval = lv;
  X.uval = uv;
  lval_t *e16;
  if ((rc = dalloc(c, &e16, n))) return rc;
  X.e = e16;
  if ((rc = dalloc(c, &X.fl, std::max(nfi, 1)))) return rc;
  if ((rc = dalloc(c, &X.fsig, std::max(nfi, 1)))) return rc;
  if ((rc = dalloc(c, &X.vec, (size_t)10 * mx_ld(n)))) return rc;
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
                                                  (lval_t *)X.uval, (lval_t *)X.e);
  FVK_HIP(hipGetLastError());
  X.fac_valid = true;
  return 0;
}

// Work vectors (FP32, n each)
enum { V_E = 0, V_R0, V_R1, V_Z, V_P0, V_P1, V_Q, V_T, V_A, V_B };
static inline float *fv(Ctx &c, int k) { return c.mx.vec + (size_t)k * mx_ld(c.n); }

// State carried from one inner solve to the next: the buffer holding the
// last search direction and the r.z it was formed with.
struct InnerCarry {
  int p_buf = -1;  // index into the direction pair, -1: none (restart)
  double s = 0.0;
};

// One inner PCG solve of Ah eh = rh (rh in V_R0, rh = D^-1/2 (b - A x) / sigma).
// carry in: the previous solve's last direction and r.z, to be scaled by f
// (rows) and f^2 (r.z); carry out: this solve's.
// alpha_pend: the factor path's last alpha, whose e += alpha p (p in P[0])
// is still to be applied.
// Every replace_every iterations (0: never) the FP32 correction so far goes
// into the FP64 solution x and the recursive residual is replaced by the
// FP64 residual of that x (reliable updates in the manner of Sleijpen and
// van der Vorst): the direction and r.z carry on, so the recurrence keeps
// its Krylov space while the residual it works on stays the true one.
template <class I>
static int inner_solve(Ctx &c, bool fac, int n_iter, double reduc, InnerCarry &carry, float f, int &iters,
                       int &broken, double &alpha_pend, KTimer *T, double *x, const double *b, double sigma) {
  Mix &X = c.mx;
  const int n = c.n, G = red_grid(n), ws = c.Ws, wl = X.WL, wu = X.WU;
  const I *col = (const I *)X.col, *lcol = (const I *)X.lcol, *ucol = (const I *)X.ucol;
  const lval_t *lval = (const lval_t *)X.lval, *uval = (const lval_t *)X.uval;
  float *r[2] = {fv(c, V_R0), fv(c, V_R1)}, *P[2] = {fv(c, V_P0), fv(c, V_P1)};
  const float *p_last = carry.p_buf >= 0 ? P[carry.p_buf] : nullptr;
  // scalars: reduc, n_iter, s = large (restart) or the carried r.z
  KS h[6];
  std::memset(h, 0, sizeof h);
  for (int s = 0; s < 2; ++s) {
    h[s * 3].reduc = reduc;
    h[s * 3].n_iter = n_iter;
    h[s * 3].s = p_last ? carry.s * (double)f * (double)f : 1.0 / SMALL;
  }
  FVK_HIP(hipMemcpyAsync(c.ksr, h, sizeof h, hipMemcpyHostToDevice, c.stream));
  float *q = fv(c, V_Q), *ev = fv(c, V_E);
  t16_t *t = (t16_t *)fv(c, V_T);
  float *t0 = fv(c, V_A);  // the start's t, FP32
  const lval_t *ee = (const lval_t *)X.e;
  // partials: [0] |r|_1 and [3] r.z of the last update, [1] p.q
  double *pU = pb(c, 0), *pQ = pb(c, 1);
#define L(W)                                                                                                   \
  k_mx_start<W, I, float><<<G, RB, 0, c.stream>>>(n, fac ? 1 : 0, r[0], ev, P[0], p_last, f, wu, ucol, uval, ee, \
                                                   t0, pU, nullptr)
  FVK_W(wu, L);
#undef L
  const int chunk = 32;
  Poller Pl(c);
  int it = 0, next_poll = chunk;
  const int R = X.replace_every;
  for (; it < n_iter; ++it) {
    const bool rep = R > 0 && it > 0 && it % R == 0;
    if (rep) {
      // the scalars of iteration it - 1 gate the three kernels
      KS *prev = c.ksr + ((it + 1) & 1) * 3;
      float *rr = fac ? r[it & 1] : r[0];
      k_mx_rep_x<<<grid_for(n), BS, 0, c.stream>>>(n, x, ev, fac ? P[0] : nullptr, c.diag, sigma, prev);
#define L(W) k_mx_res<W><<<G, RB, 0, c.stream>>>(c.M, c.diag, c.spv, x, b, rr, pb(c, 3), prev, 1.0 / sigma)
      FVK_W(c.Ws, L);
#undef L
#define L(W)                                                                                                 \
  k_mx_start<W, I, float><<<G, RB, 0, c.stream>>>(n, fac ? 1 : 0, rr, nullptr, nullptr, nullptr, 1.f, wu, ucol, \
                                                   uval, ee, t0, pU, prev)
      FVK_W(wu, L);
#undef L
    }
    hipEvent_t ea, eb;
    const bool stamp = T && T->take(ea, eb);
    if (!fac) {
      // Jacobi: z is r itself, updated in place; p double-buffered
      const float *pold = P[it & 1], *sp = X.spv, *z = r[0];
      float *pnew = P[(it + 1) & 1];
      if (stamp) {
#define L(W)                                                                                                       \
  hipExtLaunchKernelGGL(k_mx_pq<W, I>, dim3(G), dim3(RB), 0, c.stream, ea, eb, 0, n, ws, it, col, sp, z, pold, pnew, \
                        q, (const double *)pU, G, pQ, c.ksr, c.ks)
        FVK_W(ws, L);
#undef L
      } else {
#define L(W) k_mx_pq<W, I><<<G, RB, 0, c.stream>>>(n, ws, it, col, sp, z, pold, pnew, q, pU, G, pQ, c.ksr, c.ks)
        FVK_W(ws, L);
#undef L
      }
      k_mx_upd_jac<<<G, RB, 0, c.stream>>>(n, it, ev, r[0], pnew, q, pQ, G, pU, c.ksr);
    } else {
      // factor: p in place in P[0]; r double-buffered
      float *p = P[0], *rin = r[it & 1], *rout = r[(it + 1) & 1];
      const float *pc = p;
      if (it == 0 || rep) {
#define L(W) k_mx_zp<W, I, float><<<G, RB, 0, c.stream>>>(n, wl, it, lcol, lval, t0, p, ev, pU, G, c.ksr, c.ks)
        FVK_W(wl, L);
#undef L
      } else {
#define L(W) k_mx_zp<W, I, t16_t><<<G, RB, 0, c.stream>>>(n, wl, it, lcol, lval, t, p, ev, pU, G, c.ksr, c.ks)
        FVK_W(wl, L);
#undef L
      }
      const float *au = X.au;
      const unsigned char *iau = X.iau;
      if (X.qfull) {
        const float *sp = X.spv;
        if (stamp) {
#define L(W)                                                                                                     \
  hipExtLaunchKernelGGL(k_mx_qf<W, I>, dim3(G), dim3(RB), 0, c.stream, ea, eb, 0, n, ws, it, col, sp, pc, q, pQ, \
                        (const KS *)c.ksr)
          FVK_W(ws, L);
#undef L
        } else {
#define L(W) k_mx_qf<W, I><<<G, RB, 0, c.stream>>>(n, ws, it, col, sp, pc, q, pQ, c.ksr)
          FVK_W(ws, L);
#undef L
        }
      } else if (stamp) {
#define L(W)                                                                                                      \
  hipExtLaunchKernelGGL(k_mx_q<W, I>, dim3(G), dim3(RB), 0, c.stream, ea, eb, 0, n, wl, it, ucol, lcol, au, iau, pc, \
                        q, pQ, (const KS *)c.ksr)
        FVK_W(wl, L);
#undef L
      } else {
#define L(W) k_mx_q<W, I><<<G, RB, 0, c.stream>>>(n, wl, it, ucol, lcol, au, iau, pc, q, pQ, c.ksr)
        FVK_W(wl, L);
#undef L
      }
#define L(W)                                                                                                        \
  k_mx_upd_fac<W, I><<<G, RB, 0, c.stream>>>(n, it, rin, rout, q, wu, ucol, uval, ee, t, pQ, G, pU, c.ksr)
      FVK_W(wu, L);
#undef L
    }
    if (it + 1 == next_poll && it + 1 < n_iter) {
      if (T) {
        KS k;
        if (T->flush()) return FVK_ERR_HIP;
        FVK_HIP(hipMemcpyAsync(&k, c.ks, sizeof k, hipMemcpyDeviceToHost, c.stream));
        FVK_HIP(hipStreamSynchronize(c.stream));
        if (k.done) { ++it; break; }
        next_poll += chunk;
        continue;
      }
      bool stop = false;
      int next = chunk;
      FVK_HIP(Pl.step(it + 1, chunk, reduc, stop, next));
      if (stop) { ++it; break; }
      next_poll = it + 1 + next;
    }
  }
  k_mx_final<<<1, RB, 0, c.stream>>>(it, pU, G, c.ksr, c.ks);
  FVK_HIP(hipGetLastError());
  if (T && T->flush()) return FVK_ERR_HIP;
  KS k, kr[2];
  FVK_HIP(hipMemcpyAsync(&k, c.ks, sizeof k, hipMemcpyDeviceToHost, c.stream));
  FVK_HIP(hipMemcpyAsync(&kr[0], c.ksr, sizeof(KS), hipMemcpyDeviceToHost, c.stream));
  FVK_HIP(hipMemcpyAsync(&kr[1], c.ksr + 3, sizeof(KS), hipMemcpyDeviceToHost, c.stream));
  FVK_HIP(hipStreamSynchronize(c.stream));
  iters = k.iter;
  broken = kr[0].pad || kr[1].pad;
  // the decision that ended the solve kept the last update's alpha (0 when
  // k_mx_zp had applied it, or when no update ran)
  alpha_pend = fac ? k.alpha : 0.0;
  // The last direction: the factor path updates P[0] in place; with Jacobi
  // iteration j writes P[(j + 1) & 1], so the last of the iters iterations
  // left it in P[iters & 1] (P[0] when none ran).
  carry.p_buf = broken || !(k.s > 0.0) || !std::isfinite(k.s) ? -1 : (fac ? 0 : (iters & 1));
  carry.s = k.s;
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
  // the factor path's SpMV reads the upper entries only (au), or the full
  // rows where those are fewer bytes
  if (!use_factor || X.qfull)
    k_mx_conv<<<grid_for(n), BS, 0, c.stream>>>(c.M, c.diag, c.spv, X.spv);
  else
    k_mx_au<<<grid_for(n), BS, 0, c.stream>>>(n, X.WU, X.ufi, X.fi_face, X.fi_n, c.ngp, c.diag, X.au);
  FVK_HIP(hipMemsetAsync(x, 0, sizeof(double) * n, c.stream));
  res = DcResult();
  InnerCarry carry;
  double sigma_prev = 0.0;
  for (int outer = 0;; ++outer) {
#define L(W) k_mx_res<W><<<G, RB, 0, c.stream>>>(c.M, c.diag, c.spv, x, b, fv(c, V_R0), pb(c, 3), nullptr, 1.0)
    FVK_W(c.Ws, L);
#undef L
    k_mx_tot4<<<1, RB, 0, c.stream>>>(pb(c, 3), G, X.tot);
    double h[4];
    FVK_HIP(hipMemcpyAsync(h, X.tot, sizeof h, hipMemcpyDeviceToHost, c.stream));
    FVK_HIP(hipStreamSynchronize(c.stream));
    if (outer == 0 && !std::isfinite(h[2])) {
      set_error("non-finite right-hand side");
      return FVK_ERR_ARG;
    }
    res.b_l1 = h[2];
    res.res_l1 = h[2] > 0 ? h[0] / h[2] : h[0];
    res.res_l2 = h[3] > 0 ? std::sqrt(h[1] / h[3]) : std::sqrt(h[1]);
    res.outer = outer;
    if (!(res.res_l1 == res.res_l1)) { res.broken = 1; break; }
    if (res.res_l1 < tol) { res.converged = true; break; }
    if (outer >= X.max_outer || res.iters >= max_iter || res.broken) break;
    // inner target: what is left to reach tol (with a margin), but no
    // deeper than inner_reduc, where FP32 rounding takes over.  With residual
    // replacements the inner solve can go all the way, as long as |r|_2^2 of
    // the normalised residual (at least n eta^2) stays far above the absolute
    // small = 1e-20 in alpha = s/(p.q + small).
    const double floor = X.replace_every > 0 ? std::sqrt(1e-12 / n) : X.inner_reduc;
    const double eta = std::min(0.5, std::max(floor, X.eta_margin * tol / res.res_l1));
    // The inner right-hand side is normalised to a mean |r| of 1: the PCG
    // recurrence's alpha = s/(p.q + small) (solv_cg.f90:69-70) carries an
    // absolute small = 1e-20 that stalls the solve once p.q is that tiny,
    // which the residuals of later correction steps reach.
    const double sigma = h[0] > 0 ? h[0] / n : 1.0;
    k_scale32<<<grid_for(n), BS, 0, c.stream>>>(n, fv(c, V_R0), (float)(1.0 / sigma));
    int its = 0, broken = 0;
    if (!X.carry) carry.p_buf = -1;
    const float f = carry.p_buf >= 0 ? (float)(sigma_prev / sigma) : 1.f;
    double alpha_pend = 0.0;
    rc = X.i16 ? inner_solve<short>(c, use_factor, max_iter - res.iters, eta, carry, f, its, broken, alpha_pend, T,
                                    x, b, sigma)
               : inner_solve<int>(c, use_factor, max_iter - res.iters, eta, carry, f, its, broken, alpha_pend, T, x,
                                  b, sigma);
    if (rc) return rc;
    sigma_prev = sigma;
    if (outer < 8) {
      res.inner[outer] = its;
      res.res_before[outer] = res.res_l1;
    }
    res.iters += its;
    // a breakdown (p.q <= 0 or non-finite) ends the inner solve with the
    // finite correction made so far, which is kept; the next correction
    // step restarts from that x (a fresh inner solve re-derives its scales
    // from the new residual).  After X.max_breakdowns restarts the solve
    // gives up and the caller's fallback takes over.
    if (broken && ++res.breakdowns > X.max_breakdowns) res.broken = 1;
    k_mx_add<<<grid_for(n), BS, 0, c.stream>>>(n, x, fv(c, V_E), use_factor ? fv(c, V_P0) : nullptr,
                                               (float)alpha_pend, c.diag, sigma);
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
#define L(W) \
  k_mx_start<W, I, float><<<G, RB, 0, c.stream>>>(n, 1, r32, ev, p0, nullptr, 1.f, wu, ucol, uval, \
                                                  (const lval_t *)X.e, t, pb(c, 0), nullptr)
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
