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
// without FMA contraction).
//
// Reductions (dot products, L1 norms) never synchronise inside a kernel.  A
// producer kernel leaves one partial per block (plain stores); the next
// kernel on the stream, where the kernel boundary makes them visible, adds
// the partials in every block in the same fixed tree, so all blocks hold the
// same bits and take the same decisions (alpha, beta, exit test) without a
// device-scope fence.  Block 0 commits the scalars into a two-slot ring
// indexed by iteration parity, so a kernel never rewrites a slot that its
// own blocks read.  The order differs from the reference's sequential sums
// by rounding only.
#include <algorithm>
#include <cstring>

#include "cgcore.h"
#include "sweep.h"

namespace fvk {

// ---------------------------------------------------------------- kernels

__global__ __launch_bounds__(BS) void k_precond_jacobi(int n, const double *__restrict__ diag,
                                                       double *__restrict__ c, double *__restrict__ cr) {
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
    const int j = M.spc[(size_t)k * M.ld + r];
    if (j < 0) break;
    if (M.rank[j] < rr) {
      const int f = M.spf[(size_t)k * M.ld + r];
      if (imat == 1) d = d - ngn[f] * ngn[f] * c[j];
      else d = d - ngp[f] * c[j] * ngn[f];
    }
  }
  cr[r] = d;
  c[r] = 1.0 / (d + SMALL);
}

__global__ __launch_bounds__(BS) void k_jacobi_apply(int n, int nvc, const double *__restrict__ c,
                                                     const double *__restrict__ b, double *__restrict__ x,
                                                     const KS *ks) {
  int i = blockIdx.x * BS + threadIdx.x;
  if (i >= n || all_done(ks, nvc)) return;
  for (int q = 0; q < nvc; ++q) x[q * n + i] = b[q * n + i] * c[i];
}

// Forward sweep over one level: x = (b - L x) * c   (solv_direct.f90:12-24)
__global__ __launch_bounds__(BS) void k_sweep_fwd(DMesh M, const int *__restrict__ rows, int cnt, int nvc,
                                                  const double *__restrict__ spv, const double *__restrict__ c,
                                                  const double *__restrict__ b, double *__restrict__ x,
                                                  const KS *ks) {
  int t = blockIdx.x * BS + threadIdx.x;
  if (t >= cnt || all_done(ks, nvc)) return;
  const int r = rows[t], n = M.n, rr = M.rank[r];
  double acc[3];
  for (int q = 0; q < nvc; ++q) acc[q] = b[q * n + r];
  for (int k = 0; k < M.Ws; ++k) {
    const int j = M.spc[(size_t)k * M.ld + r];
    if (j < 0) break;
    if (M.rank[j] < rr) {
      const double a = spv[(size_t)k * M.ld + r];
      for (int q = 0; q < nvc; ++q) acc[q] = acc[q] - a * x[q * n + j];
    }
  }
  const double cc = c[r];
  for (int q = 0; q < nvc; ++q) x[q * n + r] = acc[q] * cc;
}

// Backward sweep over one level: x = (x*c_r - U x) * c   (solv_direct.f90:26-38)
__global__ __launch_bounds__(BS) void k_sweep_bwd(DMesh M, const int *__restrict__ rows, int cnt, int nvc,
                                                  const double *__restrict__ spv, const double *__restrict__ c,
                                                  const double *__restrict__ cr, double *__restrict__ x,
                                                  const KS *ks) {
  int t = blockIdx.x * BS + threadIdx.x;
  if (t >= cnt || all_done(ks, nvc)) return;
  const int r = rows[t], n = M.n, rr = M.rank[r];
  double acc[3];
  const double crr = cr[r];
  for (int q = 0; q < nvc; ++q) acc[q] = x[q * n + r] * crr;
  for (int k = 0; k < M.Ws; ++k) {
    const int j = M.spc[(size_t)k * M.ld + r];
    if (j < 0) break;
    if (M.rank[j] > rr) {
      const double a = spv[(size_t)k * M.ld + r];
      for (int q = 0; q < nvc; ++q) acc[q] = acc[q] - a * x[q * n + j];
    }
  }
  const double cc = c[r];
  for (int q = 0; q < nvc; ++q) x[q * n + r] = acc[q] * cc;
}

// ---------------------------------------------------------------- one-launch sweeps
// The factorisation and the triangular sweeps in one launch each.  Rows are
// taken in level order (every row after the rows it depends on) in chunks
// of BS positions; a resident grid of a few blocks per CU takes the chunks
// from a ticket counter, so the rows in flight stay near the front of the
// order and every row a chunk waits for belongs to a chunk taken earlier,
// by a block that is running or has finished.  An output starts as
// kSweepMark (a signalling NaN: no arithmetic result has this pattern) and
// a row waits until each value it needs has been replaced, taking the values
// in slot order, so each row performs the same operations in the same order
// as the level-by-level kernels above.  Waiting lanes never block their
// wave: the wave loops until all its lanes are finished, and lanes whose
// values have arrived finish in the same pass.  Every wait is bounded
// (kSweepSpin passes per chunk); past the bound the kernel records an error
// and its blocks stop taking chunks.  lmask / umask: bit k of a row is set
// when slot k holds a lower / upper neighbour (rank order).
// out[0 .. cnt) = kSweepMark and *ticket = 0, unless the gate is closed
__global__ __launch_bounds__(BS) void k_sweep_mark(size_t cnt, double *__restrict__ out, int *__restrict__ ticket,
                                                   const KS *ks, int nvc) {
  if (all_done(ks, nvc)) return;
  unsigned long long *o = (unsigned long long *)out;
  for (size_t i = blockIdx.x * (size_t)BS + threadIdx.x; i < cnt; i += (size_t)gridDim.x * BS) o[i] = kSweepMark;
  if (blockIdx.x == 0 && threadIdx.x == 0) *ticket = 0;
}

__global__ __launch_bounds__(BS) void k_sweep_masks(DMesh M, unsigned char *__restrict__ lmask,
                                                    unsigned char *__restrict__ umask) {
  const int r = blockIdx.x * BS + threadIdx.x;
  if (r >= M.n) return;
  const int rr = M.rank[r];
  unsigned lo = 0, up = 0;
  for (int k = 0; k < M.Ws; ++k) {
    const int j = M.spc[(size_t)k * M.ld + r];
    if (j < 0) break;
    if (M.rank[j] < rr) lo |= 1u << k;
    else if (M.rank[j] > rr) up |= 1u << k;
  }
  lmask[r] = (unsigned char)lo;
  umask[r] = (unsigned char)up;
}

// A row keeps its slots' columns and coefficients in registers (read once,
// all slots at a time) and each pass issues the loads of all values it still
// waits for at once, then takes them in slot order up to the first missing
// one: one memory round trip per pass.  At most kSweepSlots slots.

// DIC (imat 1) / DILU (imat 2) factor (solv_precond.f90:10-29)
struct PrecondRow {
  const int *spc, *spf, *order;
  size_t ld;
  int Ws;
  const unsigned char *lmask;
  int imat;
  const double *diag, *ngp, *ngn;
  double *c, *cr;
  int r, k;
  unsigned m;
  double d;
  int j[kSweepSlots];
  double a1[kSweepSlots], a2[kSweepSlots];
  __device__ void start(int pos) {
    r = order[pos];
    m = lmask[r];
    d = diag[r];
    k = 0;
#pragma unroll
    for (int s = 0; s < kSweepSlots; ++s) {
      j[s] = 0;
      a1[s] = a2[s] = 0.0;
      if (s < Ws && ((m >> s) & 1u)) {
        j[s] = spc[(size_t)s * ld + r];
        const int f = spf[(size_t)s * ld + r];
        a1[s] = imat == 1 ? ngn[f] : ngp[f];
        a2[s] = ngn[f];
      }
    }
  }
  __device__ bool advance() {
    unsigned long long v[kSweepSlots];
#pragma unroll
    for (int s = 0; s < kSweepSlots; ++s)
      if (s >= k && s < Ws && ((m >> s) & 1u)) v[s] = ld_agent(c + j[s]);
#pragma unroll
    for (int s = 0; s < kSweepSlots; ++s) {
      if (s < k || s >= Ws || !((m >> s) & 1u)) continue;
      if (v[s] == kSweepMark) break;
      const double cj = __longlong_as_double((long long)v[s]);
      if (imat == 1) d = d - a1[s] * a2[s] * cj;
      else d = d - a1[s] * cj * a2[s];
      k = s + 1;
    }
    // k: the slot after the last one taken; done when no lower slot remains
    if ((m >> k) != 0u) return false;
    cr[r] = d;
    st_agent(c + r, 1.0 / (d + SMALL));
    return true;
  }
};

__global__ __launch_bounds__(BS) void k_precond_sf(DMesh M, const int *__restrict__ order,
                                                   const unsigned char *__restrict__ lmask, int imat,
                                                   const double *__restrict__ diag, const double *__restrict__ ngp,
                                                   const double *__restrict__ ngn, double *c,
                                                   double *__restrict__ cr, int *ticket, int *err) {
  PrecondRow row;
  row.spc = M.spc; row.spf = M.spf; row.order = order; row.ld = M.ld; row.Ws = M.Ws;
  row.lmask = lmask; row.imat = imat; row.diag = diag; row.ngp = ngp; row.ngn = ngn; row.c = c; row.cr = cr;
  row.r = 0; row.k = 0; row.m = 0; row.d = 0.0;
  sweep_chunks(M.n, ticket, err, row);
}

// Forward sweep t = (b - L t) * c (solv_direct.f90:12-24) and backward
// sweep x = (t*c_r - U x) * c (:26-38) for NVC right-hand sides: out is
// written, src read (b or t), the waits are on out.
template <bool FWD, int NVC>
struct SweepRow {
  const int *spc, *order;
  size_t ld;
  int Ws, n;
  const unsigned char *mask;
  const double *spv, *c, *cr, *src;
  double *out;
  int r, k;
  unsigned m;
  double acc[NVC];
  int j[kSweepSlots];
  double a[kSweepSlots];
  __device__ void start(int pos) {
    r = order[pos];
    m = mask[r];
    k = 0;
#pragma unroll
    for (int s = 0; s < kSweepSlots; ++s) {
      j[s] = 0;
      a[s] = 0.0;
      if (s < Ws && ((m >> s) & 1u)) {
        j[s] = spc[(size_t)s * ld + r];
        a[s] = spv[(size_t)s * ld + r];
      }
    }
    if (FWD) {
#pragma unroll
      for (int q = 0; q < NVC; ++q) acc[q] = src[q * n + r];
    } else {
      const double crr = cr[r];
#pragma unroll
      for (int q = 0; q < NVC; ++q) acc[q] = src[q * n + r] * crr;
    }
  }
  __device__ bool advance() {
    unsigned long long v[kSweepSlots][NVC];
#pragma unroll
    for (int s = 0; s < kSweepSlots; ++s)
      if (s >= k && s < Ws && ((m >> s) & 1u))
#pragma unroll
        for (int q = 0; q < NVC; ++q) v[s][q] = ld_agent(out + q * n + j[s]);
#pragma unroll
    for (int s = 0; s < kSweepSlots; ++s) {
      if (s < k || s >= Ws || !((m >> s) & 1u)) continue;
      bool ready = true;
#pragma unroll
      for (int q = 0; q < NVC; ++q) ready = ready && v[s][q] != kSweepMark;
      if (!ready) break;
#pragma unroll
      for (int q = 0; q < NVC; ++q) acc[q] = acc[q] - a[s] * __longlong_as_double((long long)v[s][q]);
      k = s + 1;
    }
    if ((m >> k) != 0u) return false;
    const double cc = c[r];
#pragma unroll
    for (int q = 0; q < NVC; ++q) st_agent(out + q * n + r, acc[q] * cc);
    return true;
  }
};

template <bool FWD, int NVC>
__device__ __forceinline__ void sweep_run(const DMesh &M, const int *order, const unsigned char *mask,
                                          const double *spv, const double *c, const double *cr,
                                          const double *src, double *out, int *ticket, int *err) {
  SweepRow<FWD, NVC> row;
  row.spc = M.spc; row.order = order; row.ld = M.ld; row.Ws = M.Ws; row.n = M.n; row.mask = mask;
  row.spv = spv; row.c = c; row.cr = cr; row.src = src; row.out = out;
  row.r = 0; row.k = 0; row.m = 0;
  sweep_chunks(M.n, ticket, err, row);
}

template <int NVC>
__global__ __launch_bounds__(BS) void k_sweep_fwd_sf(DMesh M, const int *__restrict__ order,
                                                     const unsigned char *__restrict__ lmask,
                                                     const double *__restrict__ spv, const double *__restrict__ c,
                                                     const double *__restrict__ b, double *t, int *ticket, int *err,
                                                     const KS *ks) {
  if (all_done(ks, NVC)) return;
  sweep_run<true, NVC>(M, order, lmask, spv, c, nullptr, b, t, ticket, err);
}

template <int NVC>
__global__ __launch_bounds__(BS) void k_sweep_bwd_sf(DMesh M, const int *__restrict__ order,
                                                     const unsigned char *__restrict__ umask,
                                                     const double *__restrict__ spv, const double *__restrict__ c,
                                                     const double *__restrict__ cr, const double *__restrict__ t,
                                                     double *x, int *ticket, int *err, const KS *ks) {
  if (all_done(ks, NVC)) return;
  sweep_run<false, NVC>(M, order, umask, spv, c, cr, t, x, ticket, err);
}

// The same factor and sweeps for sweep_parts: a position's descriptor holds
// the row, its slot mask (first), the number of lower (upper) slots and the
// columns of those slots in slot order.
__device__ __forceinline__ void slots_of(unsigned m, int (&sl)[kSweepSlots]) {
#pragma unroll
  for (int t = 0; t < kSweepSlots; ++t) {
    sl[t] = m ? __builtin_ctz(m) : 0;
    m &= m - 1u;
  }
}

struct PtPrecond {
  const int *spf;
  size_t ld;
  int imat;
  const double *diag, *ngp, *ngn;
  double *cr;
  double d;
  double a1[kSweepSlots], a2[kSweepSlots];
  __device__ void fetch(const SweepDesc &q, int MW) {
    int sl[kSweepSlots];
    slots_of((unsigned)q.first, sl);
    d = diag[q.r];
#pragma unroll
    for (int t = 0; t < kSweepSlots; ++t)
      if (t < MW) {
        const int f0 = spf[(size_t)sl[t] * ld + q.r];
        const int f = t < q.nl ? f0 : 0;
        a1[t] = imat == 1 ? ngn[f] : ngp[f];
        a2[t] = ngn[f];
      }
  }
  __device__ void begin() {}
  __device__ void take(int t, const double *w) {
    if (imat == 1) d = d - a1[t] * a2[t] * w[0];
    else d = d - a1[t] * w[0] * a2[t];
  }
  __device__ double result(int) const { return 1.0 / (d + SMALL); }
  __device__ void done(const SweepDesc &q) { cr[q.r] = d; }
};

template <bool FWD, int NVC>
struct PtSweep {
  const double *spv, *c, *cr, *src;
  size_t ld;
  int n;
  double cc, crr, acc[NVC];
  double a[kSweepSlots];
  __device__ void fetch(const SweepDesc &q, int MW) {
    int sl[kSweepSlots];
    slots_of((unsigned)q.first, sl);
#pragma unroll
    for (int t = 0; t < kSweepSlots; ++t)
      if (t < MW) a[t] = spv[(size_t)sl[t] * ld + q.r];
    cc = c[q.r];
    if (!FWD) crr = cr[q.r];
#pragma unroll
    for (int v = 0; v < NVC; ++v) acc[v] = src[v * n + q.r];
  }
  __device__ void begin() {
    if (!FWD) {
#pragma unroll
      for (int v = 0; v < NVC; ++v) acc[v] = acc[v] * crr;
    }
  }
  __device__ void take(int t, const double *w) {
#pragma unroll
    for (int v = 0; v < NVC; ++v) acc[v] = acc[v] - a[t] * w[v];
  }
  __device__ double result(int v) const { return acc[v] * cc; }
  __device__ void done(const SweepDesc &) {}
};

template <int MW, int TB>
__global__ __launch_bounds__(TB) void k_precond_pt(int n, int R, int P, SweepTabs T, const int *__restrict__ step,
                                                   const int *__restrict__ sp, DMesh M, int imat,
                                                   const double *__restrict__ diag, const double *__restrict__ ngp,
                                                   const double *__restrict__ ngn, double *c, double *__restrict__ cr,
                                                   int *ticket, int *err) {
  PtPrecond row;
  row.spf = M.spf; row.ld = M.ld; row.imat = imat; row.diag = diag; row.ngp = ngp; row.ngn = ngn; row.cr = cr;
  sweep_parts<1, MW, PtPrecond, TB>(n, R, P, T, step, sp, c, ticket, err, row);
}

template <bool FWD, int NVC, int MW, int TB>
__global__ __launch_bounds__(TB) void k_sweep_pt(int n, int R, int P, SweepTabs T, const int *__restrict__ step,
                                                 const int *__restrict__ sp, size_t ld, const double *__restrict__ spv,
                                                 const double *__restrict__ c, const double *__restrict__ cr,
                                                 const double *__restrict__ src, double *out, int *ticket, int *err,
                                                 const KS *ks) {
  if (all_done(ks, NVC)) return;
  PtSweep<FWD, NVC> row;
  row.spv = spv; row.c = c; row.cr = cr; row.src = src; row.ld = ld; row.n = n;
  sweep_parts<NVC, MW, PtSweep<FWD, NVC>, TB>(n, R, P, T, step, sp, out, ticket, err, row);
}

template <int WS>
__global__ __launch_bounds__(BS) void k_spmv(DMesh M, int nvc, const double *__restrict__ diag,
                                             const double *__restrict__ spv, const double *__restrict__ x,
                                             double *__restrict__ y) {
  int i = blockIdx.x * BS + threadIdx.x;
  if (i >= M.n) return;
  for (int q = 0; q < nvc; ++q) y[q * M.n + i] = row_spmv<WS>(M, diag, spv, x + q * M.n, i);
}

// r = b - A x (or r = b) with partials {|r|_1[q], s2[q]} where s2 = r.(c r)
// and z = c r (mode 1, Jacobi PCG), s2 = r.r (mode 2, BiCGStab) or 0.
// (solv_res.f90; solv_cg.f90:19-33; solv_bicg.f90:19-38)
template <int WS>
__global__ __launch_bounds__(RB64) void k_residual(DMesh M, int nvc, const double *__restrict__ diag,
                                                 const double *__restrict__ spv, const double *__restrict__ x,
                                                 double *__restrict__ r, int not_r0, int mode,
                                                 const double *__restrict__ c, double *__restrict__ z,
                                                 double *__restrict__ part) {
  const int n = M.n;
  double a[6] = {0, 0, 0, 0, 0, 0};
  int r0, r1;
  block_rows<RB64>(n, r0, r1);
  for (int i = r0 + threadIdx.x; i < r1; i += RB64)
    for (int q = 0; q < nvc; ++q) {
      double v = r[q * n + i];
      if (!not_r0) {
        const double *xq = x + q * n;
        v = v - diag[i] * xq[i];
#pragma unroll
        for (int k = 0; k < nslots<WS>(M); ++k) {
          const int j = M.spc[(size_t)k * M.ld + i];
          const double a = spv[(size_t)k * M.ld + i];
          if (j >= 0) v = v - a * xq[j];
        }
        r[q * n + i] = v;
      }
      a[q] += fabs(v);
      if (mode == 1) {
        const double zi = v * c[i];
        z[q * n + i] = zi;
        a[3 + q] += v * zi;
      } else if (mode == 2) {
        a[3 + q] += v * v;
      }
    }
  store_partials<6, RB64>(a, part);
}

// Decide iteration it, then p = z + beta p ; q = A p ; partial p.q.  p of the
// neighbours is formed on the fly from (z, p_old): p_new is written once.
// Each thread takes two rows per pass so that both rows' loads are in flight
// together.
template <int NVC, int WS>
__device__ __forceinline__ void cg_pq_row(const DMesh &M, int i, const double *__restrict__ diag,
                                          const double *__restrict__ spv, const double *__restrict__ zq,
                                          const double *__restrict__ po, double beta, double *__restrict__ pn,
                                          double *__restrict__ qq, double &pq) {
  const int n = M.n;
  const double pi = zq[i] + beta * po[i];
  double y = diag[i] * pi;
#pragma unroll
  for (int k = 0; k < nslots<WS>(M); ++k) {
    const int j = M.spc[(size_t)k * M.ld + i];
    const double a = spv[(size_t)k * M.ld + i];
    if (j >= 0) y = y + a * (zq[j] + beta * po[j]);
  }
  pn[i] = pi;
  qq[i] = y;
  pq += pi * y;
}

template <int NVC, int WS>
__global__ __launch_bounds__(RB64) void k_cg_pq(DMesh M, int it, const double *__restrict__ diag,
                                              const double *__restrict__ spv, const double *__restrict__ z,
                                              const double *__restrict__ pold, double *__restrict__ pnew,
                                              double *__restrict__ qv, const double *__restrict__ partU, int PU,
                                              const double *__restrict__ partR, int PR, double *__restrict__ partQ,
                                              KS *ring, KS *out) {
  // |r|_1 (components 0..NVC) and, for Jacobi, r.z (components 3..3+NVC)
  double L1[NVC], rz[NVC];
  if (partR) {
    load_totals<NVC, RB64>(partU, PU, L1);
    load_totals<NVC, RB64>(partR, PR, rz);
  } else {
    int cs[2 * NVC];
    double t[2 * NVC];
    for (int q = 0; q < NVC; ++q) { cs[q] = q; cs[NVC + q] = 3 + q; }
    load_sel<2 * NVC, RB64>(partU, PU, cs, t);
    for (int q = 0; q < NVC; ++q) { L1[q] = t[q]; rz[q] = t[NVC + q]; }
  }
  __shared__ KS sk[3];
  if (threadIdx.x == 0) {
    cg_decide(it, NVC, L1, rz, ring + ((it + 1) & 1) * 3, sk);
    if (blockIdx.x == 0)
      for (int q = 0; q < NVC; ++q) ring[(it & 1) * 3 + q] = out[q] = sk[q];
  }
  __syncthreads();
  double beta[NVC];
  int dn[NVC];
  bool any = false;
  for (int q = 0; q < NVC; ++q) {
    beta[q] = sk[q].beta;
    dn[q] = sk[q].done;
    any = any || !dn[q];
  }
  if (!any) return;
  const int n = M.n;
  double pq[NVC];
  for (int q = 0; q < NVC; ++q) pq[q] = 0.0;
  int r0, r1;
  block_rows<RB64>(n, r0, r1);
  int i = r0 + threadIdx.x;
  for (; i + RB64 < r1; i += 2 * RB64)
#pragma unroll
    for (int q = 0; q < NVC; ++q) {
      if (dn[q]) continue;
      const double *zq = z + q * n, *po = pold + q * n;
      double *pn = pnew + q * n, *qq = qv + q * n;
      cg_pq_row<NVC, WS>(M, i, diag, spv, zq, po, beta[q], pn, qq, pq[q]);
      cg_pq_row<NVC, WS>(M, i + RB64, diag, spv, zq, po, beta[q], pn, qq, pq[q]);
    }
  if (i < r1)
#pragma unroll
    for (int q = 0; q < NVC; ++q) {
      if (dn[q]) continue;
      cg_pq_row<NVC, WS>(M, i, diag, spv, z + q * n, pold + q * n, beta[q], pnew + q * n, qv + q * n, pq[q]);
    }
  store_partials<NVC, RB64>(pq, partQ);
}

// alpha = s/(p.q + small) ; x += alpha p ; r -= alpha q ; partials |r|_1 and
// (Jacobi) r.z with z = c r.
template <int NVC>
__global__ __launch_bounds__(RB64) void k_cg_upd(int n, int it, double *__restrict__ x, double *__restrict__ r,
                                               const double *__restrict__ p, const double *__restrict__ qv,
                                               const double *__restrict__ c, int jacobi, double *__restrict__ z,
                                               const double *__restrict__ partQ, int PQ, double *__restrict__ partU,
                                               KS *ring) {
  const KS *cur = ring + (it & 1) * 3;
  bool any = false;
  for (int q = 0; q < NVC; ++q) any = any || !cur[q].done;
  if (!any) return;
  double t[NVC];
  load_totals<NVC, RB64>(partQ, PQ, t);
  double alpha[NVC];
  int dn[NVC];
  for (int q = 0; q < NVC; ++q) {
    dn[q] = cur[q].done;
    alpha[q] = cur[q].s / (t[q] + SMALL);
  }
  double a[6] = {0, 0, 0, 0, 0, 0};
  int r0, r1;
  block_rows<RB64>(n, r0, r1);
  for (int i = r0 + threadIdx.x; i < r1; i += RB64)
#pragma unroll
    for (int q = 0; q < NVC; ++q) {
      if (dn[q]) continue;
      const int o = q * n + i;
      x[o] = x[o] + alpha[q] * p[o];
      const double ri = r[o] - alpha[q] * qv[o];
      r[o] = ri;
      a[q] += fabs(ri);
      if (jacobi) {
        const double zi = ri * c[i];
        z[o] = zi;
        a[3 + q] += ri * zi;
      }
    }
  store_partials<6, RB64>(a, partU);
  if (blockIdx.x == 0 && threadIdx.x == 0)
    for (int q = 0; q < NVC; ++q) ring[(it & 1) * 3 + q].alpha = alpha[q];
}

// r.z after the sweeps (DIC / DILU)
__global__ __launch_bounds__(RB64) void k_dot(int n, int nvc, const double *__restrict__ a,
                                            const double *__restrict__ b, double *__restrict__ part, const KS *gate) {
  double s[3] = {0, 0, 0};
  if (!all_done(gate, nvc)) {
    int r0, r1;
    block_rows<RB64>(n, r0, r1);
    for (int i = r0 + threadIdx.x; i < r1; i += RB64)
      for (int q = 0; q < nvc; ++q) s[q] += a[q * n + i] * b[q * n + i];
  }
  store_partials<3, RB64>(s, part);
}

// Final decision after the last queued iteration (1 block).
__global__ __launch_bounds__(RB64) void k_cg_final(int it, int nvc, const double *partU, int PU, const double *partR,
                                                 int PR, KS *ring, KS *out) {
  double t[6];
  load_totals<6, RB64>(partU, PU, t);
  double rz[3] = {t[3], t[4], t[5]};
  if (partR) {
    double tr[3];
    load_totals<3, RB64>(partR, PR, tr);
    rz[0] = tr[0]; rz[1] = tr[1]; rz[2] = tr[2];
  }
  if (threadIdx.x == 0) {
    KS cur[3];
    cg_decide(it, nvc, t, rz, ring + ((it + 1) & 1) * 3, cur);
    for (int q = 0; q < nvc; ++q) out[q] = cur[q];
  }
}

// ---- BiCGStab (solv_bicg.f90; w_sc: res_n=1 alpha=2 beta0=3 beta=4 gam=5 omeg=6) ----
__device__ __forceinline__ void bi_decide(int it, int nvc, const double *L1, const double *rr0, const KS *prev,
                                          KS *cur) {
  double mx = 0.0;
  for (int q = 0; q < nvc; ++q) {
    KS k = prev[q];
    if (it == 0) {
      k.res0_n = L1[q];
      k.iter = 0;
      mx = q == 0 ? L1[q] : fmax(mx, L1[q]);
    } else if (!k.done) {
      k.w1 = L1[q] / (k.res0_n + SMALL);
      k.res_n = k.w1;
      k.iter = it;
      if (k.w1 < k.reduc || it >= k.n_iter) k.done = 1;
    }
    cur[q] = k;
  }
  if (it == 0)
    for (int q = 0; q < nvc; ++q)
      if (mx < SMALL || cur[q].n_iter <= 0) cur[q].done = 1;  // solv_bicg.f90:21
  for (int q = 0; q < nvc; ++q) {
    KS &k = cur[q];
    if (k.done) continue;
    k.w4 = rr0[q];
    k.w6 = k.w4 * k.w5 / (k.w2 * k.w3 + SMALL);
    k.w3 = k.w4;
  }
}

// Decide iteration it ; p = r + w6 (p - w2 u) ; Jacobi: z = c p
__global__ __launch_bounds__(RB64) void k_bi_p(int n, int it, int nvc, const double *__restrict__ r,
                                             double *__restrict__ p, const double *__restrict__ u,
                                             const double *__restrict__ c, int jacobi, double *__restrict__ z,
                                             const double *__restrict__ partU, int PU, KS *ring, KS *out) {
  double t[6];
  load_totals<6, RB64>(partU, PU, t);
  __shared__ KS sk[3];
  if (threadIdx.x == 0) {
    bi_decide(it, nvc, t, t + 3, ring + ((it + 1) & 1) * 3, sk);
    if (blockIdx.x == 0)
      for (int q = 0; q < nvc; ++q) ring[(it & 1) * 3 + q] = out[q] = sk[q];
  }
  __syncthreads();
  double w6[3] = {0, 0, 0}, w2[3] = {0, 0, 0};
  int dn[3] = {1, 1, 1};
  for (int q = 0; q < nvc; ++q) {
    w6[q] = sk[q].w6;
    w2[q] = sk[q].w2;
    dn[q] = sk[q].done;
  }
  int r0, r1;
  block_rows<RB64>(n, r0, r1);
  for (int i = r0 + threadIdx.x; i < r1; i += RB64)
    for (int q = 0; q < nvc; ++q) {
      if (dn[q]) continue;
      const int o = q * n + i;
      const double pi = r[o] + w6[q] * (p[o] - w2[q] * u[o]);
      p[o] = pi;
      if (jacobi) z[o] = pi * c[i];
    }
}

// y = A z with partials y.other (and y.y when mode = 1)
template <int WS>
__global__ __launch_bounds__(RB64) void k_bi_spmv(DMesh M, int it, int nvc, int mode, const double *__restrict__ diag,
                                                const double *__restrict__ spv, const double *__restrict__ z,
                                                double *__restrict__ y, const double *__restrict__ other,
                                                double *__restrict__ part, const KS *ring) {
  const KS *cur = ring + (it & 1) * 3;
  double a[6] = {0, 0, 0, 0, 0, 0};
  if (!all_done(cur, nvc)) {
    const int n = M.n;
    int r0, r1;
    block_rows<RB64>(n, r0, r1);
    for (int i = r0 + threadIdx.x; i < r1; i += RB64)
      for (int q = 0; q < nvc; ++q) {
        if (cur[q].done) continue;
        const double yi = row_spmv<WS>(M, diag, spv, z + q * n, i);
        y[q * n + i] = yi;
        a[q] += yi * other[q * n + i];
        if (mode) a[3 + q] += yi * yi;
      }
  }
  store_partials<6, RB64>(a, part);
}

// gam = beta/(u.r0 + small) ; x += gam z ; r -= gam u ; Jacobi: z = c r
__global__ __launch_bounds__(RB64) void k_bi_upd1(int n, int it, int nvc, double *__restrict__ x,
                                                double *__restrict__ r, double *__restrict__ z,
                                                const double *__restrict__ u, const double *__restrict__ c, int jacobi,
                                                const double *__restrict__ partB, int PB, KS *ring) {
  KS *cur = ring + (it & 1) * 3;
  if (all_done(cur, nvc)) return;
  double t[6];
  load_totals<6, RB64>(partB, PB, t);
  double w5[3];
  int dn[3];
  for (int q = 0; q < nvc; ++q) {
    dn[q] = cur[q].done;
    w5[q] = t[q];
    w5[q] = cur[q].w4 / (w5[q] + SMALL);
  }
  // every block has read cur[] above; only w5 is written, which no block of
  // this kernel reads
  if (blockIdx.x == 0 && threadIdx.x == 0)
    for (int q = 0; q < nvc; ++q)
      if (!dn[q]) cur[q].w5 = w5[q];
  int r0, r1;
  block_rows<RB64>(n, r0, r1);
  for (int i = r0 + threadIdx.x; i < r1; i += RB64)
    for (int q = 0; q < nvc; ++q) {
      if (dn[q]) continue;
      const int o = q * n + i;
      x[o] = x[o] + w5[q] * z[o];
      const double ri = r[o] - w5[q] * u[o];
      r[o] = ri;
      if (jacobi) z[o] = ri * c[i];
    }
}

// omega = (v.r)/(v.v + small) ; x += omega z ; r -= omega v ; partials |r|_1, r.r0
__global__ __launch_bounds__(RB64) void k_bi_upd2(int n, int it, int nvc, double *__restrict__ x,
                                                double *__restrict__ r, const double *__restrict__ z,
                                                const double *__restrict__ v, const double *__restrict__ r0v,
                                                const double *__restrict__ partD, int PD, double *__restrict__ partU,
                                                KS *ring) {
  KS *cur = ring + (it & 1) * 3;
  double a[6] = {0, 0, 0, 0, 0, 0};
  if (!all_done(cur, nvc)) {
    double t[6];
    load_totals<6, RB64>(partD, PD, t);
    double w1[3], w2[3];
    int dn[3];
    for (int q = 0; q < nvc; ++q) {
      dn[q] = cur[q].done;
      w1[q] = t[q];
      w2[q] = t[3 + q];
      w2[q] = w1[q] / (w2[q] + SMALL);
    }
    if (blockIdx.x == 0 && threadIdx.x == 0)
      for (int q = 0; q < nvc; ++q)
        if (!dn[q]) { cur[q].w1 = w1[q]; cur[q].w2 = w2[q]; }
    int r0, r1;
    block_rows<RB64>(n, r0, r1);
    for (int i = r0 + threadIdx.x; i < r1; i += RB64)
      for (int q = 0; q < nvc; ++q) {
        if (dn[q]) continue;
        const int o = q * n + i;
        x[o] = x[o] + w2[q] * z[o];
        const double ri = r[o] - w2[q] * v[o];
        r[o] = ri;
        a[q] += fabs(ri);
        a[3 + q] += ri * r0v[o];
      }
  }
  store_partials<6, RB64>(a, partU);
}

__global__ __launch_bounds__(RB64) void k_bi_final(int it, int nvc, const double *partU, int PU, KS *ring, KS *out) {
  double t[6];
  load_totals<6, RB64>(partU, PU, t);
  if (threadIdx.x == 0) {
    KS cur[3];
    bi_decide(it, nvc, t, t + 3, ring + ((it + 1) & 1) * 3, cur);
    for (int q = 0; q < nvc; ++q) out[q] = cur[q];
  }
}

// res0_n only (fvk_solv_res)
__global__ __launch_bounds__(RB64) void k_res_final(int nvc, const double *partU, int PU, KS *out) {
  double t[6];
  load_totals<6, RB64>(partU, PU, t);
  if (threadIdx.x == 0)
    for (int q = 0; q < nvc; ++q) out[q].res0_n = t[q];
}

// FP64 true residual {|b-Ax|_1, |b-Ax|_2^2, |b|_1, |b|_2^2} as partials
template <int WS>
__global__ __launch_bounds__(RB64) void k_true_res(DMesh M, const double *__restrict__ diag,
                                                 const double *__restrict__ spv, const double *__restrict__ x,
                                                 const double *__restrict__ b, double *__restrict__ part) {
  double a[4] = {0, 0, 0, 0};
  int r0, r1;
  block_rows<RB64>(M.n, r0, r1);
  for (int i = r0 + threadIdx.x; i < r1; i += RB64) {
    const double bi = b[i], ri = bi - row_spmv<WS>(M, diag, spv, x, i);
    a[0] += fabs(ri);
    a[1] += ri * ri;
    a[2] += fabs(bi);
    a[3] += bi * bi;
  }
  store_partials<4, RB64>(a, part);
}

__global__ __launch_bounds__(RB64) void k_totals4(const double *part, int P, double *out) {
  double s[4];
  load_totals<4, RB64>(part, P, s);
  if (threadIdx.x == 0)
    for (int j = 0; j < 4; ++j) out[j] = s[j];
}

// Fills and copies as kernels, so that a launch sequence can be captured
// into a HIP graph (no memset or host-to-device copy nodes).
__global__ __launch_bounds__(BS) void k_fill(size_t n, double *__restrict__ a, double v) {
  for (size_t i = blockIdx.x * (size_t)BS + threadIdx.x; i < n; i += (size_t)gridDim.x * BS) a[i] = v;
}
__global__ __launch_bounds__(BS) void k_copy_d(size_t n, const double *__restrict__ a, double *__restrict__ b) {
  for (size_t i = blockIdx.x * (size_t)BS + threadIdx.x; i < n; i += (size_t)gridDim.x * BS) b[i] = a[i];
}
// Both ring slots of the Krylov scalars: reduc, n_iter and the recurrences'
// initial values (solv_cg.f90:28 s = large; solv_bicg.f90:30 w_sc = 1)
__global__ void k_ks_init(KS *ring, int nvc, double reduc, int n_iter, int bicg) {
  const int t = threadIdx.x;
  if (t >= 6) return;
  KS k;
  k.res0_n = k.res_n = k.s = k.alpha = k.beta = 0.0;
  k.w1 = k.w2 = k.w3 = k.w4 = k.w5 = k.w6 = 0.0;
  k.reduc = 0.0;
  k.iter = k.n_iter = k.done = k.pad = 0;
  if (t % 3 < nvc) {
    k.reduc = reduc;
    k.n_iter = n_iter;
    if (bicg) k.w2 = k.w3 = k.w5 = 1.0;
    else k.s = 1.0 / SMALL;  // large (precision.f90:9)
  }
  ring[t] = k;
}

int fill(Ctx &c, double *a, size_t n, double v) {
  if (n == 0) return 0;
  const size_t g = std::min<size_t>((n + BS - 1) / BS, 4096);
  k_fill<<<(int)g, BS, 0, c.stream>>>(n, a, v);
  FVK_HIP(hipGetLastError());
  return 0;
}
int copy_d(Ctx &c, const double *a, double *b, size_t n) {
  if (n == 0) return 0;
  const size_t g = std::min<size_t>((n + BS - 1) / BS, 4096);
  k_copy_d<<<(int)g, BS, 0, c.stream>>>(n, a, b);
  FVK_HIP(hipGetLastError());
  return 0;
}

// ------------------------------------------------------------ host side

static int mark_grid(size_t cnt) { return (int)std::min<size_t>((cnt + BS - 1) / BS, 4096); }
// resident grid of the one-launch kernels: sweep_blocks_per_cu blocks per CU,
// no more than there are chunks
static int sweep_grid(const Ctx &c) {
  const int chunks = std::max(1, (c.n + BS - 1) / BS);
  return std::max(1, std::min(chunks, c.n_cu * c.sweep_blocks_per_cu));
}

int sweep_masks_build(Ctx &c) {
  k_sweep_masks<<<grid_for(c.n), BS, 0, c.stream>>>(c.M, c.sweep_lmask, c.sweep_umask);
  FVK_HIP(hipGetLastError());
  return 0;
}

static void free_owned(Ctx &c, void *p) {
  if (!p) return;
  for (size_t k = 0; k < c.owned.size(); ++k)
    if (c.owned[k] == p) {
      hipFree(p);
      c.owned.erase(c.owned.begin() + (long)k);
      return;
    }
}

template <class T>
static int upload_v(Ctx &c, T **dst, const std::vector<T> &src) {
  int rc = dalloc(c, dst, src.size());
  if (rc) return rc;
  if (!src.empty()) FVK_HIP(hipMemcpy(*dst, src.data(), src.size() * sizeof(T), hipMemcpyHostToDevice));
  return 0;
}

// Ranges of R rows consecutive in the reference numbering (the order of
// the sweeps: forward ascending, backward descending) and, in each, its
// rows by level (ties by rank), with the descriptors of sweep_parts.
static int parts_build(Ctx &c, bool fwd, const std::vector<int> &rank, const std::vector<int> &spc,
                       const std::vector<unsigned char> &mask, Ctx::SweepParts &pt) {
  const int n = c.n, R = c.sweep_part_rows, Ws = c.M.Ws;
  const size_t ld = (size_t)c.M.ld;
  std::vector<int> row_of(n);
  for (int r = 0; r < n; ++r) row_of[rank[r]] = r;
  // sweep position of each row (0 = first reached) and its level
  auto pos_of = [&](int r) { return fwd ? rank[r] : n - 1 - rank[r]; };
  std::vector<int> lev(n, 0);
  for (int k = 0; k < n; ++k) {
    const int r = row_of[fwd ? k : n - 1 - k];
    int l = 0;
    for (int s = 0; s < Ws; ++s)
      if ((mask[r] >> s) & 1u) l = std::max(l, lev[spc[(size_t)s * ld + r]] + 1);
    lev[r] = l;
  }
  pt.R = R;
  pt.P = (n + R - 1) / R;
  std::vector<int> rows, first(n), cnt(n), aux(n, 0), lrow(n), step, sp;
  rows.reserve(n);
  std::vector<std::pair<int, int>> buf;
  for (int q = 0; q < pt.P; ++q) {
    buf.clear();
    for (int k = q * R; k < std::min(n, (q + 1) * R); ++k) {
      const int r = row_of[fwd ? k : n - 1 - k];
      buf.emplace_back(lev[r], k);
    }
    std::sort(buf.begin(), buf.end());
    sp.push_back((int)step.size());
    for (size_t t = 0; t < buf.size(); ++t) {
      if (t == 0 || buf[t].first != buf[t - 1].first) step.push_back((int)rows.size());
      rows.push_back(row_of[fwd ? buf[t].second : n - 1 - buf[t].second]);
    }
    step.push_back((int)rows.size());
  }
  sp.push_back((int)step.size());
  pt.wmax = 0;
  for (int q = 0; q < pt.P; ++q)
    for (int k = sp[q]; k + 1 < sp[q + 1]; ++k) pt.wmax = std::max(pt.wmax, step[k + 1] - step[k]);
  int W = 1;
  for (int p = 0; p < n; ++p) {
    first[p] = mask[rows[p]];
    cnt[p] = __builtin_popcount(mask[rows[p]]);
    W = std::max(W, cnt[p]);
  }
  pt.W = W;
  std::vector<int> cols((size_t)W * n, 0), lcol((size_t)W * n, -1);
  for (int p = 0; p < n; ++p) {
    const int r = rows[p], q = pos_of(r) / R;
    lrow[p] = pos_of(r) - q * R;
    int t = 0;
    for (int s = 0; s < Ws; ++s)
      if ((mask[r] >> s) & 1u) {
        const int j = spc[(size_t)s * ld + r];
        cols[(size_t)t * n + p] = j;
        if (pos_of(j) / R == q) lcol[(size_t)t * n + p] = pos_of(j) - q * R;
        ++t;
      }
  }
  std::vector<int> hd, jc, lc;
  sweep_pack(n, W, rows, first, cnt, aux, lrow, cols, lcol, hd, jc, lc);
  int rc = upload_v(c, &pt.hd, hd);
  if (!rc) rc = upload_v(c, &pt.jc, jc);
  if (!rc) rc = upload_v(c, &pt.lc, lc);
  if (!rc) rc = upload_v(c, &pt.step, step);
  if (!rc) rc = upload_v(c, &pt.sp, sp);
  return rc;
}

int sweep_parts_build(Ctx &c) {
  for (Ctx::SweepParts *pt : {&c.spf, &c.spb}) {
    for (int *q : {pt->hd, pt->jc, pt->lc, pt->step, pt->sp}) free_owned(c, q);
    *pt = Ctx::SweepParts();
  }
  if (!c.sweep_one_launch) return 0;
  const int n = c.n;
  const size_t ld = (size_t)c.M.ld;
  std::vector<int> rank(n), spc((size_t)c.M.Ws * ld);
  std::vector<unsigned char> lm(n), um(n);
  FVK_HIP(hipStreamSynchronize(c.stream));
  FVK_HIP(hipMemcpy(rank.data(), c.M.rank, (size_t)n * sizeof(int), hipMemcpyDeviceToHost));
  FVK_HIP(hipMemcpy(spc.data(), c.M.spc, spc.size() * sizeof(int), hipMemcpyDeviceToHost));
  FVK_HIP(hipMemcpy(lm.data(), c.sweep_lmask, n, hipMemcpyDeviceToHost));
  FVK_HIP(hipMemcpy(um.data(), c.sweep_umask, n, hipMemcpyDeviceToHost));
  int rc = parts_build(c, true, rank, spc, lm, c.spf);
  if (!rc) rc = parts_build(c, false, rank, spc, um, c.spb);
  if (rc) return rc;
  const int bytes = 3 * c.sweep_part_rows * 8;
  if (bytes > 65536) {
    const void *k[] = {
#define K_(D, V, MW) (const void *)k_sweep_pt<D, V, MW, 64>, (const void *)k_sweep_pt<D, V, MW, BS>
#define K2_(V, MW) K_(true, V, MW), K_(false, V, MW)
        K2_(1, 4), K2_(2, 4), K2_(3, 4), K2_(1, 8), K2_(2, 8), K2_(3, 8)
#undef K2_
#undef K_
    };
    for (const void *f : k) FVK_HIP(hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, bytes));
  }
  return 0;
}

static SweepTabs tabs_of(const Ctx::SweepParts &p) {
  return SweepTabs{(const int4 *)p.hd, (const int4 *)p.jc, (const int4 *)p.lc};
}
// blocks of the partitioned kernels: as many per CU as the LDS of a block
// lets reside, at most four of 64 threads or two of 256 (the ranges a grid
// cannot hold at once wait for a block to finish)
static int pt_grid(const Ctx &c, const Ctx::SweepParts &p, size_t lds, int tb) {
  const int per_cu = std::max(1, std::min<int>(tb == 64 ? 4 : 2, (int)(163840 / (lds + 1024))));
  return std::max(1, std::min(p.P, c.n_cu * per_cu));
}

// one partitioned sweep: one wave a block when no level of a range has more
// than 64 rows
template <bool FWD, int V, int MW>
static void sweep_pt_launch(Ctx &c, const Ctx::SweepParts &X, size_t lds, const double *cr, const double *src,
                            double *out, int *ticket, const KS *gate) {
  const size_t ld = (size_t)c.M.ld;
  if (X.wmax <= 64)
    k_sweep_pt<FWD, V, MW, 64><<<pt_grid(c, X, lds, 64), 64, lds, c.stream>>>(
        c.n, X.R, X.P, tabs_of(X), X.step, X.sp, ld, c.spv, c.pc_c, cr, src, out, ticket, c.sweep_int + 3, gate);
  else
    k_sweep_pt<FWD, V, MW, BS><<<pt_grid(c, X, lds, BS), BS, lds, c.stream>>>(
        c.n, X.R, X.P, tabs_of(X), X.step, X.sp, ld, c.spv, c.pc_c, cr, src, out, ticket, c.sweep_int + 3, gate);
}

template <bool FWD>
static void sweep_pt_any(Ctx &c, int nvc, const Ctx::SweepParts &X, size_t lds, const double *cr, const double *src,
                         double *out, int *ticket, const KS *gate) {
  if (X.W <= 4) {
    if (nvc == 1) sweep_pt_launch<FWD, 1, 4>(c, X, lds, cr, src, out, ticket, gate);
    else if (nvc == 2) sweep_pt_launch<FWD, 2, 4>(c, X, lds, cr, src, out, ticket, gate);
    else sweep_pt_launch<FWD, 3, 4>(c, X, lds, cr, src, out, ticket, gate);
  } else {
    if (nvc == 1) sweep_pt_launch<FWD, 1, 8>(c, X, lds, cr, src, out, ticket, gate);
    else if (nvc == 2) sweep_pt_launch<FWD, 2, 8>(c, X, lds, cr, src, out, ticket, gate);
    else sweep_pt_launch<FWD, 3, 8>(c, X, lds, cr, src, out, ticket, gate);
  }
}

int precond_build(Ctx &c, int imat) {
  const int n = c.n;
  if (imat == 0) {
    k_precond_jacobi<<<grid_for(n), BS, 0, c.stream>>>(n, c.diag, c.pc_c, c.pc_cr);
  } else if (c.sweep_one_launch && c.sweep_parts) {
    const Ctx::SweepParts &P = c.spf;
    const size_t lds = (size_t)P.R * 8;
    k_sweep_mark<<<mark_grid(n), BS, 0, c.stream>>>(n, c.pc_c, c.sweep_int + 2, nullptr, 1);
#define PC_(MW, TB) k_precond_pt<MW, TB><<<pt_grid(c, P, lds, TB), TB, lds, c.stream>>>(n, P.R, P.P, tabs_of(P), \
    P.step, P.sp, c.M, imat, c.diag, c.ngp, c.ngn, c.pc_c, c.pc_cr, c.sweep_int + 2, c.sweep_int + 3)
    if (P.W <= 4) {
      if (P.wmax <= 64) PC_(4, 64); else PC_(4, BS);
    } else {
      if (P.wmax <= 64) PC_(8, 64); else PC_(8, BS);
    }
#undef PC_
  } else if (c.sweep_one_launch) {
    k_sweep_mark<<<mark_grid(n), BS, 0, c.stream>>>(n, c.pc_c, c.sweep_int + 2, nullptr, 1);
    k_precond_sf<<<sweep_grid(c), BS, 0, c.stream>>>(c.M, c.lf.rows, c.sweep_lmask, imat, c.diag, c.ngp, c.ngn,
                                                       c.pc_c, c.pc_cr, c.sweep_int + 2, c.sweep_int + 3);
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
  } else if (c.sweep_one_launch && c.sweep_parts) {
    const size_t cnt = (size_t)nvc * n;
    const size_t lds = (size_t)nvc * c.sweep_part_rows * 8;
    k_sweep_mark<<<mark_grid(cnt), BS, 0, c.stream>>>(cnt, c.sweep_t, c.sweep_int, gate, nvc);
    sweep_pt_any<true>(c, nvc, c.spf, lds, nullptr, b, c.sweep_t, c.sweep_int, gate);
    k_sweep_mark<<<mark_grid(cnt), BS, 0, c.stream>>>(cnt, x, c.sweep_int + 1, gate, nvc);
    sweep_pt_any<false>(c, nvc, c.spb, lds, c.pc_cr, c.sweep_t, x, c.sweep_int + 1, gate);
  } else if (c.sweep_one_launch) {
    const size_t cnt = (size_t)nvc * n;
    k_sweep_mark<<<mark_grid(cnt), BS, 0, c.stream>>>(cnt, c.sweep_t, c.sweep_int, gate, nvc);
    const int G = sweep_grid(c);
#define FWD(V) k_sweep_fwd_sf<V><<<G, BS, 0, c.stream>>>(c.M, c.lf.rows, c.sweep_lmask, c.spv, c.pc_c, b, \
                                                         c.sweep_t, c.sweep_int, c.sweep_int + 3, gate)
#define BWD(V) k_sweep_bwd_sf<V><<<G, BS, 0, c.stream>>>(c.M, c.lb.rows, c.sweep_umask, c.spv, c.pc_c, c.pc_cr, \
                                                         c.sweep_t, x, c.sweep_int + 1, c.sweep_int + 3, gate)
    if (nvc == 1) FWD(1); else if (nvc == 2) FWD(2); else FWD(3);
    k_sweep_mark<<<mark_grid(cnt), BS, 0, c.stream>>>(cnt, x, c.sweep_int + 1, gate, nvc);
    if (nvc == 1) BWD(1); else if (nvc == 2) BWD(2); else BWD(3);
#undef FWD
#undef BWD
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

// Launch F(WS) with the slot count of the mesh's table as a template
// argument where an instantiation exists.
#define FVK_WS(ws, F) \
  switch (ws) {       \
    case 3: F(3); break; \
    case 4: F(4); break; \
    case 6: F(6); break; \
    default: F(0); break; \
  }

int spmv(Ctx &c, int nvc, const double *x, double *y) {
#define L(W) k_spmv<W><<<grid_for(c.n), BS, 0, c.stream>>>(c.M, nvc, c.diag, c.spv, x, y)
  FVK_WS(c.M.Ws, L);
#undef L
  FVK_HIP(hipGetLastError());
  return 0;
}

static inline double *wv(Ctx &c, int k) { return wk_slot(c, k); }
static inline double *pbuf(Ctx &c, int k) { return c.part + (size_t)k * RG * 8; }

static int ks_init(Ctx &c, int nvc, double reduc, int n_iter, bool bicg) {
  k_ks_init<<<1, 64, 0, c.stream>>>(c.ksr, nvc, reduc, n_iter, bicg ? 1 : 0);
  FVK_HIP(hipGetLastError());
  return 0;
}

int residual(Ctx &c, int nvc, const double *x, double *r, bool not_r0) {
  const int G = red_grid64(c.n);
#define L(W) \
  k_residual<W><<<G, RB64, 0, c.stream>>>(c.M, nvc, c.diag, c.spv, x, r, not_r0 ? 1 : 0, 0, nullptr, nullptr, pbuf(c, 0))
  FVK_WS(c.M.Ws, L);
#undef L
  k_res_final<<<1, RB64, 0, c.stream>>>(nvc, pbuf(c, 0), G, c.ks);
  FVK_HIP(hipGetLastError());
  return 0;
}

__global__ __launch_bounds__(RB64) void k_res_final_to(int nvc, const double *partU, int PU, double *out) {
  double t[6];
  load_totals<6, RB64>(partU, PU, t);
  if (threadIdx.x == 0)
    for (int q = 0; q < nvc; ++q) out[q] = t[q];
}

// residual() with the per-block partials in `part` (6 RG doubles) and the
// L1 norms written to the device array out_dev (nvc doubles): nothing
// outside the inputs, the outputs and `part` is touched, and nothing waits
// for the device.
int residual_to(Ctx &c, int nvc, const double *x, double *r, bool not_r0, double *part, double *out_dev) {
  const int G = red_grid64(c.n);
  if (G > RG) {
    set_error("partials exceed the scratch");
    return FVK_ERR_ARG;
  }
#define L(W) \
  k_residual<W><<<G, RB64, 0, c.stream>>>(c.M, nvc, c.diag, c.spv, x, r, not_r0 ? 1 : 0, 0, nullptr, nullptr, part)
  FVK_WS(c.M.Ws, L);
#undef L
  k_res_final_to<<<1, RB64, 0, c.stream>>>(nvc, part, G, out_dev);
  FVK_HIP(hipGetLastError());
  return 0;
}

int sweep_status(Ctx &c) {
  if (!c.sweep_int) return 0;
  int e = 0;
  FVK_HIP(hipMemcpyAsync(&e, c.sweep_int + 3, sizeof e, hipMemcpyDeviceToHost, c.stream));
  FVK_HIP(hipStreamSynchronize(c.stream));
  if (e) {
    set_error("triangular sweep: a wait exceeded its bound; the preconditioner's results are invalid");
    return FVK_ERR_HIP;
  }
  return 0;
}

static int all_done_host(Ctx &c, int nvc, bool &done) {
  KS h[3];
  FVK_HIP(hipMemcpyAsync(h, c.ks, sizeof(KS) * nvc, hipMemcpyDeviceToHost, c.stream));
  FVK_HIP(hipStreamSynchronize(c.stream));
  done = true;
  for (int q = 0; q < nvc; ++q) done = done && h[q].done;
  return 0;
}

template <int NVC>
static void launch_cg_pq(Ctx &c, int G, int it, double *z, double *pold, double *pnew, double *q, const double *pU,
                         const double *pR, KTimer *T) {
  hipEvent_t a, b;
  if (T && T->take(a, b)) {
    const double *zc = z, *pc = pold;
#define L(W)                                                                                                   \
  hipExtLaunchKernelGGL(k_cg_pq<NVC, W>, dim3(G), dim3(RB64), 0, c.stream, a, b, 0, c.M, it, (const double *)c.diag, \
                        (const double *)c.spv, zc, pc, pnew, q, pU, G, pR, G, pbuf(c, 1), c.ksr, c.ks)
    FVK_WS(c.M.Ws, L);
#undef L
    return;
  }
#define L(W)                                                                                                    \
  k_cg_pq<NVC, W><<<G, RB64, 0, c.stream>>>(c.M, it, c.diag, c.spv, z, pold, pnew, q, pU, G, pR, G, pbuf(c, 1), \
                                          c.ksr, c.ks)
  FVK_WS(c.M.Ws, L);
#undef L
}
template <int NVC>
static void launch_cg_upd(Ctx &c, int G, int it, double *x, double *r, double *p, double *q, int jac, double *z) {
  k_cg_upd<NVC><<<G, RB64, 0, c.stream>>>(c.n, it, x, r, p, q, c.pc_c, jac, z, pbuf(c, 1), G, pbuf(c, 0), c.ksr);
}

// PCG: r0 holds b on entry, b - A x0 (the initial residual) on exit; x is
// updated.  chunk > 0: stop queueing iterations once converged.
static int cg_run(Ctx &c, int imat, double reduc, int n_iter, int nvc, double *x, double *r0, bool not_r0, int chunk,
                  KTimer *T) {
  const int n = c.n, G = red_grid64(n);
  double *r = wv(c, 0), *z = wv(c, 1), *P = wv(c, 2), *q = wv(c, 4);
  const int jac = imat == 0;
  if (!c.pc_valid || c.pc_imat != imat) {
    int rc = precond_build(c, imat);
    if (rc) return rc;
  }
  // both direction buffers (slots 2 and 3)
  if (fill(c, P, wk_ld(n) + (size_t)3 * n, 0.0)) return FVK_ERR_HIP;
#define L(W)                                                                                                \
  k_residual<W><<<G, RB64, 0, c.stream>>>(c.M, nvc, c.diag, c.spv, x, r0, not_r0 ? 1 : 0, jac ? 1 : 0, c.pc_c, z, \
                                        pbuf(c, 0))
  FVK_WS(c.M.Ws, L);
#undef L
  if (copy_d(c, r0, r, (size_t)nvc * n)) return FVK_ERR_HIP;
  const double *pR = nullptr;
  if (!jac) {
    int rc = precond_apply(c, imat, nvc, r, z, nullptr);
    if (rc) return rc;
    k_dot<<<G, RB64, 0, c.stream>>>(n, nvc, r, z, pbuf(c, 2), nullptr);
    pR = pbuf(c, 2);
  }
  Poller Pl(c);
  int it = 0, next_poll = chunk;
  for (; it < n_iter; ++it) {
    double *pold = P + (size_t)(it & 1) * wk_ld(n), *pnew = P + (size_t)((it + 1) & 1) * wk_ld(n);
    if (nvc == 1) launch_cg_pq<1>(c, G, it, z, pold, pnew, q, pbuf(c, 0), pR, T);
    else if (nvc == 2) launch_cg_pq<2>(c, G, it, z, pold, pnew, q, pbuf(c, 0), pR, T);
    else launch_cg_pq<3>(c, G, it, z, pold, pnew, q, pbuf(c, 0), pR, T);
    if (nvc == 1) launch_cg_upd<1>(c, G, it, x, r, pnew, q, jac, z);
    else if (nvc == 2) launch_cg_upd<2>(c, G, it, x, r, pnew, q, jac, z);
    else launch_cg_upd<3>(c, G, it, x, r, pnew, q, jac, z);
    if (!jac) {
      int rc = precond_apply(c, imat, nvc, r, z, c.ksr + (it & 1) * 3);
      if (rc) return rc;
      k_dot<<<G, RB64, 0, c.stream>>>(n, nvc, r, z, pbuf(c, 2), c.ksr + (it & 1) * 3);
    }
    if (chunk > 0 && it + 1 == next_poll && it + 1 < n_iter) {
      bool done;
      int next = chunk;
      if (nvc == 1 && !T) {
        // the next chunk is queued before the host waits for this one
        FVK_HIP(Pl.step(it + 1, chunk, reduc, done, next));
      } else {
        if (T && T->flush()) return FVK_ERR_HIP;
        if (all_done_host(c, nvc, done)) return FVK_ERR_HIP;
      }
      if (done) { ++it; break; }
      next_poll = it + 1 + next;
    }
  }
  k_cg_final<<<1, RB64, 0, c.stream>>>(it, nvc, pbuf(c, 0), G, pR, G, c.ksr, c.ks);
  FVK_HIP(hipGetLastError());
  if (T && T->flush()) return FVK_ERR_HIP;
  return 0;
}

static int bicg_run(Ctx &c, int imat, int n_iter, int nvc, double *x, double *r0, bool not_r0) {
  const int n = c.n, G = red_grid64(n);
  double *r = wv(c, 0), *p = wv(c, 1), *u = wv(c, 2), *v = wv(c, 3), *z = wv(c, 4);
  const int jac = imat == 0;
  if (!c.pc_valid || c.pc_imat != imat) {
    int rc = precond_build(c, imat);
    if (rc) return rc;
  }
  // slots 1-4
  if (fill(c, p, (size_t)3 * wk_ld(n) + (size_t)3 * n, 0.0)) return FVK_ERR_HIP;
#define L(W)                                                                                              \
  k_residual<W><<<G, RB64, 0, c.stream>>>(c.M, nvc, c.diag, c.spv, x, r0, not_r0 ? 1 : 0, 2, nullptr, nullptr, \
                                        pbuf(c, 0))
  FVK_WS(c.M.Ws, L);
#undef L
  if (copy_d(c, r0, r, (size_t)nvc * n)) return FVK_ERR_HIP;
  int it = 0;
  for (; it < n_iter; ++it) {
    const KS *gate = c.ksr + (it & 1) * 3;
    k_bi_p<<<G, RB64, 0, c.stream>>>(n, it, nvc, r, p, u, c.pc_c, jac, z, pbuf(c, 0), G, c.ksr, c.ks);
    if (!jac && precond_apply(c, imat, nvc, p, z, gate)) return FVK_ERR_HIP;
#define L(W) k_bi_spmv<W><<<G, RB64, 0, c.stream>>>(c.M, it, nvc, 0, c.diag, c.spv, z, u, r0, pbuf(c, 1), c.ksr)
    FVK_WS(c.M.Ws, L);
#undef L
    k_bi_upd1<<<G, RB64, 0, c.stream>>>(n, it, nvc, x, r, z, u, c.pc_c, jac, pbuf(c, 1), G, c.ksr);
    if (!jac && precond_apply(c, imat, nvc, r, z, gate)) return FVK_ERR_HIP;
#define L(W) k_bi_spmv<W><<<G, RB64, 0, c.stream>>>(c.M, it, nvc, 1, c.diag, c.spv, z, v, r, pbuf(c, 2), c.ksr)
    FVK_WS(c.M.Ws, L);
#undef L
    k_bi_upd2<<<G, RB64, 0, c.stream>>>(n, it, nvc, x, r, z, v, r0, pbuf(c, 2), G, pbuf(c, 0), c.ksr);
  }
  k_bi_final<<<1, RB64, 0, c.stream>>>(it, nvc, pbuf(c, 0), G, c.ksr, c.ks);
  FVK_HIP(hipGetLastError());
  return 0;
}

int krylov(Ctx &c, int kind, int imat, double reduc, int n_iter, int nvc, double *x, double *r0, bool not_r0,
           int chunk, KTimer *T) {
  int rc = ks_init(c, nvc, reduc, n_iter, kind == 1);
  if (rc) return rc;
  if (kind == 0) return cg_run(c, imat, reduc, n_iter, nvc, x, r0, not_r0, chunk, T);
  return bicg_run(c, imat, n_iter, nvc, x, r0, not_r0);
}

__global__ __launch_bounds__(BS) void k_scale64(size_t n, double *__restrict__ a, const double *__restrict__ b,
                                                double s) {
  for (size_t i = blockIdx.x * (size_t)BS + threadIdx.x; i < n; i += (size_t)gridDim.x * BS) a[i] = s * b[i];
}

int pcg_tol(Ctx &c, int imat, const double *b, double *x, double tol, int max_iter, int chunk, KTimer *T,
            KS &out) {
  const int n = c.n;
  int rc;
  if ((rc = fill(c, x, n, 0.0))) return rc;
  double tr[4];
  if ((rc = true_residual(c, x, b, c.stage, tr))) return rc;
  const double l1b = tr[2], sigma = l1b > 0 ? l1b / n : 1.0;
  if (!std::isfinite(l1b)) {
    set_error("non-finite right-hand side");
    return FVK_ERR_ARG;
  }
  double *bs = wk_slot(c, 7);
  const int g = (int)std::min<size_t>(((size_t)n + BS - 1) / BS, 4096);
  k_scale64<<<g, BS, 0, c.stream>>>(n, bs, b, 1.0 / sigma);
  if ((rc = krylov(c, 0, imat, tol, max_iter, 1, x, bs, true, chunk, T))) return rc;
  k_scale64<<<g, BS, 0, c.stream>>>(n, x, x, sigma);
  FVK_HIP(hipGetLastError());
  FVK_HIP(hipMemcpyAsync(&out, c.ks, sizeof out, hipMemcpyDeviceToHost, c.stream));
  FVK_HIP(hipStreamSynchronize(c.stream));
  // the caller sees the norms of its own right-hand side
  out.res0_n = l1b;
  FVK_HIP(hipMemcpy(c.ks, &out, sizeof out, hipMemcpyHostToDevice));
  return 0;
}

int true_residual(Ctx &c, const double *x, const double *b, double *dev_out4, double *host_out4) {
  const int G = red_grid64(c.n);
#define L(W) k_true_res<W><<<G, RB64, 0, c.stream>>>(c.M, c.diag, c.spv, x, b, pbuf(c, 3))
  FVK_WS(c.M.Ws, L);
#undef L
  k_totals4<<<1, RB64, 0, c.stream>>>(pbuf(c, 3), G, dev_out4);
  FVK_HIP(hipGetLastError());
  FVK_HIP(hipMemcpyAsync(host_out4, dev_out4, 4 * sizeof(double), hipMemcpyDeviceToHost, c.stream));
  FVK_HIP(hipStreamSynchronize(c.stream));
  return 0;
}

}  // namespace fvk
