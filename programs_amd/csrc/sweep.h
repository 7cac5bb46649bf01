// Device helpers of the one-launch triangular sweeps (krylov.hip: the
// spray-simulation factor on the slot tables; flowlab.hip: flowlab's factor
// on CSR rows), in two forms: rows taken in level order in chunks
// (sweep_chunks) and ranges of rows walked level by level by one block each
// (sweep_parts, the default).  In both, work is taken from a ticket counter,
// so everything a block waits for was taken earlier by a block that is
// running or has finished.  An output starts as kSweepMark (a signalling
// NaN: no arithmetic result has this pattern) and a row waits until each
// value it needs from another block has been replaced.  Every wait is
// bounded (kSweepSpin passes); past the bound the kernel records an error
// and its blocks stop taking work.
//
// sweep_chunks: chunks of BS positions of the level order; waiting lanes
// never block their wave (the wave loops until all its lanes are finished).
#pragma once
#include <vector>

#include "dev.h"

namespace fvk {

constexpr unsigned long long kSweepMark = 0x7FF4DEADBEEF0001ull;
constexpr unsigned kSweepSpin = 1u << 20;

__device__ __forceinline__ unsigned long long ld_agent(const double *p) {
  return __hip_atomic_load((const unsigned long long *)p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st_agent(double *p, double v) {
  __hip_atomic_store((unsigned long long *)p, (unsigned long long)__double_as_longlong(v), __ATOMIC_RELAXED,
                     __HIP_MEMORY_SCOPE_AGENT);
}

// The chunk loop: row.start(pos) loads the row at level-order position pos,
// row.advance() takes the values that have arrived (in the row's order) and
// returns true once the row's result is written.
template <class Row>
__device__ __forceinline__ void sweep_chunks(int n, int *ticket, int *err, Row &row) {
  __shared__ int s_chunk, s_abort;
  if (threadIdx.x == 0) s_abort = 0;
  for (;;) {
    if (threadIdx.x == 0) s_chunk = atomicAdd(ticket, 1);
    __syncthreads();
    const int chunk = s_chunk;
    const int abort = s_abort;
    __syncthreads();
    if (abort || (long)chunk * BS >= (long)n) return;
    const int pos = chunk * BS + threadIdx.x;
    bool fin = pos >= n;
    if (!fin) row.start(pos);
    unsigned spin = 0, idle = 0;
    while (!__all(fin)) {
      bool moved = false;
      if (!fin) {
        fin = row.advance();
        moved = fin;
      }
      if (++spin > kSweepSpin) {
        if (!fin) { *err = 1; s_abort = 1; }
        break;
      }
      idle = __any(moved) ? 0 : idle + 1;
      if (!__all(fin)) {
        if (idle < 8) __builtin_amdgcn_s_sleep(1);
        else __builtin_amdgcn_s_sleep(8);
      }
    }
    __syncthreads();
  }
}

// at most this many lower (upper) entries per row are kept in registers
constexpr int kSweepSlots = 8;

// Partitioned sweep.  The rows are cut into P ranges of R rows consecutive
// in the sweep's order (range q holds the rows the sweep reaches q-th); a
// block takes the ranges in ticket order (so a range waits only for ranges
// taken earlier) and walks its rows level by level, a barrier between
// levels.  Values of rows inside the range pass through LDS
// (NC x R doubles); values from other ranges are read from `out` with
// agent-scope loads until they are no longer the mark.  Range q's levels
// are positions [step[k], step[k+1]) for k in [sp[q], sp[q+1]-1).
//
// Each position p has a descriptor: its row, a row-specific first-entry
// word, the entry count, one more row-specific index, the row's own LDS
// slot (its index in the range), and per entry the column and the column's
// LDS slot (-1: another range), packed in 16-byte words (SweepTabs).  The
// first TB positions of each level are software-pipelined (descriptors four
// levels ahead, the row's coefficients two and its values from other ranges
// one where three stages fit in registers, else two, one and one), all
// loads issued without branches (a lane past the level's end loads the
// level's last position and is not run), so a level waits for memory only
// when a value from another range has not yet arrived.
//
// Row interface: fetch(d, MW) loads the row's coefficients (entries t < MW,
// entries t >= d.nl from a valid address, unused); begin() does the
// arithmetic on them that precedes the entries; take(t, v) adds entry t with
// the NC values v[] (called in entry order, t < d.nl); result(c) is
// component c of the row's value; done(d) runs after the row's value is
// stored.
// Descriptor tables of n positions: hd[p] = {row, first-entry word,
// count | LDS slot << 4, aux}; for each group g of four entry slots,
// jc[g n + p] the columns and lc[g n + p] their LDS slots (-1: another
// range).  Slots past the row's count hold the row and 0 (never read).
struct SweepTabs {
  const int4 *hd, *jc, *lc;
};

// Packs per-position arrays (cols / lcol slot-major, W slots of n) into the
// int words of hd (4 n) and of jc, lc (4 n per group of four slots, MW / 4
// groups: MW = 4 when W <= 4, else 8).
inline void sweep_pack(int n, int W, const std::vector<int> &rows, const std::vector<int> &first,
                       const std::vector<int> &cnt, const std::vector<int> &aux, const std::vector<int> &lrow,
                       const std::vector<int> &cols, const std::vector<int> &lcol, std::vector<int> &hd,
                       std::vector<int> &jc, std::vector<int> &lc) {
  const int groups = W <= 4 ? 1 : 2;
  hd.assign((size_t)4 * n, 0);
  jc.assign((size_t)4 * groups * n, 0);
  lc.assign((size_t)4 * groups * n, 0);
  for (int p = 0; p < n; ++p) {
    hd[(size_t)4 * p] = rows[p];
    hd[(size_t)4 * p + 1] = first[p];
    hd[(size_t)4 * p + 2] = cnt[p] | lrow[p] << 4;
    hd[(size_t)4 * p + 3] = aux[p];
    for (int t = 0; t < 4 * groups; ++t) {
      const size_t o = ((size_t)(t / 4) * n + p) * 4 + t % 4;
      jc[o] = t < cnt[p] ? cols[(size_t)t * n + p] : rows[p];
      lc[o] = t < cnt[p] ? lcol[(size_t)t * n + p] : 0;
    }
  }
}

struct SweepDesc {
  int r, first, nl, aux, lr;
  int j[kSweepSlots], li[kSweepSlots];
};

// read-only tables with a uniform index: scalar loads, which do not wait
// behind the vector loads in flight
typedef const __attribute__((address_space(4))) int *const_int_p;
__device__ __forceinline__ int ld_uniform(const int *p, int i) { return ((const_int_p)p)[i]; }

// descriptor of position p of the level [pb, pe) (p >= pe: a copy of the
// level's last position with no entries, not run): 1 + 2 MW / 4 loads of
// 16 bytes
template <int MW>
__device__ __forceinline__ void sweep_desc(const SweepTabs &T, int n, int p, int pe, SweepDesc &d) {
  const bool live = p < pe;
  const int q = live ? p : pe - 1;
  const int4 h = T.hd[q];
  d.r = h.x;
  d.first = h.y;
  d.nl = live ? (h.z & 15) : -1;
  d.lr = h.z >> 4;
  d.aux = h.w;
#pragma unroll
  for (int g = 0; g < MW / 4; ++g) {
    const int4 J = T.jc[(size_t)g * n + q], L = T.lc[(size_t)g * n + q];
    d.j[4 * g] = J.x;
    d.j[4 * g + 1] = J.y;
    d.j[4 * g + 2] = J.z;
    d.j[4 * g + 3] = J.w;
    d.li[4 * g] = L.x;
    d.li[4 * g + 1] = L.y;
    d.li[4 * g + 2] = L.z;
    d.li[4 * g + 3] = L.w;
  }
}

// values of entries in other ranges: one agent-scope load per entry slot
// t < MW (slots that need none load the row's own value, unused)
template <int NC, int MW>
__device__ __forceinline__ void sweep_ext(const SweepDesc &d, int n, const double *out,
                                          unsigned long long (&v)[kSweepSlots][NC]) {
#pragma unroll
  for (int t = 0; t < MW; ++t) {
    const bool ext = t < d.nl && d.li[t] < 0;
    const int jj = ext ? d.j[t] : d.r;
#pragma unroll
    for (int c = 0; c < NC; ++c) v[t][c] = ld_agent(out + (size_t)c * n + jj);
  }
}

// run one row (d.nl >= 0): values inside the range from LDS, the others from
// v[] (polled again while they are the mark); false if the wait bound ran
// out.  Every slot t < MW is read and taken, and a slot past the row's count
// is dropped by a select, so that the entries cost no branches.
template <int NC, int MW, class Row>
__device__ __forceinline__ bool sweep_row(const SweepDesc &d, Row &row,
                                          unsigned long long (&v)[kSweepSlots][NC], int R, int n, double *out,
                                          double *lv, int *err) {
  row.begin();
  unsigned need = 0;
#pragma unroll
  for (int t = 0; t < MW; ++t) {
    const int l = d.li[t];
    const bool loc = l >= 0;
    const int la = loc ? l : 0;
    bool mk = false;
#pragma unroll
    for (int c = 0; c < NC; ++c) {
      const unsigned long long x = (unsigned long long)__double_as_longlong(lv[c * R + la]);
      v[t][c] = loc ? x : v[t][c];
      mk = mk | (v[t][c] == kSweepMark);
    }
    need |= (t < d.nl && mk) ? 1u << t : 0u;
  }
  unsigned spin = 0;
  while (need) {
    if (++spin > kSweepSpin ||
        ((spin & 255u) == 0 && __hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))) {
      __hip_atomic_store(err, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      return false;
    }
    __builtin_amdgcn_s_sleep(1);
#pragma unroll
    for (int t = 0; t < MW; ++t)
      if ((need >> t) & 1u)
#pragma unroll
        for (int c = 0; c < NC; ++c) v[t][c] = ld_agent(out + (size_t)c * n + d.j[t]);
#pragma unroll
    for (int t = 0; t < MW; ++t) {
      if (!((need >> t) & 1u)) continue;
      bool ready = true;
#pragma unroll
      for (int c = 0; c < NC; ++c) ready = ready && v[t][c] != kSweepMark;
      if (ready) need &= ~(1u << t);
    }
  }
#pragma unroll
  for (int t = 0; t < MW; ++t) {
    double w[NC];
#pragma unroll
    for (int c = 0; c < NC; ++c) w[c] = __longlong_as_double((long long)v[t][c]);
    Row next = row;
    next.take(t, w);
    if (t < d.nl) row = next;
  }
#pragma unroll
  for (int c = 0; c < NC; ++c) {
    const double x = row.result(c);
    lv[c * R + d.lr] = x;
    st_agent(out + (size_t)c * n + d.r, x);
  }
  row.done(d);
  return true;
}

template <int NC, class Row>
struct SweepStage {
  SweepDesc d;
  Row row;
  unsigned long long v[kSweepSlots][NC];
};

template <int NC, int MW, class Row, int TB = BS>
__device__ __forceinline__ void sweep_parts(int n, int R, int P, const SweepTabs &T,
                                            const int *__restrict__ step, const int *__restrict__ sp, double *out,
                                            int *ticket, int *err, const Row &proto) {
  // descriptors four levels ahead, coefficients two and values from other
  // ranges one where the three stages fit in registers, else two, one and one
  constexpr bool deep = NC * MW <= 12;
  extern __shared__ double lv[];
  __shared__ int s_q, s_abort;
  if (threadIdx.x == 0) s_abort = 0;
  const int tid = threadIdx.x;
  for (;;) {
    if (tid == 0) s_q = atomicAdd(ticket, 1);
    __syncthreads();
    const int q = s_q;
    if (q >= P || s_abort) return;
    const int k0 = ld_uniform(sp, q), k1 = ld_uniform(sp, q + 1) - 1;
    // descriptors of level k (past the last level: dead copies of it)
    auto desc = [&](int k, SweepDesc &d) {
      const bool in = k < k1;
      const int kk = in ? k : k1 - 1;
      const int pb = ld_uniform(step, kk), pe = ld_uniform(step, kk + 1);
      sweep_desc<MW>(T, n, in ? pb + tid : pe, pe, d);
    };
    auto coef = [&](const SweepDesc &d, SweepStage<NC, Row> &S) {
      S.d = d;
      S.row.fetch(S.d, MW);
    };
    auto ext = [&](SweepStage<NC, Row> &S) { sweep_ext<NC, MW>(S.d, n, out, S.v); };
    auto stage = [&](const SweepDesc &d, SweepStage<NC, Row> &S) {
      coef(d, S);
      ext(S);
    };
    // positions of level k beyond the first TB, not pipelined
    auto rest = [&](int k) -> bool {
      const int e = ld_uniform(step, k + 1);
      bool ok = true;
      for (int p = ld_uniform(step, k) + tid + TB; ok && p < e; p += TB) {
        SweepDesc d;
        SweepStage<NC, Row> Z;
        Z.row = proto;
        sweep_desc<MW>(T, n, p, e, d);
        stage(d, Z);
        ok = sweep_row<NC, MW>(Z.d, Z.row, Z.v, R, n, out, lv, err);
      }
      return ok;
    };
    bool ok = true;
    // one level: STAGE (loads for the levels ahead, the coefficients from the
    // descriptors in DB), load the descriptors of level K + AHEAD into DB,
    // run RUN (level K), barrier.  Levels past the range's last have no live
    // positions; a failed wait is acted on once per unrolled group.
#define SWEEP_LEVEL(RUN, STAGE, DB, K, AHEAD)                                                          \
    STAGE;                                                                                             \
    desc((K) + (AHEAD), DB);                                                                           \
    if (ok && RUN.d.nl >= 0) ok = sweep_row<NC, MW>(RUN.d, RUN.row, RUN.v, R, n, out, lv, err);        \
    if ((K) < k1 && ld_uniform(step, (K) + 1) - ld_uniform(step, K) > TB && ok) ok = rest(K);          \
    __syncthreads();
    if constexpr (deep) {
      SweepStage<NC, Row> S0, S1, S2;
      S0.row = proto;
      S1.row = proto;
      S2.row = proto;
      SweepDesc D0, D1;
      desc(k0, D0);
      stage(D0, S0);
      desc(k0 + 1, D1);
      coef(D1, S1);
      desc(k0 + 2, D0);
      desc(k0 + 3, D1);
      // level k: coefficients of k + 2, values from other ranges of k + 1
      for (int k = k0; k < k1; k += 6) {
        SWEEP_LEVEL(S0, (coef(D0, S2), ext(S1)), D0, k, 4)
        SWEEP_LEVEL(S1, (coef(D1, S0), ext(S2)), D1, k + 1, 4)
        SWEEP_LEVEL(S2, (coef(D0, S1), ext(S0)), D0, k + 2, 4)
        SWEEP_LEVEL(S0, (coef(D1, S2), ext(S1)), D1, k + 3, 4)
        SWEEP_LEVEL(S1, (coef(D0, S0), ext(S2)), D0, k + 4, 4)
        SWEEP_LEVEL(S2, (coef(D1, S1), ext(S0)), D1, k + 5, 4)
        if (!ok) s_abort = 1;
        __syncthreads();
        if (s_abort) return;
      }
    } else {
      SweepStage<NC, Row> X, Y;
      X.row = proto;
      Y.row = proto;
      SweepDesc D;
      desc(k0, D);
      stage(D, X);
      desc(k0 + 1, D);
      for (int k = k0; k < k1; k += 2) {
        SWEEP_LEVEL(X, stage(D, Y), D, k, 2)
        SWEEP_LEVEL(Y, stage(D, X), D, k + 1, 2)
        if (!ok) s_abort = 1;
        __syncthreads();
        if (s_abort) return;
      }
    }
#undef SWEEP_LEVEL
    __syncthreads();
  }
}

}  // namespace fvk
