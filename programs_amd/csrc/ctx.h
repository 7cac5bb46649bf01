// Host-side state of an fvk device context.
#pragma once
#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>

#include <cmath>
#include <string>
#include <vector>

#include "../../include/fvk.h"
#include "dev.h"

namespace fvk {

void set_error(const std::string &msg);

#define FVK_HIP(call)                                                                    \
  do {                                                                                   \
    hipError_t e_ = (call);                                                              \
    if (e_ != hipSuccess) {                                                              \
      ::fvk::set_error(std::string(#call) + ": " + hipGetErrorString(e_));              \
      return FVK_ERR_HIP;                                                                \
    }                                                                                    \
  } while (0)

// Per-component Krylov scalars, kept in device memory so that a whole solve
// runs without host round trips (done flags gate every kernel).
struct KS {
  double res0_n;  // initial L1 norm
  double res_n;   // current relative L1 norm
  double s, alpha, beta;                 // PCG (solv_cg.f90:28)
  double w1, w2, w3, w4, w5, w6;         // BiCGStab w_sc(1..6) (solv_bicg.f90:30)
  double reduc;
  int iter, n_iter, done, pad;
};

struct Levels {
  std::vector<int> ptr;  // host: level offsets into rows
  int *rows = nullptr;   // device
  int count() const { return ptr.empty() ? 0 : (int)ptr.size() - 1; }
};

// FP32 inner solver of defect correction and the factor G = D^-1/2 (I + L) Delta
// of the learned preconditioner z = G G^T r.  L sits on the lower-triangular
// pattern of A in the reference numbering (row = neighbour cell, column =
// owner cell of an interior face); its slots are split into per-row lists of
// lower (faces whose neighbour is the row) and upper (faces the row owns)
// entries.  Column indices are 16-bit offsets from the row when every
// neighbour lies within +-32767 rows, else 32-bit rows; padding slots point
// at the row itself with value 0.
struct Mix {
  bool built = false, i16 = false;
  bool qfull = false;  // factor path's SpMV from the full rows (spv, col): fewer bytes than au/iau
  int WL = 0, WU = 0, nfi = 0;
  void *col = nullptr;                 // A's interior slots (Ws*n)
  void *lcol = nullptr, *ucol = nullptr;  // factor slots (WL*n, WU*n)
  int *lfi = nullptr, *ufi = nullptr;  // interior face of each factor slot, -1 padding
  int *fi_face = nullptr;              // interior face -> face id (nfi)
  int *face_fi = nullptr;              // face id -> interior face or -1 (nf)
  int *fi_p = nullptr, *fi_n = nullptr;  // rows of an interior face's owner / neighbour
  float *spv = nullptr;                // Jacobi-scaled off-diagonals a_ij / sqrt(a_ii a_jj), FP32
  float *au = nullptr;                 // the same towards the upper neighbours only ((WU + 1) x ld)
  unsigned char *iau = nullptr;        // lower slot -> slot in the neighbour's upper list
  void *lval = nullptr, *uval = nullptr;  // L per slot, FP16
  void *e = nullptr;                   // Delta^2 per row, FP16
  void *geo = nullptr;                 // geometric MLP features per interior face (bfloat16)
  float *fl = nullptr, *fsig = nullptr;  // per interior face l_f, sig_f
  float *vec = nullptr;                // 10 x n FP32 work vectors
  double *tot = nullptr;               // 8 doubles: residual totals
  bool fac_valid = false;
  double inner_reduc = 1e-5;           // deepest relative L1 reduction of an inner solve
  double eta_margin = 0.8;             // a later inner solve aims at eta_margin x what tol needs
  int max_outer = 30;
  bool carry = false;                  // continue the Krylov recurrence across correction steps
  int replace_every = 160;             // >0: FP64 residual replacement every that many inner iterations
  int max_breakdowns = 1;              // inner breakdowns followed by a restart before the solve gives up
};

struct Ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  int n = 0, nf = 0, nbf = 0, W = 0, Ws = 0, ndx = 0;
  std::vector<int> perm, iperm;  // perm[row] = reference cell (0-based)
  std::vector<int> h_fip, h_fin, h_fib, h_fbty;  // rows
  DMesh M{};
  std::vector<void *> owned;      // device allocations freed at destroy
  // matrix on the device
  double *diag = nullptr, *spv = nullptr;  // ELL values (Ws*n)
  double *ngp = nullptr, *ngn = nullptr;   // face coefficients (nf)
  // preconditioner (per imat: c, c_r)
  double *pc_c = nullptr, *pc_cr = nullptr;
  int pc_imat = -1;
  bool pc_valid = false;
  Levels lf, lb;                  // forward / backward level sets
  // one-launch factorisation and sweeps (else one launch per level)
  bool sweep_one_launch = true;
  int sweep_blocks_per_cu = 1;    // resident blocks per CU of the chunked one-launch kernels (1: measured fastest)
  int n_cu = 256;
  double *sweep_t = nullptr;      // forward sweep output (3 x n)
  int *sweep_int = nullptr;       // tickets: forward, backward, factor; [3] = wait-bound error
  unsigned char *sweep_lmask = nullptr, *sweep_umask = nullptr;  // lower / upper slots of each row
  // partitioned one-launch sweeps (sweep_parts in sweep.h): ranges of
  // sweep_part_rows rows consecutive in the reference numbering, each walked
  // level by level by one block
  struct SweepParts {
    int P = 0, R = 0, W = 0, wmax = 0;  // wmax: the most rows of one level of a range
    int *hd = nullptr, *jc = nullptr, *lc = nullptr;  // descriptors (sweep_pack), 16-byte words
    int *step = nullptr, *sp = nullptr;
  };
  SweepParts spf, spb;
  bool sweep_parts = true;
  int sweep_part_rows = 4096;
  // Krylov workspace (3 components)
  double *wk = nullptr;           // 8 vectors x 3 components x n
  KS *ks = nullptr;               // 3 components: state after the last kernel / solve
  KS *ksr = nullptr;              // 2 x 3: scalar ring indexed by iteration parity
  double *part = nullptr;         // 4 partial buffers of RG x 8 doubles
  // scratch for reference-order transfers
  double *stage = nullptr;        // max(n, nf) * 3 doubles
  double *sys_b = nullptr, *sys_x = nullptr, *sys_r = nullptr;  // stand-alone system
  double *bnd_buf = nullptr;      // 3 x n_bf boundary values (fvk_solv_grad)
  Mix mx;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  // host polls of a solve's decisions: two pinned slots and their events
  KS *poll_h = nullptr;
  hipEvent_t poll_ev[2] = {nullptr, nullptr};
};

// Host polls of the done flag that keep the device busy: the decision after
// a chunk of iterations is copied to pinned memory behind an event, the next
// chunk is queued, and only then does the host wait for the copy.  When the
// observed convergence rate says that the solve ends within the queued work,
// the host waits for the latest poll before queueing more (a chunk queued
// after the end costs one launch per kernel that returns at its gate).
struct Poller {
  Ctx &c;
  int j = 0;            // poll index
  bool pend = false;    // poll j - 1 is posted and not taken
  // the last observations of log(relative residual) and their iterations
  static constexpr int H = 8;
  int n_obs = 0, hit[H] = {};
  double hlr[H] = {};
  int last_it = 0;
  explicit Poller(Ctx &cc) : c(cc) {}
  hipError_t post() {
    hipError_t e = hipMemcpyAsync(&c.poll_h[j & 1], c.ks, sizeof(KS), hipMemcpyDeviceToHost, c.stream);
    return e != hipSuccess ? e : hipEventRecord(c.poll_ev[j & 1], c.stream);
  }
  hipError_t take(int jj, KS &k) {
    const hipError_t e = hipEventSynchronize(c.poll_ev[jj & 1]);
    k = c.poll_h[jj & 1];
    return e;
  }
  void note(const KS &k) {
    if (!(k.res_n > 0.0)) return;
    hit[n_obs % H] = k.iter;
    hlr[n_obs % H] = std::log(k.res_n);
    last_it = k.iter;
    ++n_obs;
  }
  // iterations from the latest observation to a relative residual of reduc,
  // at the rate since the newest observation at least `span` iterations
  // older (else the oldest kept); large when not known yet
  double predict(double reduc, int span) const {
    if (n_obs < 2) return 1e30;
    const int l = (n_obs - 1) % H;
    int a = -1;
    for (int q = 1; q < std::min(n_obs, H); ++q) {
      const int m = (n_obs - 1 - q) % H;
      a = m;
      if (hit[l] - hit[m] >= span) break;
    }
    if (a < 0 || hit[l] <= hit[a]) return 1e30;
    const double rate = (hlr[l] - hlr[a]) / (hit[l] - hit[a]);
    if (!(rate < 0.0)) return 1e30;
    return (std::log(reduc) - hlr[l]) / rate;
  }
  // After the iterations queued so far (`queued`): post this poll, take the
  // previous one, and take this one too when the solve is predicted to end
  // within the queued work.  stop: a taken poll says done.  next: the
  // iterations to queue before the next poll, a chunk while the end is far,
  // else the predicted remainder and two more (so that few iterations are
  // queued after the end, each of which still costs its launches).
  hipError_t step(int queued, int chunk, double reduc, bool &stop, int &next) {
    stop = false;
    next = chunk;
    hipError_t e = post();
    if (e != hipSuccess) return e;
    KS k;
    if (pend) {
      if ((e = take(j - 1, k)) != hipSuccess) return e;
      if (k.done) stop = true;
      else note(k);
    }
    if (!stop) {
      if (predict(reduc, chunk) - (queued - last_it) < chunk) {
        if ((e = take(j, k)) != hipSuccess) return e;
        pend = false;
        if (k.done) {
          stop = true;
        } else {
          note(k);
          const double left = predict(reduc, chunk) - (queued - last_it);
          next = left < chunk ? std::max(2, (int)std::ceil(left) + 2) : chunk;
        }
      } else {
        pend = true;
      }
    }
    ++j;
    return hipSuccess;
  }
};

// Slots of the FP64 work buffer c.wk (8 slots of up to 3n doubles): their
// stride is 3n plus 64 KiB + 256 B, so that the same row of different vectors
// does not sit at the same offset of its memory page (vectors a power of two
// apart collide in the same HBM channels).
inline size_t wk_ld(int n) { return (size_t)3 * n + 8224; }
inline double *wk_slot(Ctx &c, int k) { return c.wk + (size_t)k * wk_ld(c.n); }

template <class T>
int dalloc(Ctx &c, T **p, size_t n) {
  void *q = nullptr;
  FVK_HIP(hipMalloc(&q, (n ? n : 1) * sizeof(T)));
  c.owned.push_back(q);
  FVK_HIP(hipMemsetAsync(q, 0, (n ? n : 1) * sizeof(T), c.stream));
  FVK_HIP(hipStreamSynchronize(c.stream));
  *p = (T *)q;
  return 0;
}

// krylov.hip
int precond_build(Ctx &c, int imat);
int precond_apply(Ctx &c, int imat, int nvc, const double *b, double *x, const KS *gate);
// FVK_ERR_HIP if a one-launch factorisation or sweep exceeded its wait bound
int sweep_status(Ctx &c);
int sweep_masks_build(Ctx &c);
// the tables of the partitioned sweeps (after sweep_masks_build); frees the
// previous ones
int sweep_parts_build(Ctx &c);
int spmv(Ctx &c, int nvc, const double *x, double *y);
// a[0..n) = v and b = a as kernels (capturable in a HIP graph)
int fill(Ctx &c, double *a, size_t n, double v);
int copy_d(Ctx &c, const double *a, double *b, size_t n);
int residual(Ctx &c, int nvc, const double *x, double *r, bool not_r0);
// Times one kernel per iteration with event pairs on the launch stream.
struct KTimer {
  std::vector<hipEvent_t> ev;
  int used = 0;
  double total_ms = 0.0;
  long launches = 0;
  void init(int pairs) {
    ev.resize(2 * pairs);
    for (auto &e : ev) hipEventCreate(&e);
  }
  // the next free event pair, passed to hipExtLaunchKernelGGL, which stamps
  // the kernel's own start and end (no marker packets around it)
  bool take(hipEvent_t &a, hipEvent_t &b) {
    if (used >= (int)ev.size() / 2) return false;
    a = ev[2 * used];
    b = ev[2 * used + 1];
    ++used;
    return true;
  }
  int flush() {
    for (int k = 0; k < used; ++k) {
      float ms = 0.f;
      if (hipEventSynchronize(ev[2 * k + 1]) != hipSuccess) return 1;
      if (hipEventElapsedTime(&ms, ev[2 * k], ev[2 * k + 1]) != hipSuccess) return 1;
      total_ms += ms;
      ++launches;
    }
    used = 0;
    return 0;
  }
  ~KTimer() {
    for (auto &e : ev) hipEventDestroy(e);
  }
};

// kind 0: PCG, 1: BiCGStab.  x, r0: nvc component vectors in device rows;
// r0 holds the RHS on entry and the initial residual on exit.  chunk > 0:
// stop queueing iterations once converged (checked every chunk iterations).
int krylov(Ctx &c, int kind, int imat, double reduc, int n_iter, int nvc, double *x, double *r0, bool not_r0,
           int chunk = 0, KTimer *T = nullptr);
int true_residual(Ctx &c, const double *x, const double *b, double *dev_out4, double *host_out4);
// Plain FP64 PCG (the reference recurrence with preconditioner imat) from
// x = 0 to a relative L1 residual tol.  The right-hand side is normalised to
// a mean |b| of 1 first: the recurrence's alpha = s/(p.q + small)
// (solv_cg.f90:69-70) has an absolute small = 1e-20 that stalls solves whose
// p.q gets that small.  out: the final scalars, res0_n = |b|_1.
int residual_to(Ctx &c, int nvc, const double *x, double *r, bool not_r0, double *part, double *out_dev);
int pcg_tol(Ctx &c, int imat, const double *b, double *x, double tol, int max_iter, int chunk, KTimer *T,
            KS &out);

// mixed.hip
struct DcResult {
  int iters = 0, outer = 0, broken = 0, breakdowns = 0;
  double res_l1 = 0.0, res_l2 = 0.0, b_l1 = 0.0;
  int inner[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  double res_before[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  bool converged = false;
};
int mix_build(Ctx &c);
// factor from per-interior-face (l, sig) already in c.mx.fl / c.mx.fsig
int factor_assemble(Ctx &c);
int factor_apply(Ctx &c, const double *r, double *z);
// FP32-inner CG in FP64 defect correction on (c.diag, c.spv), b -> x (FP64
// rows); use_factor: G G^T preconditioner (else Jacobi)
int dc_solve(Ctx &c, bool use_factor, const double *b, double *x, double tol, int max_iter, DcResult &res,
             KTimer *T);

// learn.hip: factor of the current matrix from the network, and Adam steps
// on the current matrix
int learned_factor(Ctx &c, fvk_model *m);
int learned_train(Ctx &c, fvk_model *m, int steps, double lr, double *last_loss);

// flow.hip
// solv_grad: n_iter Green-Gauss sweeps; g0 is scratch of the size of g;
// on_flag (device, may be null): all-zero gradient unless *on_flag > small.
int grad(Ctx &c, int nvc, const double *phi, const double *phib, double *g, int n_iter, int neu,
         const double *on_flag, double *g0);

}  // namespace fvk

// The opaque handle of the C ABI
struct fvk_ctx {
  fvk::Ctx c;
};
