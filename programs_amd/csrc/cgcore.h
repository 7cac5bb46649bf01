// Scalar decisions of the PCG recurrence (solv_cg.f90), shared by the FP64
// reference-order solver and the FP32 inner solver of defect correction.
#pragma once
#include "ctx.h"

namespace fvk {

__device__ __forceinline__ bool all_done(const KS *ks, int nvc) {
  if (!ks) return false;
  for (int q = 0; q < nvc; ++q)
    if (!ks[q].done) return false;
  return true;
}

// y = diag*x + sum over the row's interior faces (solv_cg.f90:56-64)
template <int WS>
__device__ __forceinline__ double row_spmv(const DMesh &M, const double *__restrict__ diag,
                                           const double *__restrict__ spv, const double *__restrict__ x,
                                           int i) {
  const size_t ld = M.ld;
  double y = diag[i] * x[i];
#pragma unroll
  for (int k = 0; k < nslots<WS>(M); ++k) {
    const int j = M.spc[k * ld + i];
    const double a = spv[k * ld + i];
    if (j >= 0) y = y + a * x[j];
  }
  return y;
}

// ---- PCG (solv_cg.f90) ----
// Scalars at the start of iteration it from the totals of the previous
// kernel: L1 = |r|_1, rz = r.z.  it = 0: L1 is res0_n; otherwise the exit
// test on res_n = L1/(res0_n+small) of iteration it (it iterations done).
__device__ __forceinline__ void cg_decide(int it, int nvc, const double *L1, const double *rz, const KS *prev,
                                          KS *cur) {
  double mx = 0.0;
  for (int q = 0; q < nvc; ++q) {
    KS k = prev[q];
    if (it == 0) {
      k.res0_n = L1[q];
      k.iter = 0;
      k.res_n = 0.0;
      mx = q == 0 ? L1[q] : fmax(mx, L1[q]);
    } else if (!k.done) {
      k.res_n = L1[q] / (k.res0_n + SMALL);
      k.iter = it;
      if (k.res_n < k.reduc || it >= k.n_iter) k.done = 1;
    }
    cur[q] = k;
  }
  if (it == 0)
    for (int q = 0; q < nvc; ++q)
      if (mx < SMALL || cur[q].n_iter <= 0) cur[q].done = 1;  // solv_cg.f90:21
  for (int q = 0; q < nvc; ++q) {
    KS &k = cur[q];
    if (k.done) continue;
    k.beta = 1.0 / k.s;
    k.s = rz[q];
    k.beta = k.beta * k.s;
  }
}

}  // namespace fvk
