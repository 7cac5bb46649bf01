// Device-side building blocks shared by the fvk kernels (gfx950, wave64).
#pragma once
#include <hip/hip_runtime.h>

#include <climits>
#include <cstdint>

namespace fvk {

constexpr int BS = 256;          // threads per block (4 waves)
constexpr int WAVE = 64;
constexpr int MAX_RED_BLOCKS = 2048;  // grid of the reduction kernels
// small = ten**(-20) as the flang build of spray-simulation/precision.f90:8
// evaluates it (bit pattern 0x3BC79CA10C924224); large = one/small.
constexpr double SMALL = 0x1.79ca10c924224p-67;
constexpr int PAD = INT_MIN;     // empty slot of the cell-face table

// Topology and geometry in HBM.  Cell arrays are in device row order; face
// arrays keep the reference's face order.  Vectors are structure-of-arrays.
struct DMesh {
  int n, nf, nbf, W, Ws, ndx;
  const double *cx, *cy, *cz, *v, *vr;
  const int *rank;                 // reference cell id (0-based) of each row
  const int *fip, *fin, *fib, *fbty;  // rows; fin = -1 and fib >= 0 on boundary faces
  const double *fs, *flpn, *fwfp;
  const double *fx, *fy, *fz, *fnx, *fny, *fnz;
  const double *dpx, *dpy, *dpz, *dnx, *dny, *dnz;
  const int *cf;    // W*n, slot-major: f (owner side), ~f (neighbour side), PAD
  const int *spc;   // Ws*n interior slots: column row (row itself on padding)
  const int *spf;   // Ws*n: face id of the slot, -1 on padding
};

__device__ __forceinline__ double fx_(const DMesh &M, const double *a, int f) { return a[f]; }

__device__ __forceinline__ void face_n(const DMesh &M, int f, double n[3]) {
  n[0] = M.fnx[f]; n[1] = M.fny[f]; n[2] = M.fnz ? M.fnz[f] : 0.0;
}
__device__ __forceinline__ void face_dxp(const DMesh &M, int f, double t[3]) {
  t[0] = M.dpx[f]; t[1] = M.dpy[f]; t[2] = M.dpz ? M.dpz[f] : 0.0;
}
__device__ __forceinline__ void face_dxn(const DMesh &M, int f, double t[3]) {
  t[0] = M.dnx[f]; t[1] = M.dny[f]; t[2] = M.dnz ? M.dnz[f] : 0.0;
}

// Sequential left-to-right dot product of length nd (0 + a0 b0 + a1 b1 ...),
// the order of Fortran's dot_product in the reference build.
__device__ __forceinline__ double dotn(const double *a, const double *b, int nd) {
  double s = 0.0;
  for (int d = 0; d < nd; ++d) s = s + a[d] * b[d];
  return s;
}
__device__ __forceinline__ double magn(const double *a, int nd) { return sqrt(dotn(a, a, nd)); }

// ---- deterministic reductions ----
// Tree over the 64 lanes of a wave, then over the 4 waves of the block in a
// fixed order: the result does not depend on scheduling.
template <int NV>
__device__ __forceinline__ void block_sum(double (&v)[NV]) {
  __shared__ double red_sm[NV][BS / WAVE];
#pragma unroll
  for (int o = WAVE / 2; o > 0; o >>= 1)
#pragma unroll
    for (int j = 0; j < NV; ++j) v[j] += __shfl_xor(v[j], o, WAVE);
  const int lane = threadIdx.x & (WAVE - 1), w = threadIdx.x / WAVE;
  __syncthreads();
  if (lane == 0)
#pragma unroll
    for (int j = 0; j < NV; ++j) red_sm[j][w] = v[j];
  __syncthreads();
  if (threadIdx.x == 0)
#pragma unroll
    for (int j = 0; j < NV; ++j) v[j] = ((red_sm[j][0] + red_sm[j][1]) + red_sm[j][2]) + red_sm[j][3];
}

template <int NV>
__device__ __forceinline__ void block_max(double (&v)[NV]) {
  __shared__ double red_mx[NV][BS / WAVE];
#pragma unroll
  for (int o = WAVE / 2; o > 0; o >>= 1)
#pragma unroll
    for (int j = 0; j < NV; ++j) v[j] = fmax(v[j], __shfl_xor(v[j], o, WAVE));
  const int lane = threadIdx.x & (WAVE - 1), w = threadIdx.x / WAVE;
  __syncthreads();
  if (lane == 0)
#pragma unroll
    for (int j = 0; j < NV; ++j) red_mx[j][w] = v[j];
  __syncthreads();
  if (threadIdx.x == 0)
#pragma unroll
    for (int j = 0; j < NV; ++j) v[j] = fmax(fmax(red_mx[j][0], red_mx[j][1]), fmax(red_mx[j][2], red_mx[j][3]));
}

// Grid-wide sum: every block stores its partial, the last block to arrive
// (ticket) adds the partials in block order and returns true in thread 0
// with the totals in v.  The ticket is reset by that block, so consecutive
// kernels on one stream can share it.
struct Red {
  double *part;       // MAX_RED_BLOCKS * 8
  unsigned *ticket;
};

template <int NV>
__device__ __forceinline__ bool grid_sum(double (&v)[NV], Red R) {
  block_sum<NV>(v);
  __shared__ int red_last;
  if (threadIdx.x == 0) {
#pragma unroll
    for (int j = 0; j < NV; ++j)
      __hip_atomic_store(&R.part[blockIdx.x * NV + j], v[j], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __threadfence();
    unsigned t = atomicAdd(R.ticket, 1u);
    red_last = (t == gridDim.x - 1);
  }
  __syncthreads();
  if (!red_last) return false;
  __threadfence();
  double acc[NV];
#pragma unroll
  for (int j = 0; j < NV; ++j) acc[j] = 0.0;
  for (int b = threadIdx.x; b < (int)gridDim.x; b += BS)
#pragma unroll
    for (int j = 0; j < NV; ++j)
      acc[j] += __hip_atomic_load(&R.part[b * NV + j], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  block_sum<NV>(acc);
  if (threadIdx.x == 0) {
    *R.ticket = 0u;
#pragma unroll
    for (int j = 0; j < NV; ++j) v[j] = acc[j];
  }
  return threadIdx.x == 0;
}

inline int red_grid(int n) {
  int g = (n + BS - 1) / BS;
  return g < MAX_RED_BLOCKS ? (g < 1 ? 1 : g) : MAX_RED_BLOCKS;
}
inline int grid_for(int n) { return n > 0 ? (n + BS - 1) / BS : 1; }

}  // namespace fvk
