// Device-side building blocks shared by the fvk kernels (gfx950, wave64).
#pragma once
#include <hip/hip_runtime.h>

#include <climits>
#include <cstdint>

namespace fvk {

constexpr int BS = 256;          // threads per block of the row kernels (4 waves)
constexpr int WAVE = 64;
constexpr int RB = 512;          // threads per block of the reduction kernels (8 waves)
constexpr int RG = 1024;         // at most RG blocks (partials) per reduction kernel
// small = ten**(-20) as the flang build of spray-simulation/precision.f90:8
// evaluates it (bit pattern 0x3BC79CA10C924224); large = one/small.
constexpr double SMALL = 0x1.79ca10c924224p-67;
constexpr int PAD = INT_MIN;     // empty slot of the cell-face table

// Topology and geometry in HBM.  Cell arrays are in device row order; face
// arrays keep the reference's face order.  Vectors are structure-of-arrays.
struct DMesh {
  int n, nf, nbf, W, Ws, ndx;
  int ld;  // stride of the interior-slot tables spc, spf and the matrix's spv
  const double *cx, *cy, *cz, *v, *vr;
  const int *rank;                 // reference cell id (0-based) of each row
  const int *fip, *fin, *fib, *fbty;  // rows; fin = -1 and fib >= 0 on boundary faces
  const double *fs, *flpn, *fwfp;
  const double *fx, *fy, *fz, *fnx, *fny, *fnz;
  const double *dpx, *dpy, *dpz, *dnx, *dny, *dnz;
  const int *cf;    // W*n, slot-major: f (owner side), ~f (neighbour side), PAD
  // The face geometry again per cell slot (W*n, slot-major like cf), for the
  // gradient's cell gather: s, n, w, dx_p, dx_n of the slot's face, and in
  // go the other row of an interior face or ~fib of a boundary face (whose
  // gp holds the face centre minus the cell centre).  The same values as the
  // face arrays, laid out so that the rows of a wave read adjacent words.
  const double *gs, *gnx, *gny, *gnz, *gw, *gpx, *gpy, *gpz, *gqx, *gqy, *gqz;
  const int *go;
  // W*n, slot-major: for the neighbour side of an interior face, the slot of
  // that face in its owner row's list (the owner's entries hold the same
  // geometry); -1 otherwise.  The gradient reads an interior face's
  // geometry from the owner's entry, which neighbouring rows have just
  // streamed, so each face's geometry comes from HBM about once per sweep.
  const signed char *gk;
  const int *spc;   // Ws x ld interior slots: column row, -1 on padding
  const int *spf;   // Ws x ld: face id of the slot, -1 on padding
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

// ---- fence-free grid reductions ----
// A producer kernel leaves one partial per block; the next kernel reads them
// (kernel boundaries make plain stores visible) and every block adds them in
// the same fixed tree.

// Row range of a reduction kernel's block: contiguous chunks, and with a
// grid that is a multiple of 8 the blocks that land on one XCD (dispatched
// round-robin by block id) own one contiguous eighth of the rows, so
// neighbour gathers stay in that XCD's L2.
template <int B = RB>
__device__ __forceinline__ void block_rows(int n, int &r0, int &r1) {
  const int G = gridDim.x;
  const int lb = (G % 8 == 0) ? (int)(blockIdx.x % 8) * (G / 8) + (int)(blockIdx.x / 8) : (int)blockIdx.x;
  const int chunk = ((n + G - 1) / G + B - 1) / B * B;
  r0 = lb * chunk;
  r1 = min(n, r0 + chunk);
}

template <int NV, int B = RB>
__device__ __forceinline__ void rblock_sum(double (&v)[NV]) {
  __shared__ double sm[NV][B / WAVE];
#pragma unroll
  for (int o = WAVE / 2; o > 0; o >>= 1)
#pragma unroll
    for (int j = 0; j < NV; ++j) v[j] += __shfl_xor(v[j], o, WAVE);
  const int lane = threadIdx.x & (WAVE - 1), w = threadIdx.x / WAVE;
  __syncthreads();
  if (lane == 0)
#pragma unroll
    for (int j = 0; j < NV; ++j) sm[j][w] = v[j];
  __syncthreads();
#pragma unroll
  for (int j = 0; j < NV; ++j) {
    double a = 0.0;
#pragma unroll
    for (int k = 0; k < B / WAVE; ++k) a += sm[j][k];
    v[j] = a;  // every thread holds the block total
  }
}

// This block's partials, component-major: part[j*RG + blockIdx.x]
template <int NV, int B = RB>
__device__ __forceinline__ void store_partials(double (&v)[NV], double *part) {
  rblock_sum<NV, B>(v);
  if (threadIdx.x == 0)
#pragma unroll
    for (int j = 0; j < NV; ++j) part[j * RG + blockIdx.x] = v[j];
}

// Totals of components comp[0..NV) of a previous kernel's P partials,
// identical in every block (same fixed tree).
template <int NV, int B = RB>
__device__ __forceinline__ void load_sel(const double *part, int P, const int (&comp)[NV], double (&v)[NV]) {
#pragma unroll
  for (int j = 0; j < NV; ++j) v[j] = 0.0;
  for (int b = threadIdx.x; b < P; b += B)
#pragma unroll
    for (int j = 0; j < NV; ++j) v[j] += part[comp[j] * RG + b];
  rblock_sum<NV, B>(v);
}

// load_sel in two halves: issue() starts the loads of this thread's (at
// most two) partials, finish() adds them in load_sel's order and reduces.
// A kernel issues them before its own row loads, so that the reduction
// waits only for the partials while the row loads are in flight.
template <int NV>
struct SelLoads {
  double a[NV], b[NV];
  __device__ __forceinline__ void issue(const double *part, int P, const int (&comp)[NV]) {
    static_assert(RG <= 2 * RB, "two partials per thread");
    const int t = threadIdx.x;
#pragma unroll
    for (int j = 0; j < NV; ++j) {
      a[j] = t < P ? part[comp[j] * RG + t] : 0.0;
      b[j] = t + RB < P ? part[comp[j] * RG + t + RB] : 0.0;
    }
  }
  __device__ __forceinline__ void finish(double (&v)[NV]) {
#pragma unroll
    for (int j = 0; j < NV; ++j) v[j] = (0.0 + a[j]) + b[j];
    rblock_sum<NV>(v);
  }
};

// Totals of components 0..NV
template <int NV, int B = RB>
__device__ __forceinline__ void load_totals(const double *part, int P, double (&v)[NV]) {
#pragma unroll
  for (int j = 0; j < NV; ++j) v[j] = 0.0;
  for (int b = threadIdx.x; b < P; b += B)
#pragma unroll
    for (int j = 0; j < NV; ++j) v[j] += part[j * RG + b];
  rblock_sum<NV, B>(v);
}

// Interior slots of a row: WS > 0 is the table width fixed at compile time
// (loops unroll and every slot's loads issue together), 0 reads it from the
// mesh.  Padding slots (column -1) sit at the end of a row and add nothing.
template <int WS>
__device__ __forceinline__ int nslots(const DMesh &M) { return WS > 0 ? WS : M.Ws; }

// Grid of the reduction kernels: at most RG blocks of RB threads, a multiple
// of 8 once there are enough rows (XCD-contiguous row chunks).
inline int red_grid(int n) {
  int g = (n + RB - 1) / RB;
  if (g < 1) g = 1;
  if (g > RG) g = RG;
  if (g >= 16) g = g / 8 * 8;
  return g;
}
inline int grid_for(int n) { return n > 0 ? (n + BS - 1) / BS : 1; }

// MI355X dispatches the workgroups of a launch to its 8 XCDs round robin,
// and each XCD has its own L2.  xcd_block() renumbers blocks so that XCD x
// runs the contiguous range [x G/8, (x+1) G/8) of a grid of G blocks
// (blocks past the last full round keep their number): rows that are
// neighbours in the mesh then meet in one L2.  A bijection on block
// numbers, so a kernel covers the same rows.
constexpr int kXcd = 8;
__device__ __forceinline__ int xcd_block() {
  const int G = gridDim.x, b = blockIdx.x, per = G / kXcd;
  return b < per * kXcd ? (b % kXcd) * per + b / kXcd : b;
}

// The FP64 Krylov kernels (krylov.hip) run 1024-thread blocks, at most 512
// of them: measured 2.6 % faster than 512 x 512 on the plain FP64 solve, and
// 4.5 % slower on the FP32 inner solve, which keeps RB.
constexpr int RB64 = 1024;
constexpr int RG64 = 512;
inline int red_grid64(int n) {
  int g = (n + RB64 - 1) / RB64;
  if (g < 1) g = 1;
  if (g > RG64) g = RG64;
  if (g >= 16) g = g / 8 * 8;
  return g;
}

// Stride of the slot-major FP64 tables: n plus 128 KiB + 512 B of doubles
// (64 KiB + 256 B of ints), so that the same row of different slots does not
// sit at the same offset of its memory page (slots a power of two apart
// collide in the same HBM channels).
inline int slot_ld(int n) { return n + 16448; }

}  // namespace fvk
