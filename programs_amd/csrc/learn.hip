// The learned part of the pressure preconditioner: a per-face MLP that maps
// twelve local features to the entries of the factor G = D^-1/2 (I + L)
// Delta (mixed.hip applies z = G G^T r), its training loss and an Adam
// optimiser.  No counterpart in the reference.
//
// Network: 12 -> 64 -> 64 -> 2, tanh, outputs (l_f, sig_f) per interior
// face.  Forward and backward run on the matrix cores with hand-written
// v_mfma_f32_16x16x16_bf16 tiles: bfloat16 operands (weights and
// activations rounded to nearest even), float32 accumulation.  A wave owns a
// tile of 16 faces and computes every layer transposed, Z^T = W^T X^T, so
// that the accumulator layout of one layer is the B-operand layout of the
// next and no shuffle is needed between layers:
//   A operand (16x16, M x K): lane l holds A[l%16][4(l/16) .. +3]
//   B operand (16x16, K x N): lane l holds B[4(l/16) .. +3][l%16]
//   D (16x16):                lane l holds D[4(l/16) .. +3][l%16]
// The weight gradients sum over faces (K = faces), for which the tile's
// activations are transposed through the wave's own LDS area.
//
// Loss: Kaporin's measure of the Jacobi-scaled preconditioned matrix,
// J = log(tr(Gh^T Ah Gh)/N) - (2/N) sum_c s_c with Gh = (I + L) Delta,
// computed column by column (column c holds Delta_c and Delta_c l_f for the
// faces c owns), in FP64.  Its minimiser over the pattern is FSAI.
#include <cmath>
#include <cstring>
#include <vector>

#include "ctx.h"

namespace fvk {

constexpr int NIN = 12, NH = 64;
constexpr int OFF_W1 = 0, OFF_B1 = 768, OFF_W2 = 832, OFF_B2 = 4928, OFF_W3 = 4992, OFF_B3 = 5120;
constexpr int NPARAM = 5122;
constexpr int MB = 256;          // threads per MLP block (4 waves)
constexpr int MLP_FWD_BLOCKS = 2048;
constexpr int MLP_BWD_BLOCKS = 256;

typedef short v4s __attribute__((ext_vector_type(4)));
typedef float v4f __attribute__((ext_vector_type(4)));

__device__ __forceinline__ v4f mfma16(v4s a, v4s b, v4f c) {
  return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
}

// float -> bfloat16 bits, round to nearest even (v_cvt_pk_bf16_f32, two
// values per instruction when the compiler can pair them)
__device__ __forceinline__ short bf16(float x) { return __builtin_bit_cast(short, (__bf16)x); }
__device__ __forceinline__ v4s pack4(v4f v) {
  v4s r;
  r[0] = bf16(v[0]); r[1] = bf16(v[1]); r[2] = bf16(v[2]); r[3] = bf16(v[3]);
  return r;
}
// tanh(x) = 1 - 2 / (exp(2x) + 1) with the hardware exp2 and reciprocal
// (absolute error ~1e-7, below the bfloat16 rounding that follows);
// saturates to +-1
__device__ __forceinline__ float tanh_fast(float x) {
  const float e = __builtin_amdgcn_exp2f(2.8853900817779268f * x);  // exp(2x)
  return 1.f - 2.f * __builtin_amdgcn_rcpf(e + 1.f);
}
__device__ __forceinline__ v4f tanh4(v4f v) {
  v4f r;
  r[0] = tanh_fast(v[0]); r[1] = tanh_fast(v[1]); r[2] = tanh_fast(v[2]); r[3] = tanh_fast(v[3]);
  return r;
}

// Weights of one block in LDS as bfloat16: transposed copies for the
// forward A operands, natural ones for the backward.
struct WLds {
  short w1t[64 * 16];  // [o][i], i < 12
  short w2t[64 * 64];  // [o][i]
  short w3t[16 * 64];  // [o][i], o < 2
  short w2[64 * 64];   // [i][o]
  short w3[64 * 16];   // [i][o], o < 2
  float b1[64], b2[64], b3[16];
};

__device__ void load_weights(const float *__restrict__ th, WLds &s, bool bwd) {
  for (int k = threadIdx.x; k < 64 * 16; k += blockDim.x) {
    const int o = k / 16, i = k % 16;
    s.w1t[k] = i < NIN ? bf16(th[OFF_W1 + i * 64 + o]) : (short)0;
    const int o3 = k / 64, i3 = k % 64;
    s.w3t[k] = o3 < 2 ? bf16(th[OFF_W3 + i3 * 2 + o3]) : (short)0;
    if (bwd) {
      const int i4 = k / 16, o4 = k % 16;
      s.w3[k] = o4 < 2 ? bf16(th[OFF_W3 + i4 * 2 + o4]) : (short)0;
    }
  }
  for (int k = threadIdx.x; k < 64 * 64; k += blockDim.x) {
    const int o = k / 64, i = k % 64;
    s.w2t[k] = bf16(th[OFF_W2 + i * 64 + o]);
    if (bwd) s.w2[k] = bf16(th[OFF_W2 + k]);
  }
  for (int k = threadIdx.x; k < 64; k += blockDim.x) {
    s.b1[k] = th[OFF_B1 + k];
    s.b2[k] = th[OFF_B2 + k];
    if (k < 16) s.b3[k] = k < 2 ? th[OFF_B3 + k] : 0.f;
  }
  __syncthreads();
}

// Inputs of an interior face q for the MLP.
struct FaceIn {
  DMesh M;
  const double *diag, *ngp;
  const int *fi_face, *fi_p, *fi_n;
  int nfi;
  // the geometric features, bfloat16, computed once per mesh: feature 3
  // (g0) and features 4-7 and 8-11 (g1, g2) of each interior face
  const short *g0;
  const v4s *g1, *g2;
};

// Geometric features 3 .. 11 of interior face q (the bfloat16 B operand
// entries: 3 alone, 4-7, 8-11).  The mesh does not change within a context:
// they are computed once (k_geo_feat) and read by every forward and
// backward pass.  Lengths in FP64 from the FP64 mesh; the cube root and the
// logarithm in FP32 (the features enter the network in bfloat16).
__device__ void geo_features(const FaceIn &F, int q, float &x3, v4f &x47, v4f &x811) {
  const int f = F.fi_face[q], p = F.fi_p[q], m = F.fi_n[q];
  // |V|: the reference's volumes (x_f s_f n_x summed from the origin,
  // grd_geom.f90:151-153,172-173) turn negative on strongly warped cells
  const double vp = fmax(fabs(F.M.v[p]), 1e-300), vn = fmax(fabs(F.M.v[m]), 1e-300);
  const double h = (double)cbrtf((float)(0.5 * (vp + vn)));
  x3 = (float)(F.M.fs[f] / (h * h));
  double nv[3];
  face_n(F.M, f, nv);
  x47[0] = (float)nv[0];
  x47[1] = (float)nv[1];
  x47[2] = (float)nv[2];
  x47[3] = (float)(F.M.flpn[f] / h);
  double dp[3], dn[3];
  face_dxp(F.M, f, dp);
  face_dxn(F.M, f, dn);
  x811[0] = (float)F.M.fwfp[f];
  x811[1] = (float)(sqrt(dp[0] * dp[0] + dp[1] * dp[1] + dp[2] * dp[2]) / h);
  x811[2] = (float)(sqrt(dn[0] * dn[0] + dn[1] * dn[1] + dn[2] * dn[2]) / h);
  x811[3] = logf((float)(vp / vn));
}

__global__ __launch_bounds__(BS) void k_geo_feat(FaceIn F, short *g0, v4s *g1, v4s *g2) {
  const int q = blockIdx.x * BS + threadIdx.x;
  if (q >= F.nfi) return;
  float x3;
  v4f a, b;
  geo_features(F, q, x3, a, b);
  g0[q] = bf16(x3);
  g1[q] = pack4(a);
  g2[q] = pack4(b);
}

__device__ v4s features4(const FaceIn &F, int q, int g);

// The features as the network reads them, as float: X[12 q + k]
__global__ __launch_bounds__(BS) void k_feat_dump(FaceIn F, float *__restrict__ X) {
  const int q = blockIdx.x * BS + threadIdx.x;
  if (q >= F.nfi) return;
  for (int g = 0; g < 3; ++g) {
    const v4s v = features4(F, q, g);
    for (int k = 0; k < 4; ++k)
      X[(size_t)12 * q + 4 * g + k] = __builtin_bit_cast(float, (unsigned)(unsigned short)v[k] << 16);
  }
}

// Features 4g .. 4g+3 of interior face q (the lane's share of the B operand):
// the coefficient ratios of group 0 from the current matrix, the rest stored
__device__ v4s features4(const FaceIn &F, int q, int g) {
  v4s r = {0, 0, 0, 0};
  if (q < F.nfi && g < 3) {
    if (g == 0) {
      const int f = F.fi_face[q], p = F.fi_p[q], m = F.fi_n[q];
      const double a = -F.ngp[f], dp = F.diag[p], dn = F.diag[m];
      r[0] = bf16((float)(a / sqrt(dp * dn)));
      r[1] = bf16((float)(a / dp));
      r[2] = bf16((float)(a / dn));
      r[3] = F.g0[q];
    } else {
      r = g == 1 ? F.g1[q] : F.g2[q];
    }
  }
  return r;
}

// Forward pass of one 16-face tile.  h1, h2: tanh activations (float, D
// layout), B1, B2: their bfloat16 B operands, out: D of the output layer.
__device__ __forceinline__ void tile_forward(const WLds &s, v4s xb, v4f (&h1)[4], v4s (&B1)[4], v4f (&h2)[4],
                                             v4s (&B2)[4], v4f &out) {
  const int l = threadIdx.x & 63, r = l & 15, g = l >> 4;
#pragma unroll
  for (int ob = 0; ob < 4; ++ob) {
    v4f acc = {s.b1[ob * 16 + 4 * g], s.b1[ob * 16 + 4 * g + 1], s.b1[ob * 16 + 4 * g + 2],
               s.b1[ob * 16 + 4 * g + 3]};
    acc = mfma16(*(const v4s *)&s.w1t[(ob * 16 + r) * 16 + 4 * g], xb, acc);
    h1[ob] = tanh4(acc);
    B1[ob] = pack4(h1[ob]);
  }
#pragma unroll
  for (int ob = 0; ob < 4; ++ob) {
    v4f acc = {s.b2[ob * 16 + 4 * g], s.b2[ob * 16 + 4 * g + 1], s.b2[ob * 16 + 4 * g + 2],
               s.b2[ob * 16 + 4 * g + 3]};
#pragma unroll
    for (int kb = 0; kb < 4; ++kb) acc = mfma16(*(const v4s *)&s.w2t[(ob * 16 + r) * 64 + kb * 16 + 4 * g], B1[kb], acc);
    h2[ob] = tanh4(acc);
    B2[ob] = pack4(h2[ob]);
  }
  v4f acc = {s.b3[4 * g], s.b3[4 * g + 1], s.b3[4 * g + 2], s.b3[4 * g + 3]};
#pragma unroll
  for (int kb = 0; kb < 4; ++kb) acc = mfma16(*(const v4s *)&s.w3t[r * 64 + kb * 16 + 4 * g], B2[kb], acc);
  out = acc;
}

// The same forward pass keeping only the bfloat16 operands (the factor build
// needs no activations for a backward pass: fewer registers, more waves).
__device__ __forceinline__ v4f tile_forward_out(const WLds &s, v4s xb) {
  const int l = threadIdx.x & 63, r = l & 15, g = l >> 4;
  v4s B1[4], B2[4];
#pragma unroll
  for (int ob = 0; ob < 4; ++ob) {
    v4f acc = {s.b1[ob * 16 + 4 * g], s.b1[ob * 16 + 4 * g + 1], s.b1[ob * 16 + 4 * g + 2],
               s.b1[ob * 16 + 4 * g + 3]};
    acc = mfma16(*(const v4s *)&s.w1t[(ob * 16 + r) * 16 + 4 * g], xb, acc);
    B1[ob] = pack4(tanh4(acc));
  }
#pragma unroll
  for (int ob = 0; ob < 4; ++ob) {
    v4f acc = {s.b2[ob * 16 + 4 * g], s.b2[ob * 16 + 4 * g + 1], s.b2[ob * 16 + 4 * g + 2],
               s.b2[ob * 16 + 4 * g + 3]};
#pragma unroll
    for (int kb = 0; kb < 4; ++kb) acc = mfma16(*(const v4s *)&s.w2t[(ob * 16 + r) * 64 + kb * 16 + 4 * g], B1[kb], acc);
    B2[ob] = pack4(tanh4(acc));
  }
  v4f acc = {s.b3[4 * g], s.b3[4 * g + 1], s.b3[4 * g + 2], s.b3[4 * g + 3]};
#pragma unroll
  for (int kb = 0; kb < 4; ++kb) acc = mfma16(*(const v4s *)&s.w3t[r * 64 + kb * 16 + 4 * g], B2[kb], acc);
  return acc;
}

// Outputs (l_f, sig_f) of every interior face.
__global__ __launch_bounds__(MB) void k_mlp_fwd(FaceIn F, const float *__restrict__ theta, float *__restrict__ fl,
                                                float *__restrict__ fsig) {
  __shared__ WLds s;
  load_weights(theta, s, false);
  const int l = threadIdx.x & 63, r = l & 15, g = l >> 4;
  const int wave = blockIdx.x * (MB / 64) + threadIdx.x / 64, nwaves = gridDim.x * (MB / 64);
  const int tiles = (F.nfi + 15) / 16;
  for (int t = wave; t < tiles; t += nwaves) {
    const int q = t * 16 + r;
    const v4f out = tile_forward_out(s, features4(F, q, g));
    if (g == 0 && q < F.nfi) {
      fl[q] = out[0];
      fsig[q] = out[1];
    }
  }
}

// Per-wave LDS area for the transposes of the weight-gradient products
struct TLds {
  short x[16 * 16], h1[64 * 16], h2[64 * 16], z1[64 * 16], z2[64 * 16], o[16 * 16];
};

// D-layout value block b of a [rows][16 faces] transposed store
__device__ __forceinline__ void put_t(short *dst, int b, v4s v) {
  const int l = threadIdx.x & 63, r = l & 15, g = l >> 4;
#pragma unroll
  for (int e = 0; e < 4; ++e) dst[(b * 16 + 4 * g + e) * 16 + r] = v[e];
}

// Parameter gradient of sum_f dO_f . out_f, one partial per block.
__global__ __launch_bounds__(MB) void k_mlp_bwd(FaceIn F, const float *__restrict__ theta,
                                                const float *__restrict__ dl, const float *__restrict__ ds,
                                                float *__restrict__ gpart) {
  __shared__ WLds s;
  __shared__ TLds tl[MB / 64];
  __shared__ float red[NPARAM];
  load_weights(theta, s, true);
  const int l = threadIdx.x & 63, r = l & 15, g = l >> 4, w = threadIdx.x / 64;
  const int wave = blockIdx.x * (MB / 64) + w, nwaves = gridDim.x * (MB / 64);
  const int tiles = (F.nfi + 15) / 16;
  TLds &T = tl[w];
  v4f aw2[4][4], aw1[4], aw3[4];
  float bb1[4][4], bb2[4][4], bb3[4];
#pragma unroll
  for (int a = 0; a < 4; ++a) {
    aw1[a] = v4f{0.f, 0.f, 0.f, 0.f};
    aw3[a] = v4f{0.f, 0.f, 0.f, 0.f};
    bb3[a] = 0.f;
#pragma unroll
    for (int b = 0; b < 4; ++b) {
      aw2[a][b] = v4f{0.f, 0.f, 0.f, 0.f};
      bb1[a][b] = 0.f;
      bb2[a][b] = 0.f;
    }
  }
  for (int t = wave; t < tiles; t += nwaves) {
    const int q = t * 16 + r;
    const v4s xb = features4(F, q, g);
    v4f h1[4], h2[4], out;
    v4s B1[4], B2[4];
    tile_forward(s, xb, h1, B1, h2, B2, out);
    // dO^T: rows (l, sig) of face r in the lanes with g = 0
    v4f dO = {0.f, 0.f, 0.f, 0.f};
    if (g == 0 && q < F.nfi) {
      dO[0] = dl[q];
      dO[1] = ds[q];
    }
    const v4s BdO = pack4(dO);
    // dZ2^T = (W3 dO^T) * (1 - h2^2)
    v4f dz2[4];
    v4s Bz2[4];
#pragma unroll
    for (int ib = 0; ib < 4; ++ib) {
      v4f acc = {0.f, 0.f, 0.f, 0.f};
      acc = mfma16(*(const v4s *)&s.w3[(ib * 16 + r) * 16 + 4 * g], BdO, acc);
#pragma unroll
      for (int e = 0; e < 4; ++e) dz2[ib][e] = acc[e] * (1.f - h2[ib][e] * h2[ib][e]);
      Bz2[ib] = pack4(dz2[ib]);
    }
    // dZ1^T = (W2 dZ2^T) * (1 - h1^2)
    v4f dz1[4];
    v4s Bz1[4];
#pragma unroll
    for (int ib = 0; ib < 4; ++ib) {
      v4f acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
      for (int ob = 0; ob < 4; ++ob) acc = mfma16(*(const v4s *)&s.w2[(ib * 16 + r) * 64 + ob * 16 + 4 * g], Bz2[ob], acc);
#pragma unroll
      for (int e = 0; e < 4; ++e) dz1[ib][e] = acc[e] * (1.f - h1[ib][e] * h1[ib][e]);
      Bz1[ib] = pack4(dz1[ib]);
    }
    // bias gradients (float values, summed over the tile's faces below)
#pragma unroll
    for (int b = 0; b < 4; ++b)
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        bb1[b][e] += dz1[b][e];
        bb2[b][e] += dz2[b][e];
      }
#pragma unroll
    for (int e = 0; e < 4; ++e) bb3[e] += dO[e];
    // transposes through the wave's LDS area
    put_t(T.x, 0, xb);
    put_t(T.o, 0, BdO);
#pragma unroll
    for (int b = 0; b < 4; ++b) {
      put_t(T.h1, b, B1[b]);
      put_t(T.h2, b, B2[b]);
      put_t(T.z1, b, Bz1[b]);
      put_t(T.z2, b, Bz2[b]);
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    // dW2[i][o] += H1^T[i][f] dZ2[f][o]
#pragma unroll
    for (int ib = 0; ib < 4; ++ib) {
      const v4s a = *(const v4s *)&T.h1[(ib * 16 + r) * 16 + 4 * g];
#pragma unroll
      for (int ob = 0; ob < 4; ++ob) aw2[ib][ob] = mfma16(a, *(const v4s *)&T.z2[(ob * 16 + r) * 16 + 4 * g], aw2[ib][ob]);
    }
    // dW1[x][i] += X^T[x][f] dZ1[f][i]
    {
      const v4s a = *(const v4s *)&T.x[r * 16 + 4 * g];
#pragma unroll
      for (int ob = 0; ob < 4; ++ob) aw1[ob] = mfma16(a, *(const v4s *)&T.z1[(ob * 16 + r) * 16 + 4 * g], aw1[ob]);
    }
    // dW3[i][o] += H2^T[i][f] dO[f][o]
    {
      const v4s b = *(const v4s *)&T.o[r * 16 + 4 * g];
#pragma unroll
      for (int ib = 0; ib < 4; ++ib) aw3[ib] = mfma16(*(const v4s *)&T.h2[(ib * 16 + r) * 16 + 4 * g], b, aw3[ib]);
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  }
  // bias sums over the 16 faces of a lane group
#pragma unroll
  for (int o = 1; o < 16; o <<= 1)
#pragma unroll
    for (int a = 0; a < 4; ++a) {
      bb3[a] += __shfl_xor(bb3[a], o, 64);
#pragma unroll
      for (int b = 0; b < 4; ++b) {
        bb1[a][b] += __shfl_xor(bb1[a][b], o, 64);
        bb2[a][b] += __shfl_xor(bb2[a][b], o, 64);
      }
    }
  // block sum, wave by wave in a fixed order
  for (int ww = 0; ww < MB / 64; ++ww) {
    if (w == ww) {
      const bool first = ww == 0;
#define ACC(idx, val) red[idx] = first ? (val) : red[idx] + (val)
#pragma unroll
      for (int ib = 0; ib < 4; ++ib)
#pragma unroll
        for (int ob = 0; ob < 4; ++ob)
#pragma unroll
          for (int e = 0; e < 4; ++e) ACC(OFF_W2 + (ib * 16 + 4 * g + e) * 64 + ob * 16 + r, aw2[ib][ob][e]);
#pragma unroll
      for (int ob = 0; ob < 4; ++ob)
#pragma unroll
        for (int e = 0; e < 4; ++e)
          if (4 * g + e < NIN) ACC(OFF_W1 + (4 * g + e) * 64 + ob * 16 + r, aw1[ob][e]);
#pragma unroll
      for (int ib = 0; ib < 4; ++ib)
#pragma unroll
        for (int e = 0; e < 4; ++e)
          if (r < 2) ACC(OFF_W3 + (ib * 16 + 4 * g + e) * 2 + r, aw3[ib][e]);
      if (r == 0) {
#pragma unroll
        for (int b = 0; b < 4; ++b)
#pragma unroll
          for (int e = 0; e < 4; ++e) {
            ACC(OFF_B1 + b * 16 + 4 * g + e, bb1[b][e]);
            ACC(OFF_B2 + b * 16 + 4 * g + e, bb2[b][e]);
          }
        if (g == 0) {
          ACC(OFF_B3, bb3[0]);
          ACC(OFF_B3 + 1, bb3[1]);
        }
      }
#undef ACC
    }
    __syncthreads();
  }
  for (int k = threadIdx.x; k < NPARAM; k += MB) gpart[(size_t)blockIdx.x * NPARAM + k] = red[k];
}

// ---- loss (FP64) ----
// Ah entry of interior face q
__device__ __forceinline__ double ahat(const FaceIn &F, int q) {
  const int f = F.fi_face[q];
  return F.ngp[f] / sqrt(F.diag[F.fi_p[q]] * F.diag[F.fi_n[q]]);
}

// FSAI factor in closed form, per column c over the faces q it owns:
// l_q = -Ah_q, sig_q = -log(1 - sum_q Ah_q^2) / (2 cnt_c).  It is the
// Kaporin minimiser (FSAI) when no two higher neighbours of a cell are
// adjacent (hexahedral and quadrilateral meshes); elsewhere an approximation
// of it.  A column whose sum reaches 1 keeps Jacobi (l = sig = 0).
__global__ __launch_bounds__(BS) void k_fsai(FaceIn F, int WU, const int *__restrict__ ufi, float *__restrict__ fl,
                                             float *__restrict__ fsig) {
  const int c = blockIdx.x * BS + threadIdx.x, n = F.M.n;
  if (c >= n) return;
  double q2 = 0.0;
  int cnt = 0;
  for (int k = 0; k < WU; ++k) {
    const int q = ufi[(size_t)k * n + c];
    if (q < 0) continue;
    const double a = ahat(F, q);
    q2 += a * a;
    ++cnt;
  }
  const bool ok = cnt > 0 && q2 < 1.0;
  const double sg = ok ? -0.5 * log1p(-q2) / cnt : 0.0;
  for (int k = 0; k < WU; ++k) {
    const int q = ufi[(size_t)k * n + c];
    if (q < 0) continue;
    fl[q] = ok ? (float)(-ahat(F, q)) : 0.f;
    fsig[q] = (float)sg;
  }
}

// Pf_q = l_q + Ah_q + sum over the other faces q2 owned by the same cell c
// whose neighbour j is adjacent to q's neighbour i of Ah_ij l_q2
__global__ __launch_bounds__(BS) void k_loss_face(FaceIn F, int WU, const int *__restrict__ ufi,
                                                  const float *__restrict__ fl, double *__restrict__ Pf) {
  const int q = blockIdx.x * BS + threadIdx.x;
  if (q >= F.nfi) return;
  const int n = F.M.n, c = F.fi_p[q], i = F.fi_n[q];
  double v = (double)fl[q] + ahat(F, q);
  for (int k = 0; k < WU; ++k) {
    const int q2 = ufi[(size_t)k * n + c];
    if (q2 < 0 || q2 == q) continue;
    const int j = F.fi_n[q2];
    for (int s = 0; s < F.M.Ws; ++s) {
      if (F.M.spc[(size_t)s * F.M.ld + i] != j) continue;
      const int f3 = F.M.spf[(size_t)s * F.M.ld + i];
      v += F.ngp[f3] / sqrt(F.diag[i] * F.diag[j]) * (double)fl[q2];
    }
  }
  Pf[q] = v;
}

// Per column c: e_c = exp(2 s_c), tau_c; partials {e tau, s}
__global__ __launch_bounds__(RB) void k_loss_col(FaceIn F, int WU, const int *__restrict__ ufi,
                                                 const float *__restrict__ fl, const float *__restrict__ fsig,
                                                 const double *__restrict__ Pf, double *__restrict__ ec,
                                                 double *__restrict__ tau, double *__restrict__ part) {
  const int n = F.M.n;
  double a[2] = {0.0, 0.0};
  int r0, r1;
  block_rows(n, r0, r1);
  for (int c = r0 + threadIdx.x; c < r1; c += RB) {
    double sc = 0.0, pc = 1.0, tq = 0.0;
    for (int k = 0; k < WU; ++k) {
      const int q = ufi[(size_t)k * n + c];
      if (q < 0) continue;
      const double lq = (double)fl[q];
      sc += (double)fsig[q];
      pc += ahat(F, q) * lq;
      tq += lq * Pf[q];
    }
    const double e = exp(2.0 * sc), t = pc + tq;
    ec[c] = e;
    tau[c] = t;
    a[0] += e * t;
    a[1] += sc;
  }
  store_partials<2>(a, part);
}

// dJ/dl_q = 2/T e_c Pf_q ; dJ/dsig_q = 2/T e_c tau_c - 2/N (c = owner of q);
// J into jout[0] (and T into jout[1]) by block 0.
__global__ __launch_bounds__(RB) void k_loss_grad(FaceIn F, const double *__restrict__ Pf,
                                                  const double *__restrict__ ec, const double *__restrict__ tau,
                                                  const double *__restrict__ part, int P, float *__restrict__ dl,
                                                  float *__restrict__ ds, double *__restrict__ jout) {
  double t[2];
  load_totals<2>(part, P, t);
  const double T = t[0], N = (double)F.M.n;
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    jout[0] = log(T / N) - 2.0 / N * t[1];
    jout[1] = T;
  }
  for (int q = blockIdx.x * RB + threadIdx.x; q < F.nfi; q += gridDim.x * RB) {
    const int c = F.fi_p[q];
    dl[q] = (float)(2.0 / T * ec[c] * Pf[q]);
    ds[q] = (float)(2.0 / T * ec[c] * tau[c] - 2.0 / N);
  }
}

// Sum of the block partials (fixed order), then Adam (Kingma and Ba 2015).
__global__ __launch_bounds__(BS) void k_adam(float *__restrict__ theta, float *__restrict__ m1,
                                             float *__restrict__ m2, const float *__restrict__ gpart, int nblk,
                                             float *__restrict__ gout, int update, float lr, float b1, float b2,
                                             float eps, float c1, float c2) {
  const int k = blockIdx.x * BS + threadIdx.x;
  if (k >= NPARAM) return;
  float gk = 0.f;
  for (int b = 0; b < nblk; ++b) gk += gpart[(size_t)b * NPARAM + k];
  if (gout) gout[k] = gk;
  if (!update) return;
  const float m = b1 * m1[k] + (1.f - b1) * gk, v = b2 * m2[k] + (1.f - b2) * gk * gk;
  m1[k] = m;
  m2[k] = v;
  theta[k] -= lr * (m / c1) / (sqrtf(v / c2) + eps);
}

}  // namespace fvk

using namespace fvk;

struct fvk_model {
  int device = 0;
  float *theta = nullptr, *m1 = nullptr, *m2 = nullptr, *gpart = nullptr, *grad = nullptr;
  int step = 0;
};

namespace {

struct LossBufs {
  double *Pf, *ec, *tau, *jout;
  float *dl, *ds;
};

FaceIn face_in(Ctx &c) {
  FaceIn F;
  F.M = c.M;
  F.diag = c.diag;
  F.ngp = c.ngp;
  F.fi_face = c.mx.fi_face;
  F.fi_p = c.mx.fi_p;
  F.fi_n = c.mx.fi_n;
  F.nfi = c.mx.nfi;
  const size_t nfi = c.mx.nfi;
  short *g = (short *)c.mx.geo;
  F.g1 = (const v4s *)g;
  F.g2 = (const v4s *)(g + 4 * nfi);
  F.g0 = g + 8 * nfi;
  return F;
}

// The geometric features of the context's mesh, once
int geo_build(Ctx &c) {
  Mix &X = c.mx;
  if (X.geo) return 0;
  short *g;
  int rc = dalloc(c, &g, (size_t)9 * std::max(X.nfi, 1));
  if (rc) return rc;
  X.geo = g;
  const FaceIn F = face_in(c);
  k_geo_feat<<<grid_for(X.nfi), BS, 0, c.stream>>>(F, (short *)F.g0, (v4s *)F.g1, (v4s *)F.g2);
  FVK_HIP(hipGetLastError());
  return 0;
}

// Loss scratch inside the context's Krylov workspace (8 x 3 x n doubles,
// unused between solves): Pf, ec, tau, J and T, dl, ds
int loss_bufs(Ctx &c, LossBufs &B) {
  const size_t n = c.n, nfi = c.mx.nfi;
  // doubles: Pf (nfi), ec and tau (n each), 8 scalars, then two float
  // arrays of nfi (nfi doubles together)
  if (2 * nfi + 2 * n + 8 > (size_t)24 * n) {
    set_error("loss scratch too small for this mesh");
    return FVK_ERR_STATE;
  }
  double *w = c.wk;
  B.Pf = w;
  B.ec = w + nfi;
  B.tau = B.ec + n;
  B.jout = B.tau + n;
  float *fw = (float *)(B.jout + 8);
  B.dl = fw;
  B.ds = fw + nfi;
  return 0;
}

int check_model(fvk_model *m, Ctx &c) {
  if (m->device != c.device) {
    set_error("model and context live on different devices");
    return FVK_ERR_ARG;
  }
  return 0;
}

int run_forward(fvk_model *m, Ctx &c) {
  int rc = mix_build(c);
  if (rc) return rc;
  if ((rc = geo_build(c))) return rc;
  const FaceIn F = face_in(c);
  const int tiles = (F.nfi + 15) / 16;
  int blocks = std::min(MLP_FWD_BLOCKS, std::max(1, (tiles + 3) / 4));
  k_mlp_fwd<<<blocks, MB, 0, c.stream>>>(F, m->theta, c.mx.fl, c.mx.fsig);
  FVK_HIP(hipGetLastError());
  return 0;
}

// forward, loss, output gradient and the parameter gradient partials
int run_loss_grad(fvk_model *m, Ctx &c, LossBufs &B, int &nblk) {
  int rc = run_forward(m, c);
  if (rc) return rc;
  if ((rc = loss_bufs(c, B))) return rc;
  const FaceIn F = face_in(c);
  const int G = red_grid(c.n);
  double *part = c.part + (size_t)3 * RG * 8;
  k_loss_face<<<grid_for(F.nfi), BS, 0, c.stream>>>(F, c.mx.WU, c.mx.ufi, c.mx.fl, B.Pf);
  k_loss_col<<<G, RB, 0, c.stream>>>(F, c.mx.WU, c.mx.ufi, c.mx.fl, c.mx.fsig, B.Pf, B.ec, B.tau, part);
  k_loss_grad<<<std::min(G, std::max(1, (F.nfi + RB - 1) / RB)), RB, 0, c.stream>>>(F, B.Pf, B.ec, B.tau, part, G,
                                                                                    B.dl, B.ds, B.jout);
  const int tiles = (F.nfi + 15) / 16;
  nblk = std::min(MLP_BWD_BLOCKS, std::max(1, (tiles + 3) / 4));
  k_mlp_bwd<<<nblk, MB, 0, c.stream>>>(F, m->theta, B.dl, B.ds, m->gpart);
  FVK_HIP(hipGetLastError());
  return 0;
}

uint64_t smix(uint64_t &s) {
  uint64_t z = (s += 0x9E3779B97F4A7C15ULL);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}

}  // namespace

namespace fvk {

int learned_factor(Ctx &c, fvk_model *m) {
  int rc = check_model(m, c);
  if (rc) return rc;
  if ((rc = run_forward(m, c))) return rc;
  return factor_assemble(c);
}

int learned_train(Ctx &c, fvk_model *m, int steps, double lr, double *last_loss) {
  int rc = check_model(m, c);
  if (rc) return rc;
  const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
  LossBufs B;
  for (int k = 0; k < steps; ++k) {
    int nblk = 0;
    if ((rc = run_loss_grad(m, c, B, nblk))) return rc;
    m->step += 1;
    const float c1 = 1.f - std::pow(b1, (float)m->step), c2 = 1.f - std::pow(b2, (float)m->step);
    k_adam<<<(NPARAM + BS - 1) / BS, BS, 0, c.stream>>>(m->theta, m->m1, m->m2, m->gpart, nblk, nullptr, 1,
                                                        (float)lr, b1, b2, eps, c1, c2);
    FVK_HIP(hipGetLastError());
  }
  if (last_loss && steps > 0) {
    double jt[2];
    FVK_HIP(hipMemcpyAsync(jt, B.jout, sizeof jt, hipMemcpyDeviceToHost, c.stream));
    FVK_HIP(hipStreamSynchronize(c.stream));
    *last_loss = jt[0];
  }
  return 0;
}

}  // namespace fvk

#define GUARD_L(x)   \
  do {               \
    int rc_ = (x);   \
    if (rc_) return rc_; \
  } while (0)

extern "C" {

int fvk_model_nparam(void) { return NPARAM; }

int fvk_model_create(int device, uint64_t seed, fvk_model **out) {
  if (!out) { set_error("null argument"); return FVK_ERR_ARG; }
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || device < 0 || device >= ndev) {
    set_error("no HIP device");
    return FVK_ERR_NODEV;
  }
  FVK_HIP(hipSetDevice(device));
  // Xavier-uniform hidden layers, zero output layer (the untrained factor
  // is Jacobi)
  std::vector<float> th(NPARAM, 0.f);
  uint64_t s = seed;
  auto uni = [&]() { return (double)(smix(s) >> 11) * 0x1.0p-53 * 2.0 - 1.0; };
  const double a1 = std::sqrt(6.0 / (NIN + NH)), a2 = std::sqrt(6.0 / (NH + NH));
  for (int k = 0; k < NIN * NH; ++k) th[OFF_W1 + k] = (float)(a1 * uni());
  for (int k = 0; k < NH * NH; ++k) th[OFF_W2 + k] = (float)(a2 * uni());
  fvk_model *m = new fvk_model();
  m->device = device;
  auto al = [&](float **p, size_t cnt) -> int {
    if (hipMalloc(p, cnt * sizeof(float)) != hipSuccess) { set_error("hipMalloc failed"); return FVK_ERR_HIP; }
    if (hipMemset(*p, 0, cnt * sizeof(float)) != hipSuccess) { set_error("hipMemset failed"); return FVK_ERR_HIP; }
    return 0;
  };
  int rc;
  if ((rc = al(&m->theta, NPARAM)) || (rc = al(&m->m1, NPARAM)) || (rc = al(&m->m2, NPARAM)) ||
      (rc = al(&m->grad, NPARAM)) || (rc = al(&m->gpart, (size_t)MLP_BWD_BLOCKS * NPARAM))) {
    fvk_model_destroy(m);
    return rc;
  }
  if (hipMemcpy(m->theta, th.data(), NPARAM * sizeof(float), hipMemcpyHostToDevice) != hipSuccess) {
    fvk_model_destroy(m);
    set_error("hipMemcpy failed");
    return FVK_ERR_HIP;
  }
  *out = m;
  return FVK_OK;
}

void fvk_model_destroy(fvk_model *m) {
  if (!m) return;
  hipSetDevice(m->device);
  hipDeviceSynchronize();
  hipFree(m->theta);
  hipFree(m->m1);
  hipFree(m->m2);
  hipFree(m->grad);
  hipFree(m->gpart);
  delete m;
}

int fvk_model_get(const fvk_model *m, float *theta) {
  if (!m || !theta) { set_error("null argument"); return FVK_ERR_ARG; }
  FVK_HIP(hipSetDevice(m->device));
  FVK_HIP(hipDeviceSynchronize());
  FVK_HIP(hipMemcpy(theta, m->theta, NPARAM * sizeof(float), hipMemcpyDeviceToHost));
  return FVK_OK;
}

int fvk_model_set(fvk_model *m, const float *theta) {
  if (!m || !theta) { set_error("null argument"); return FVK_ERR_ARG; }
  FVK_HIP(hipSetDevice(m->device));
  FVK_HIP(hipDeviceSynchronize());
  FVK_HIP(hipMemcpy(m->theta, theta, NPARAM * sizeof(float), hipMemcpyHostToDevice));
  FVK_HIP(hipMemset(m->m1, 0, NPARAM * sizeof(float)));
  FVK_HIP(hipMemset(m->m2, 0, NPARAM * sizeof(float)));
  m->step = 0;
  return FVK_OK;
}

int fvk_factor_build(fvk_ctx *cc, const fvk_model *mc) {
  if (!cc || !mc) { set_error("null argument"); return FVK_ERR_ARG; }
  return learned_factor(cc->c, const_cast<fvk_model *>(mc));
}

int fvk_factor_fsai(fvk_ctx *cc) {
  if (!cc) { set_error("null argument"); return FVK_ERR_ARG; }
  Ctx &c = cc->c;
  GUARD_L(mix_build(c));
  const FaceIn F = face_in(c);
  k_fsai<<<grid_for(c.n), BS, 0, c.stream>>>(F, c.mx.WU, c.mx.ufi, c.mx.fl, c.mx.fsig);
  FVK_HIP(hipGetLastError());
  return factor_assemble(c);
}

int fvk_factor_features(fvk_ctx *cc, float *X) {
  if (!cc || !X) { set_error("null argument"); return FVK_ERR_ARG; }
  Ctx &c = cc->c;
  GUARD_L(mix_build(c));
  GUARD_L(geo_build(c));
  const FaceIn F = face_in(c);
  // the loss scratch holds 24 n doubles: 12 floats per interior face fit
  if ((size_t)6 * F.nfi > (size_t)24 * c.n) { set_error("feature scratch too small"); return FVK_ERR_STATE; }
  float *d = (float *)c.wk;
  std::vector<float> h((size_t)12 * F.nfi);
  if (F.nfi) {
    k_feat_dump<<<grid_for(F.nfi), BS, 0, c.stream>>>(F, d);
    FVK_HIP(hipGetLastError());
    FVK_HIP(hipMemcpyAsync(h.data(), d, h.size() * sizeof(float), hipMemcpyDeviceToHost, c.stream));
  }
  FVK_HIP(hipStreamSynchronize(c.stream));
  int q = 0;
  for (int f = 0; f < c.nf; ++f) {
    float *o = X + (size_t)12 * f;
    if (c.h_fin[f] >= 0) { for (int k = 0; k < 12; ++k) o[k] = h[(size_t)12 * q + k]; ++q; }
    else { for (int k = 0; k < 12; ++k) o[k] = 0.f; }
  }
  return FVK_OK;
}

int fvk_factor_get(fvk_ctx *cc, float *l, float *sig) {
  if (!cc || !l || !sig) { set_error("null argument"); return FVK_ERR_ARG; }
  Ctx &c = cc->c;
  if (!c.mx.built) { set_error("no factor"); return FVK_ERR_STATE; }
  std::vector<float> fl(c.mx.nfi), fs(c.mx.nfi);
  FVK_HIP(hipStreamSynchronize(c.stream));
  if (c.mx.nfi) {
    FVK_HIP(hipMemcpy(fl.data(), c.mx.fl, fl.size() * sizeof(float), hipMemcpyDeviceToHost));
    FVK_HIP(hipMemcpy(fs.data(), c.mx.fsig, fs.size() * sizeof(float), hipMemcpyDeviceToHost));
  }
  int q = 0;
  for (int f = 0; f < c.nf; ++f) {
    if (c.h_fin[f] >= 0) { l[f] = fl[q]; sig[f] = fs[q]; ++q; }
    else { l[f] = 0.f; sig[f] = 0.f; }
  }
  return FVK_OK;
}

int fvk_model_loss(fvk_model *m, fvk_ctx *cc, double *J, float *grad, float *dl, float *dsig) {
  if (!m || !cc || !J) { set_error("null argument"); return FVK_ERR_ARG; }
  Ctx &c = cc->c;
  GUARD_L(check_model(m, c));
  LossBufs B;
  int nblk = 0;
  GUARD_L(run_loss_grad(m, c, B, nblk));
  k_adam<<<(NPARAM + BS - 1) / BS, BS, 0, c.stream>>>(m->theta, m->m1, m->m2, m->gpart, nblk, m->grad, 0, 0.f, 0.f,
                                                      0.f, 0.f, 1.f, 1.f);
  FVK_HIP(hipGetLastError());
  double jt[2];
  FVK_HIP(hipMemcpyAsync(jt, B.jout, sizeof jt, hipMemcpyDeviceToHost, c.stream));
  FVK_HIP(hipStreamSynchronize(c.stream));
  *J = jt[0];
  if (grad) FVK_HIP(hipMemcpy(grad, m->grad, NPARAM * sizeof(float), hipMemcpyDeviceToHost));
  if (dl || dsig) {
    std::vector<float> a(c.mx.nfi), b(c.mx.nfi);
    if (c.mx.nfi) {
      FVK_HIP(hipMemcpy(a.data(), B.dl, a.size() * sizeof(float), hipMemcpyDeviceToHost));
      FVK_HIP(hipMemcpy(b.data(), B.ds, b.size() * sizeof(float), hipMemcpyDeviceToHost));
    }
    int q = 0;
    for (int f = 0; f < c.nf; ++f) {
      const bool in = c.h_fin[f] >= 0;
      if (dl) dl[f] = in ? a[q] : 0.f;
      if (dsig) dsig[f] = in ? b[q] : 0.f;
      q += in;
    }
  }
  return FVK_OK;
}

int fvk_model_train(fvk_model *m, fvk_ctx *cc, int steps, double lr, double *loss) {
  if (!m || !cc || steps < 0 || !(lr > 0)) { set_error("bad argument"); return FVK_ERR_ARG; }
  Ctx &c = cc->c;
  GUARD_L(check_model(m, c));
  std::vector<double> js(steps);
  const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
  for (int k = 0; k < steps; ++k) {
    LossBufs B;
    int nblk = 0;
    GUARD_L(run_loss_grad(m, c, B, nblk));
    m->step += 1;
    const float c1 = 1.f - std::pow(b1, (float)m->step), c2 = 1.f - std::pow(b2, (float)m->step);
    k_adam<<<(NPARAM + BS - 1) / BS, BS, 0, c.stream>>>(m->theta, m->m1, m->m2, m->gpart, nblk, nullptr, 1,
                                                        (float)lr, b1, b2, eps, c1, c2);
    FVK_HIP(hipGetLastError());
    if (loss) {
      double jt[2];
      FVK_HIP(hipMemcpyAsync(jt, B.jout, sizeof jt, hipMemcpyDeviceToHost, c.stream));
      FVK_HIP(hipStreamSynchronize(c.stream));
      js[k] = jt[0];
    }
  }
  FVK_HIP(hipStreamSynchronize(c.stream));
  if (loss) std::memcpy(loss, js.data(), sizeof(double) * steps);
  return FVK_OK;
}

}  // extern "C"
