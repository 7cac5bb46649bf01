// Connectivity and geometry of a mesh on the GPU: the same arrays, bit for
// bit, as the host's connect() and geometry() (mesh.cpp), which restate
//   grd_conn  spray-simulation/grd_conn.f90:7-238 (face order: the (cell,
//             local face) entries in cell order, the first occurrence of
//             each face kept; owner = the cell of that entry)
//   grd_geom  spray-simulation/grd_geom.f90:9-225
// For the meshes of millions of cells whose host set-up takes seconds.
//
// Matching: each (cell, local face) entry hashes its sorted vertex ids into
// an open-addressing table whose buckets keep the least and the greatest
// entry with that key (atomicMin / atomicMax) and a count; an entry is kept
// when it is its bucket's least.  A bucket with more than two entries (a
// non-manifold face or a 64-bit hash collision) or partners whose vertex
// sets differ make the build fail with an error; the host path is then the
// one to use.  Face numbers and boundary slots are exclusive scans over the
// entries in cell order; each cell gathers its faces in ascending face
// number; geometry keeps the host's expression order (this file is compiled
// without floating-point contraction), and every per-cell sum that the
// reference forms by scattering over faces is gathered over the cell's face
// list, which is in ascending face order: the same additions in the same
// order.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <string>
#include <vector>

#include "mesh.h"

namespace fvk {

namespace {

constexpr int TB = 256;

struct DevElem {
  int n_fce;
  int fv[6][5];
};

// the element face tables (Gambit local vertex numbering, as mesh.cpp)
__constant__ DevElem kDevElem[7] = {
    {0, {{0}}},
    {0, {{0}}},
    {4, {{2, 1, 2}, {2, 2, 3}, {2, 3, 4}, {2, 4, 1}}},
    {3, {{2, 1, 2}, {2, 2, 3}, {2, 3, 1}}},
    {6, {{4, 1, 2, 6, 5}, {4, 2, 4, 8, 6}, {4, 4, 3, 7, 8}, {4, 3, 1, 5, 7}, {4, 2, 1, 3, 4}, {4, 5, 6, 8, 7}}},
    {5, {{4, 1, 2, 5, 4}, {4, 2, 3, 6, 5}, {4, 3, 1, 4, 6}, {3, 1, 3, 2}, {3, 4, 5, 6}}},
    {4, {{3, 2, 1, 3}, {3, 1, 2, 4}, {3, 2, 3, 4}, {3, 3, 1, 4}}},
};

struct Bucket {
  unsigned long long key;  // 0: empty
  int smin, smax, cnt, pad;
};

struct DIn {
  int n_dx, n_vol, n_slot, n_vrt;
  const double *vx;
  const int *ty, *vptr, *vrt, *fptr, *bnd;
};

// vertices of entry s (cell c, local face k), in the element's local order
__device__ int entry_vertices(const DIn &in, const int *slot_cell, int s, int *v) {
  const int c = slot_cell[s], k = s - in.fptr[c];
  const int *f = kDevElem[in.ty[c]].fv[k];
  for (int a = 0; a < f[0]; ++a) v[a] = in.vrt[in.vptr[c] + f[1 + a] - 1];
  return f[0];
}

__device__ void sort_small(int *v, int n) {
  for (int a = 1; a < n; ++a)
    for (int b = a; b > 0 && v[b - 1] > v[b]; --b) { const int t = v[b - 1]; v[b - 1] = v[b]; v[b] = t; }
}

__device__ unsigned long long key_hash(const int *k, int n) {
  unsigned long long h = 0x9E3779B97F4A7C15ULL;
  for (int a = 0; a < n; ++a) {
    h ^= (unsigned long long)(unsigned)k[a] + 0x9E3779B97F4A7C15ULL + (h << 6) + (h >> 2);
    h *= 0xBF58476D1CE4E5B9ULL;
  }
  h ^= h >> 31;
  return h | 1ULL;  // never 0 (the empty mark)
}

__global__ void k_slot_cell(DIn in, int *slot_cell) {
  const int c = blockIdx.x * TB + threadIdx.x;
  if (c >= in.n_vol) return;
  for (int s = in.fptr[c]; s < in.fptr[c + 1]; ++s) slot_cell[s] = c;
}

__global__ void k_table_init(Bucket *tab, size_t cap) {
  const size_t i = blockIdx.x * (size_t)TB + threadIdx.x;
  if (i >= cap) return;
  tab[i].key = 0;
  tab[i].smin = INT_MAX;
  tab[i].smax = -1;
  tab[i].cnt = 0;
}

__global__ void k_insert(DIn in, const int *slot_cell, Bucket *tab, size_t mask, unsigned long long *hs) {
  const int s = blockIdx.x * TB + threadIdx.x;
  if (s >= in.n_slot) return;
  int v[4];
  const int n = entry_vertices(in, slot_cell, s, v);
  sort_small(v, n);
  const unsigned long long h = key_hash(v, n);
  hs[s] = h;
  size_t i = h & mask;
  for (;;) {
    const unsigned long long old = atomicCAS(&tab[i].key, 0ULL, h);
    if (old == 0ULL || old == h) {
      atomicMin(&tab[i].smin, s);
      atomicMax(&tab[i].smax, s);
      atomicAdd(&tab[i].cnt, 1);
      return;
    }
    i = (i + 1) & mask;
  }
}

// kept flag, partner entry (-1: none), boundary flag; err[0] counts buckets
// of more than two entries, err[1] partners whose vertex sets differ
__global__ void k_lookup(DIn in, const int *slot_cell, const Bucket *tab, size_t mask, const unsigned long long *hs,
                         int *kept, int *partner, int *isb, int *err) {
  const int s = blockIdx.x * TB + threadIdx.x;
  if (s >= in.n_slot) return;
  const unsigned long long h = hs[s];
  size_t i = h & mask;
  while (tab[i].key != h) i = (i + 1) & mask;
  const Bucket b = tab[i];
  if (b.cnt > 2) atomicAdd(&err[0], 1);
  int p = -1;
  if (b.cnt == 2) p = s == b.smin ? b.smax : b.smin;
  kept[s] = (s == b.smin) ? 1 : 0;
  partner[s] = p;
  isb[s] = (p < 0) ? 1 : 0;
  if (p >= 0) {
    int v[4], w[4];
    const int n = entry_vertices(in, slot_cell, s, v), m = entry_vertices(in, slot_cell, p, w);
    sort_small(v, n);
    sort_small(w, m);
    bool same = n == m;
    for (int a = 0; a < n && same; ++a) same = v[a] == w[a];
    if (!same) atomicAdd(&err[1], 1);
  }
}

struct DOut {
  int *f_ip, *f_in, *f_ib, *f_bty, *f_nv, *f_vptr, *f_vrt, *f_slot;
  int *fce_lst, *ng_lst;
  double *vol_x, *vol_v, *vol_vr, *f_x, *f_n, *f_dxp, *f_dxn, *f_s, *f_lpn, *f_wfp;
};

// faces from their kept entries: owner, neighbour, boundary index and code,
// vertex count
__global__ void k_faces(DIn in, const int *slot_cell, const int *kept, const int *partner, const int *fidx,
                        const int *bidx, DOut o, int *err) {
  const int s = blockIdx.x * TB + threadIdx.x;
  if (s >= in.n_slot || !kept[s]) return;
  const int g = fidx[s], c = slot_cell[s], p = partner[s];
  o.f_ip[g] = c;
  o.f_in[g] = p >= 0 ? slot_cell[p] : -1;
  o.f_slot[g] = s;
  const int k = s - in.fptr[c];
  o.f_nv[g] = kDevElem[in.ty[c]].fv[k][0];
  const int code = in.bnd[s];
  if (p < 0) {
    if (code <= 0) atomicAdd(&err[2], 1);  // boundary face without a boundary code
    o.f_bty[g] = code;
    o.f_ib[g] = bidx[s];
  } else {
    if (code > 0 || in.bnd[p] > 0) atomicAdd(&err[3], 1);  // a code on an interior face
    o.f_bty[g] = 0;
    o.f_ib[g] = -1;
  }
}

__global__ void k_face_vrt(DIn in, const int *slot_cell, int nf, DOut o) {
  const int g = blockIdx.x * TB + threadIdx.x;
  if (g >= nf) return;
  int v[4];
  const int n = entry_vertices(in, slot_cell, o.f_slot[g], v);
  for (int a = 0; a < n; ++a) o.f_vrt[o.f_vptr[g] + a] = v[a];
}

// per cell: its faces in ascending face number (grd_conn.f90:128-152), the
// neighbour per slot (c_neig_lst, :218-238), the centre (vertex mean)
__global__ void k_cells(DIn in, const int *kept, const int *partner, const int *fidx, DOut o) {
  const int c = blockIdx.x * TB + threadIdx.x;
  if (c >= in.n_vol) return;
  const int a = in.fptr[c], b = in.fptr[c + 1];
  int f[6];
  for (int s = a; s < b; ++s) f[s - a] = kept[s] ? fidx[s] : fidx[partner[s]];
  sort_small(f, b - a);
  for (int s = a; s < b; ++s) {
    const int g = f[s - a];
    o.fce_lst[s] = g;
    const int ip = o.f_ip[g], in_ = o.f_in[g];
    o.ng_lst[s] = ip != c ? ip : (in_ >= 0 && in_ != c ? in_ : -1);
  }
  double acc[3] = {0.0, 0.0, 0.0};
  const int v0 = in.vptr[c], v1 = in.vptr[c + 1];
  for (int k = v0; k < v1; ++k)
    for (int d = 0; d < in.n_dx; ++d) acc[d] = acc[d] + in.vx[3 * (size_t)in.vrt[k] + d];
  for (int d = 0; d < in.n_dx; ++d) o.vol_x[3 * (size_t)c + d] = acc[d] / (double)(v1 - v0);
}

__device__ double dmag(const double *v, int n) {
  double s = 0.0;
  for (int d = 0; d < n; ++d) s = s + v[d] * v[d];
  return sqrt(s);
}
__device__ double ddot(const double *a, const double *b, int n) {
  double s = 0.0;
  for (int d = 0; d < n; ++d) s = s + a[d] * b[d];
  return s;
}
__device__ void dcross(const double *a, const double *b, double *c) {
  const int c1[3] = {1, 2, 0}, c2[3] = {2, 0, 1};
  double r[3];
  for (int d = 0; d < 3; ++d) r[d] = a[c1[d]] * b[c2[d]] - b[c1[d]] * a[c2[d]];
  c[0] = r[0]; c[1] = r[1]; c[2] = r[2];
}

// per face (grd_geom.f90:71-210, the host's geometry() loop body)
__global__ void k_face_geom(DIn in, int nf, DOut o) {
  const int g = blockIdx.x * TB + threadIdx.x;
  if (g >= nf) return;
  const int nd = in.n_dx, ip = o.f_ip[g];
  double xf[3] = {0, 0, 0}, n[3] = {0, 0, 0}, sg;
  const int *fv = o.f_vrt + o.f_vptr[g];
  const int nfv = o.f_vptr[g + 1] - o.f_vptr[g];
  if (nd == 2) {
    const double *x1 = &in.vx[3 * (size_t)fv[0]], *x2 = &in.vx[3 * (size_t)fv[1]];
    double e[3] = {x2[0] - x1[0], x2[1] - x1[1], 0.0}, z[3] = {0.0, 0.0, 1.0}, cr[3];
    const double s0 = dmag(e, 3);
    dcross(e, z, cr);
    const double mg = dmag(cr, 2);
    n[0] = cr[0] / mg;
    n[1] = cr[1] / mg;
    sg = s0;
    xf[0] = 0.5 * (x1[0] + x2[0]);
    xf[1] = 0.5 * (x1[1] + x2[1]);
  } else {
    double vec[5][3] = {{0}}, s0 = 0.0;
    const double *x1 = &in.vx[3 * (size_t)fv[0]];
    for (int k = 2; k <= nfv; ++k) {
      const double *x2 = &in.vx[3 * (size_t)fv[k - 1]];
      for (int d = 0; d < 3; ++d) vec[k][d] = x2[d] - x1[d];
      if (k > 2) {
        double cr[3];
        dcross(vec[k - 1], vec[k], cr);
        for (int d = 0; d < 3; ++d) vec[1][d] = 0.5 * cr[d];
        const double s1 = dmag(vec[1], 3);
        for (int d = 0; d < 3; ++d) vec[0][d] = vec[0][d] + vec[1][d] / s1;
        s0 = s0 + s1;
      }
    }
    for (int d = 0; d < 3; ++d) n[d] = vec[0][d] / (double)(nfv - 2);
    sg = s0;
    double acc[3] = {0, 0, 0};
    for (int k = 0; k < nfv; ++k)
      for (int d = 0; d < 3; ++d) acc[d] = acc[d] + in.vx[3 * (size_t)fv[k] + d];
    for (int d = 0; d < 3; ++d) xf[d] = acc[d] / (double)nfv;
  }
  for (int d = 0; d < 3; ++d) { o.f_x[3 * (size_t)g + d] = xf[d]; o.f_n[3 * (size_t)g + d] = n[d]; }
  o.f_s[g] = sg;
  const double *xP = &o.vol_x[3 * (size_t)ip];
  double t[3], xpa[3] = {0, 0, 0};
  for (int d = 0; d < nd; ++d) t[d] = xf[d] - xP[d];
  const double dd = ddot(t, n, nd);
  for (int d = 0; d < nd; ++d) xpa[d] = xf[d] - n[d] * dd;
  for (int d = 0; d < 3; ++d) o.f_dxp[3 * (size_t)g + d] = d < nd ? xpa[d] - xP[d] : 0.0;
  const int in_ = o.f_in[g];
  double dfp[3];
  for (int d = 0; d < nd; ++d) dfp[d] = xf[d] - xpa[d];
  if (in_ >= 0) {
    const double *xN = &o.vol_x[3 * (size_t)in_];
    double xna[3], dpn[3];
    for (int d = 0; d < nd; ++d) t[d] = xf[d] - xN[d];
    const double dn = ddot(t, n, nd);
    for (int d = 0; d < nd; ++d) xna[d] = xf[d] - n[d] * dn;
    for (int d = 0; d < 3; ++d) o.f_dxn[3 * (size_t)g + d] = d < nd ? xna[d] - xN[d] : 0.0;
    for (int d = 0; d < nd; ++d) dpn[d] = xna[d] - xpa[d];
    o.f_lpn[g] = dmag(dpn, nd);
    o.f_wfp[g] = dmag(dfp, nd) / o.f_lpn[g];
  } else {
    for (int d = 0; d < 3; ++d) o.f_dxn[3 * (size_t)g + d] = 0.0;
    o.f_lpn[g] = dmag(dfp, nd);
    o.f_wfp[g] = 0.0;
  }
}

// volumes by the x-flux, gathered over each cell's faces in ascending face
// number (grd_geom.f90:151-153,172-173 scatter them in face order)
__global__ void k_cell_vol(DIn in, DOut o) {
  const int c = blockIdx.x * TB + threadIdx.x;
  if (c >= in.n_vol) return;
  double v = 0.0;
  for (int s = in.fptr[c]; s < in.fptr[c + 1]; ++s) {
    const int g = o.fce_lst[s];
    const double vp = o.f_x[3 * (size_t)g] * o.f_s[g] * o.f_n[3 * (size_t)g];
    v = o.f_ip[g] == c ? v + vp : v - vp;
  }
  o.vol_v[c] = v;
  o.vol_vr[c] = 1.0 / v;
}

template <class T>
struct DBuf {
  T *p = nullptr;
  ~DBuf() { if (p) (void)hipFree(p); }
  hipError_t alloc(size_t n) { return hipMalloc(&p, (n ? n : 1) * sizeof(T)); }
};

#define MD_HIP(call)                                                      \
  do {                                                                    \
    hipError_t e_ = (call);                                               \
    if (e_ != hipSuccess) return std::string(#call) + ": " + hipGetErrorString(e_); \
  } while (0)

template <class T>
std::string up(DBuf<T> &d, const std::vector<T> &v) {
  MD_HIP(d.alloc(v.size()));
  if (!v.empty()) MD_HIP(hipMemcpy(d.p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice));
  return "";
}

template <class T>
std::string down(std::vector<T> &v, const T *d, size_t n) {
  v.resize(n);
  if (n) MD_HIP(hipMemcpy(v.data(), d, n * sizeof(T), hipMemcpyDeviceToHost));
  return "";
}

std::string exclusive_scan(const int *in, int *out, int n) {
  size_t tmp = 0;
  MD_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, tmp, in, out, n));
  DBuf<char> t;
  MD_HIP(t.alloc(tmp));
  MD_HIP(hipcub::DeviceScan::ExclusiveSum(t.p, tmp, in, out, n));
  return "";
}

inline int blocks(size_t n) { return (int)((n + TB - 1) / TB); }

}  // namespace

std::string finish_device(Mesh &m, int device) {
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || device < 0 || device >= ndev) return "no HIP device";
  MD_HIP(hipSetDevice(device));
  const int nv = m.n_vol, ns = m.n_slot;
  if (ns == 0) return "mesh has no faces";
  DBuf<double> vx;
  DBuf<int> ty, vptr, vrt, fptr, bnd;
  std::string e;
  if (!(e = up(vx, m.vx)).empty() || !(e = up(ty, m.v_ty)).empty() || !(e = up(vptr, m.v_vptr)).empty() ||
      !(e = up(vrt, m.v_vrt)).empty() || !(e = up(fptr, m.v_fptr)).empty() || !(e = up(bnd, m.v_bnd)).empty())
    return e;
  DIn in{m.n_dx, nv, ns, m.n_vrt, vx.p, ty.p, vptr.p, vrt.p, fptr.p, bnd.p};
  DBuf<int> slot_cell, kept, partner, isb, fidx, bidx, err;
  DBuf<unsigned long long> hs;
  DBuf<Bucket> tab;
  size_t cap = 1;
  while (cap < 2 * (size_t)ns + 2) cap <<= 1;
  MD_HIP(slot_cell.alloc(ns)); MD_HIP(kept.alloc(ns)); MD_HIP(partner.alloc(ns)); MD_HIP(isb.alloc(ns));
  MD_HIP(fidx.alloc(ns)); MD_HIP(bidx.alloc(ns)); MD_HIP(err.alloc(4)); MD_HIP(hs.alloc(ns)); MD_HIP(tab.alloc(cap));
  MD_HIP(hipMemset(err.p, 0, 4 * sizeof(int)));
  k_slot_cell<<<blocks(nv), TB>>>(in, slot_cell.p);
  k_table_init<<<blocks(cap), TB>>>(tab.p, cap);
  k_insert<<<blocks(ns), TB>>>(in, slot_cell.p, tab.p, cap - 1, hs.p);
  k_lookup<<<blocks(ns), TB>>>(in, slot_cell.p, tab.p, cap - 1, hs.p, kept.p, partner.p, isb.p, err.p);
  MD_HIP(hipGetLastError());
  if (!(e = exclusive_scan(kept.p, fidx.p, ns)).empty() || !(e = exclusive_scan(isb.p, bidx.p, ns)).empty()) return e;
  int last[2], errh[4];
  MD_HIP(hipMemcpy(&last[0], fidx.p + ns - 1, sizeof(int), hipMemcpyDeviceToHost));
  MD_HIP(hipMemcpy(&last[1], kept.p + ns - 1, sizeof(int), hipMemcpyDeviceToHost));
  MD_HIP(hipMemcpy(errh, err.p, sizeof errh, hipMemcpyDeviceToHost));
  if (errh[0] || errh[1]) return "device connectivity: a face key is shared by more than two entries or collides";
  const int nf = last[0] + last[1];
  if (nf != m.n_fce) return "mesh is not conforming: face count differs from the boundary listing";
  DBuf<int> f_ip, f_in, f_ib, f_bty, f_nv, f_vptr, f_vrt, f_slot, fce_lst, ng_lst;
  DBuf<double> vol_x, vol_v, vol_vr, f_x, f_n, f_dxp, f_dxn, f_s, f_lpn, f_wfp;
  MD_HIP(f_ip.alloc(nf)); MD_HIP(f_in.alloc(nf)); MD_HIP(f_ib.alloc(nf)); MD_HIP(f_bty.alloc(nf));
  MD_HIP(f_nv.alloc(nf + 1)); MD_HIP(f_vptr.alloc(nf + 1)); MD_HIP(f_slot.alloc(nf));
  MD_HIP(fce_lst.alloc(ns)); MD_HIP(ng_lst.alloc(ns));
  MD_HIP(vol_x.alloc(3 * (size_t)nv)); MD_HIP(vol_v.alloc(nv)); MD_HIP(vol_vr.alloc(nv));
  MD_HIP(f_x.alloc(3 * (size_t)nf)); MD_HIP(f_n.alloc(3 * (size_t)nf)); MD_HIP(f_dxp.alloc(3 * (size_t)nf));
  MD_HIP(f_dxn.alloc(3 * (size_t)nf)); MD_HIP(f_s.alloc(nf)); MD_HIP(f_lpn.alloc(nf)); MD_HIP(f_wfp.alloc(nf));
  MD_HIP(hipMemset(vol_x.p, 0, 3 * (size_t)nv * sizeof(double)));
  MD_HIP(hipMemset(f_nv.p, 0, (nf + 1) * sizeof(int)));
  DOut o{f_ip.p, f_in.p, f_ib.p, f_bty.p, f_nv.p, f_vptr.p, nullptr, f_slot.p, fce_lst.p, ng_lst.p,
         vol_x.p, vol_v.p, vol_vr.p, f_x.p, f_n.p, f_dxp.p, f_dxn.p, f_s.p, f_lpn.p, f_wfp.p};
  k_faces<<<blocks(ns), TB>>>(in, slot_cell.p, kept.p, partner.p, fidx.p, bidx.p, o, err.p);
  MD_HIP(hipGetLastError());
  if (!(e = exclusive_scan(f_nv.p, f_vptr.p, nf + 1)).empty()) return e;
  int nfv = 0;
  MD_HIP(hipMemcpy(&nfv, f_vptr.p + nf, sizeof(int), hipMemcpyDeviceToHost));
  MD_HIP(f_vrt.alloc(nfv));
  o.f_vrt = f_vrt.p;
  k_face_vrt<<<blocks(nf), TB>>>(in, slot_cell.p, nf, o);
  k_cells<<<blocks(nv), TB>>>(in, kept.p, partner.p, fidx.p, o);
  k_face_geom<<<blocks(nf), TB>>>(in, nf, o);
  k_cell_vol<<<blocks(nv), TB>>>(in, o);
  MD_HIP(hipGetLastError());
  MD_HIP(hipDeviceSynchronize());
  MD_HIP(hipMemcpy(errh, err.p, sizeof errh, hipMemcpyDeviceToHost));
  if (errh[2]) return "boundary face without a boundary code";
  if (errh[3]) return "boundary code on an interior face";
  int nb = 0;
  MD_HIP(hipMemcpy(&nb, bidx.p + ns - 1, sizeof(int), hipMemcpyDeviceToHost));
  int lastb = 0;
  MD_HIP(hipMemcpy(&lastb, isb.p + ns - 1, sizeof(int), hipMemcpyDeviceToHost));
  m.n_bf = nb + lastb;
  if (!(e = down(m.f_ip, f_ip.p, nf)).empty() || !(e = down(m.f_in, f_in.p, nf)).empty() ||
      !(e = down(m.f_ib, f_ib.p, nf)).empty() || !(e = down(m.f_bty, f_bty.p, nf)).empty() ||
      !(e = down(m.f_vptr, f_vptr.p, nf + 1)).empty() || !(e = down(m.f_vrt, f_vrt.p, nfv)).empty() ||
      !(e = down(m.fce_lst, fce_lst.p, ns)).empty() || !(e = down(m.ng_lst, ng_lst.p, ns)).empty() ||
      !(e = down(m.vol_x, vol_x.p, 3 * (size_t)nv)).empty() || !(e = down(m.vol_v, vol_v.p, nv)).empty() ||
      !(e = down(m.vol_vr, vol_vr.p, nv)).empty() || !(e = down(m.f_x, f_x.p, 3 * (size_t)nf)).empty() ||
      !(e = down(m.f_n, f_n.p, 3 * (size_t)nf)).empty() || !(e = down(m.f_dxp, f_dxp.p, 3 * (size_t)nf)).empty() ||
      !(e = down(m.f_dxn, f_dxn.p, 3 * (size_t)nf)).empty() || !(e = down(m.f_s, f_s.p, nf)).empty() ||
      !(e = down(m.f_lpn, f_lpn.p, nf)).empty() || !(e = down(m.f_wfp, f_wfp.p, nf)).empty())
    return e;
  return "";
}

}  // namespace fvk
