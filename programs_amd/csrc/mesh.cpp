// Host mesh layer: Gambit neutral reader, box generator, O(N) connectivity
// and finite-volume geometry.  Produces exactly the arrays the reference's
// grd_read / grd_conn / grd_geom leave in its grd structure:
//   grd_read  spray-simulation/grd_read.f90:9-128
//   grd_conn  spray-simulation/grd_conn.f90:7-238 (O(n^2) matching replaced
//             by a hash on sorted vertex tuples; same face order)
//   grd_geom  spray-simulation/grd_geom.f90:9-225
// Floating-point expressions keep the reference's evaluation order; this file
// is compiled with -ffp-contract=off.
#include "mesh.h"

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <unordered_map>
#include <algorithm>

namespace fvk {

// Element face tables (Gambit local vertex numbering, gambit.vrt:4-47).  The
// wedge's third side is 3,1,4,6: the reference table lists vertex 0 there
// (gambit.vrt:30), which cannot describe a face.
struct Elem {
  int n_dx, n_vrt, n_fce;
  int fv[6][5];  // fv[k][0] = vertex count, then 1-based local vertices
};
static const Elem kElem[7] = {
    {0, 0, 0, {{0}}},
    {0, 0, 0, {{0}}},
    {2, 4, 4, {{2, 1, 2}, {2, 2, 3}, {2, 3, 4}, {2, 4, 1}}},
    {2, 3, 3, {{2, 1, 2}, {2, 2, 3}, {2, 3, 1}}},
    {3, 8, 6, {{4, 1, 2, 6, 5}, {4, 2, 4, 8, 6}, {4, 4, 3, 7, 8}, {4, 3, 1, 5, 7}, {4, 2, 1, 3, 4}, {4, 5, 6, 8, 7}}},
    {3, 6, 5, {{4, 1, 2, 5, 4}, {4, 2, 3, 6, 5}, {4, 3, 1, 4, 6}, {3, 1, 3, 2}, {3, 4, 5, 6}}},
    {3, 4, 4, {{3, 2, 1, 3}, {3, 1, 2, 4}, {3, 2, 3, 4}, {3, 3, 1, 4}}},
};

int elem_nfce(int ty) { return (ty >= 2 && ty <= 6) ? kElem[ty].n_fce : 0; }

static int elem_face(int ty, int k, int *loc) {
  const int *f = kElem[ty].fv[k];
  for (int a = 0; a < f[0]; ++a) loc[a] = f[1 + a];
  return f[0];
}

static void cell_face_vertices(const Mesh &m, int c, int k, int *out, int *n) {
  int loc[5];
  *n = elem_face(m.v_ty[c], k, loc);
  for (int a = 0; a < *n; ++a) out[a] = m.v_vrt[m.v_vptr[c] + loc[a] - 1];
}

// Same vertex set (grd_conn.f90:161-181 on conforming meshes).
static bool same_face(const int *a, int na, const int *b, int nb) {
  if (na != nb) return false;
  long sa = 0, sb = 0;
  for (int k = 0; k < na; ++k) { sa += a[k]; sb += b[k]; }
  if (sa != sb) return false;
  for (int i = 0; i < na; ++i) {
    bool hit = false;
    for (int j = 0; j < na && !hit; ++j) hit = a[i] == b[j];
    if (!hit) return false;
  }
  return true;
}

std::string init_cells(Mesh &m, int n_dx, int n_vrt, const double *vx3, int n_vol, const int *v_ty,
                       const int *v_vptr, const int *v_vrt1, const int *bnd) {
  m.n_dx = n_dx;
  m.n_vrt = n_vrt;
  m.n_vol = n_vol;
  m.vx.assign(vx3, vx3 + 3 * (size_t)n_vrt);
  m.v_ty.assign(v_ty, v_ty + n_vol);
  m.v_vptr.assign(v_vptr, v_vptr + n_vol + 1);
  m.v_vrt.resize(v_vptr[n_vol]);
  for (int k = 0; k < v_vptr[n_vol]; ++k) m.v_vrt[k] = v_vrt1[k] - 1;
  m.v_fptr.resize(n_vol + 1);
  m.v_fptr[0] = 0;
  m.max_fce = 0;
  for (int c = 0; c < n_vol; ++c) {
    int nf = elem_nfce(v_ty[c]);
    if (nf == 0) return "unsupported element type " + std::to_string(v_ty[c]);
    if (kElem[v_ty[c]].n_dx != n_dx) return "element dimension does not match n_dx";
    if (v_vptr[c + 1] - v_vptr[c] != kElem[v_ty[c]].n_vrt) return "vertex count does not match element type";
    m.v_fptr[c + 1] = m.v_fptr[c] + nf;
    if (nf > m.max_fce) m.max_fce = nf;
  }
  m.n_slot = m.v_fptr[n_vol];
  m.v_bnd.assign(bnd, bnd + m.n_slot);
  for (int k = 0; k < v_vptr[n_vol]; ++k)
    if (m.v_vrt[k] < 0 || m.v_vrt[k] >= n_vrt) return "vertex id out of range";
  int nb = 0;
  for (int s = 0; s < m.n_slot; ++s) nb += m.v_bnd[s] > 0;
  if ((m.n_slot + nb) % 2) return "face count is odd: boundary list does not close the mesh";
  m.n_fce = (m.n_slot + nb) / 2;  // grd_read.f90:94-99
  return "";
}

// ---- gmsh 2.2 ASCII reader ----
// The parse follows flowlab's read_gmsh_mesh (flowlab/src/mesh.cpp:39-187,
// serial mode): whitespace-separated tokens; "$Nodes": count, then id x y z
// per node; "$Elements": count, then id, type, tag count, the tags (when a
// fourth tag, the partition count, is below 1 only three tags are kept) and
// the type's vertices (0, 2, 3, 4, 4, 8, 6, 5 for types 0-7).  Elements of
// dimension n_dx become cells, in file order; elements of dimension n_dx - 1
// mark boundary faces, whose code is their physical tag (the first tag)
// mapped through code_map (pairs physical -> code; unmapped tags keep their
// value).  gmsh vertex orders are converted to the Gambit orders of the
// element face table: the hexahedron's 1 2 3 4 5 6 7 8 become 1 2 4 3 5 6 8
// 7 (Gambit numbers a brick lexicographically); quadrilateral, triangle,
// prism and tetrahedron agree.  Pyramids are not supported.
std::string parse_gmsh(GmshRaw &g, const char *path) {
  FILE *fp = std::fopen(path, "r");
  if (!fp) return std::string("cannot open ") + path;
  static const int kNvrt[8] = {0, 2, 3, 4, 4, 8, 6, 5};
  std::vector<int> &node_id = g.node_id, &e_idx = g.e_idx, &e_ty = g.e_ty, &e_tag0 = g.e_tag0,
                   &e_vptr = g.e_vptr, &e_vrt = g.e_vrt;
  std::vector<double> &node_x = g.node_x;
  e_vptr.assign(1, 0);
  char buf[256];
  std::string err;
  bool have_nodes = false, have_elems = false;
  while (err.empty() && std::fscanf(fp, "%255s", buf) == 1) {
    if (!std::strcmp(buf, "$Nodes")) {
      int n = 0;
      if (std::fscanf(fp, "%d", &n) != 1 || n < 0) { err = "bad $Nodes count"; break; }
      node_id.resize(n);
      node_x.resize(3 * (size_t)n);
      for (int i = 0; i < n; ++i)
        if (std::fscanf(fp, "%d %lg %lg %lg", &node_id[i], &node_x[3 * (size_t)i], &node_x[3 * (size_t)i + 1],
                        &node_x[3 * (size_t)i + 2]) != 4) {
          err = "bad $Nodes record";
          break;
        }
      have_nodes = true;
    } else if (!std::strcmp(buf, "$Elements")) {
      int n = 0;
      if (std::fscanf(fp, "%d", &n) != 1 || n < 0) { err = "bad $Elements count"; break; }
      for (int i = 0; i < n && err.empty(); ++i) {
        int idx, ty, nt;
        if (std::fscanf(fp, "%d %d %d", &idx, &ty, &nt) != 3 || ty < 0 || ty > 7 || nt < 0 || nt > 11) {
          err = "bad $Elements record";
          break;
        }
        int tag[12] = {0};
        for (int j = 0; j < nt; ++j)
          if (std::fscanf(fp, "%d", &tag[j]) != 1) { err = "bad element tags"; break; }
        const int nv = kNvrt[ty];
        int v[8];
        for (int j = 0; j < nv; ++j)
          if (std::fscanf(fp, "%d", &v[j]) != 1) { err = "bad element vertices"; break; }
        if (!err.empty()) break;
        e_idx.push_back(idx);
        e_ty.push_back(ty);
        e_tag0.push_back(nt > 0 ? tag[0] : 0);
        for (int j = 0; j < nv; ++j) e_vrt.push_back(v[j]);
        e_vptr.push_back((int)e_vrt.size());
      }
      have_elems = true;
    }
  }
  std::fclose(fp);
  if (!err.empty()) return err;
  if (!have_nodes || !have_elems) return "no $Nodes or $Elements section";
  return "";
}

std::string read_gmsh(Mesh &m, const char *path, int n_dx, double l_sc, const int *code_map, int n_map) {
  static const int kDim[8] = {0, 1, 2, 2, 3, 3, 3, 3};
  static const int kGambit[8] = {0, 0, 3, 2, 6, 4, 5, 0};  // gmsh type -> Gambit NTYPE
  GmshRaw g;
  {
    std::string err = parse_gmsh(g, path);
    if (!err.empty()) return err;
  }
  const std::vector<int> &node_id = g.node_id, &e_idx = g.e_idx, &e_ty = g.e_ty, &e_tag0 = g.e_tag0,
                         &e_vptr = g.e_vptr, &e_vrt = g.e_vrt;
  const std::vector<double> &node_x = g.node_x;
  // node labels -> positions
  std::unordered_map<int, int> pos;
  pos.reserve(node_id.size() * 2);
  for (size_t i = 0; i < node_id.size(); ++i)
    if (!pos.emplace(node_id[i], (int)i).second) return "duplicate node id " + std::to_string(node_id[i]);
  if (n_dx <= 0) {
    n_dx = 2;
    for (int t : e_ty) if (kDim[t] == 3) n_dx = 3;
  }
  if (n_dx != 2 && n_dx != 3) return "n_dx must be 2 or 3";
  // boundary faces: sorted vertex positions -> code
  std::unordered_map<uint64_t, std::vector<std::pair<std::vector<int>, int>>> bmap;
  auto key_of = [](std::vector<int> &k) {
    std::sort(k.begin(), k.end());
    uint64_t h = 0x9E3779B97F4A7C15ULL;
    for (int a : k) { h ^= (uint64_t)(uint32_t)a + 0x9E3779B97F4A7C15ULL + (h << 6) + (h >> 2); h *= 0xBF58476D1CE4E5B9ULL; }
    return h;
  };
  std::vector<int> cty, cvptr{0}, cvrt, cidx;
  for (size_t e = 0; e < e_ty.size(); ++e) {
    const int t = e_ty[e], a = e_vptr[e], b = e_vptr[e + 1];
    std::vector<int> vv;
    for (int k = a; k < b; ++k) {
      auto it = pos.find(e_vrt[k]);
      if (it == pos.end()) return "element " + std::to_string(e_idx[e]) + " names an unknown node";
      vv.push_back(it->second);
    }
    if (kDim[t] == n_dx) {
      if (!kGambit[t]) return "unsupported cell type " + std::to_string(t) + " (element " + std::to_string(e_idx[e]) + ")";
      static const int kHex[8] = {0, 1, 3, 2, 4, 5, 7, 6};
      cty.push_back(kGambit[t]);
      for (int k = 0; k < (int)vv.size(); ++k) cvrt.push_back(1 + (t == 5 ? vv[kHex[k]] : vv[k]));
      cvptr.push_back((int)cvrt.size());
      cidx.push_back(e_idx[e]);
    } else if (kDim[t] == n_dx - 1) {
      int code = e_tag0[e];
      for (int q = 0; q < n_map; ++q)
        if (code_map[2 * q] == code) { code = code_map[2 * q + 1]; break; }
      std::vector<int> k = vv;
      const uint64_t h = key_of(k);
      bmap[h].push_back({k, code});
    }
  }
  const int n_vol = (int)cty.size();
  if (!n_vol) return "no cells of dimension " + std::to_string(n_dx);
  // boundary code of each (cell, local face)
  std::vector<int> bnd;
  for (int c = 0; c < n_vol; ++c) {
    const int ty = cty[c];
    for (int k = 0; k < kElem[ty].n_fce; ++k) {
      int loc[5];
      const int nv = elem_face(ty, k, loc);
      std::vector<int> key(nv);
      for (int a = 0; a < nv; ++a) key[a] = cvrt[cvptr[c] + loc[a] - 1] - 1;
      const uint64_t h = key_of(key);
      int code = 0;
      auto it = bmap.find(h);
      if (it != bmap.end())
        for (auto &pr : it->second)
          if (pr.first == key) { code = pr.second; break; }
      bnd.push_back(code);
    }
  }
  std::vector<double> vx3(3 * node_id.size(), 0.0);
  for (size_t i = 0; i < node_id.size(); ++i)
    for (int d = 0; d < n_dx; ++d) vx3[3 * i + d] = node_x[3 * i + d] * l_sc;
  std::string e = init_cells(m, n_dx, (int)node_id.size(), vx3.data(), n_vol, cty.data(), cvptr.data(),
                             cvrt.data(), bnd.data());
  if (!e.empty()) return e;
  m.elm_idx = cidx;
  return "";
}

// ---- grd_conn ----
namespace {
struct Entry { int key[4]; int n; int idx; };
uint64_t mix(const int *k, int n) {
  uint64_t h = 0x9E3779B97F4A7C15ULL;
  for (int a = 0; a < n; ++a) {
    h ^= (uint64_t)(uint32_t)k[a] + 0x9E3779B97F4A7C15ULL + (h << 6) + (h >> 2);
    h *= 0xBF58476D1CE4E5B9ULL;
  }
  return h ^ (h >> 31);
}
}  // namespace

std::string connect(Mesh &m) {
  const int ns = m.n_slot;
  // (cell, local face) entries in reference order; partner of each entry
  std::vector<int> fv(4 * (size_t)ns), fn(ns), owner(ns), partner(ns, -1);
  for (int c = 0; c < m.n_vol; ++c)
    for (int s = m.v_fptr[c]; s < m.v_fptr[c + 1]; ++s) {
      cell_face_vertices(m, c, s - m.v_fptr[c], &fv[4 * (size_t)s], &fn[s]);
      owner[s] = c;
    }
  size_t cap = 1;
  while (cap < 2 * (size_t)ns + 2) cap <<= 1;
  std::vector<Entry> H(cap);
  for (auto &e : H) e.idx = -1;
  std::vector<char> copy(ns, 0);
  for (int s = 0; s < ns; ++s) {
    int key[4], kn = fn[s];
    for (int a = 0; a < kn; ++a) key[a] = fv[4 * (size_t)s + a];
    for (int a = 1; a < kn; ++a)
      for (int b = a; b > 0 && key[b - 1] > key[b]; --b) std::swap(key[b - 1], key[b]);
    size_t h = mix(key, kn) & (cap - 1);
    for (;;) {
      Entry &e = H[h];
      if (e.idx < 0) {
        e.idx = s; e.n = kn;
        std::memcpy(e.key, key, sizeof key);
        break;
      }
      if (e.n == kn && !std::memcmp(e.key, key, sizeof(int) * kn)) {
        int f = e.idx;
        if (partner[f] < 0) {  // first occurrence keeps the face
          partner[f] = s;
          partner[s] = f;
          copy[s] = 1;
        }
        break;
      }
      h = (h + 1) & (cap - 1);
    }
  }
  // kept faces in entry order (grd_conn.f90:100-113)
  int nf = 0;
  for (int s = 0; s < ns; ++s) nf += !copy[s];
  if (nf != m.n_fce) return "mesh is not conforming: face count differs from the boundary listing";
  m.f_ip.resize(nf); m.f_in.resize(nf); m.f_vptr.resize(nf + 1);
  m.f_vrt.clear();
  m.f_vrt.reserve(4 * (size_t)nf);
  int f = 0;
  m.f_vptr[0] = 0;
  for (int s = 0; s < ns; ++s) {
    if (copy[s]) continue;
    m.f_ip[f] = owner[s];
    m.f_in[f] = partner[s] >= 0 ? owner[partner[s]] : -1;
    for (int a = 0; a < fn[s]; ++a) m.f_vrt.push_back(fv[4 * (size_t)s + a]);
    m.f_vptr[f + 1] = (int)m.f_vrt.size();
    ++f;
  }
  // per-cell face lists in ascending face index (grd_conn.f90:128-152)
  m.fce_lst.assign(ns, -1);
  std::vector<int> fill(m.n_vol, 0);
  for (int g = 0; g < nf; ++g) {
    int ip = m.f_ip[g], in = m.f_in[g];
    m.fce_lst[m.v_fptr[ip] + fill[ip]++] = g;
    if (in >= 0) m.fce_lst[m.v_fptr[in] + fill[in]++] = g;
  }
  for (int c = 0; c < m.n_vol; ++c)
    if (fill[c] != m.v_fptr[c + 1] - m.v_fptr[c]) return "cell with unmatched, unlisted face";
  m.ng_lst.assign(ns, -1);  // c_neig_lst, grd_conn.f90:218-238
  for (int c = 0; c < m.n_vol; ++c)
    for (int s = m.v_fptr[c]; s < m.v_fptr[c + 1]; ++s) {
      int g = m.fce_lst[s];
      if (m.f_ip[g] != c) m.ng_lst[s] = m.f_ip[g];
      else if (m.f_in[g] >= 0 && m.f_in[g] != c) m.ng_lst[s] = m.f_in[g];
    }
  return "";
}

// ---- grd_geom ----
static inline double mag(const double *v, int n) {
  double s = 0.0;
  for (int d = 0; d < n; ++d) s = s + v[d] * v[d];
  return std::sqrt(s);
}
static inline double dot(const double *a, const double *b, int n) {
  double s = 0.0;
  for (int d = 0; d < n; ++d) s = s + a[d] * b[d];
  return s;
}
static inline void cross(const double *a, const double *b, double *c) {
  static const int c1[3] = {1, 2, 0}, c2[3] = {2, 0, 1};
  double r[3];
  for (int d = 0; d < 3; ++d) r[d] = a[c1[d]] * b[c2[d]] - b[c1[d]] * a[c2[d]];
  c[0] = r[0]; c[1] = r[1]; c[2] = r[2];
}

std::string geometry(Mesh &m) {
  const int nd = m.n_dx, nv = m.n_vol, nf = m.n_fce;
  m.vol_x.assign(3 * (size_t)nv, 0.0);
  m.vol_v.assign(nv, 0.0);
  m.vol_vr.assign(nv, 0.0);
  m.f_x.assign(3 * (size_t)nf, 0.0); m.f_n.assign(3 * (size_t)nf, 0.0);
  m.f_dxp.assign(3 * (size_t)nf, 0.0); m.f_dxn.assign(3 * (size_t)nf, 0.0);
  m.f_s.assign(nf, 0.0); m.f_lpn.assign(nf, 0.0); m.f_wfp.assign(nf, 0.0);
  m.f_ib.assign(nf, -1); m.f_bty.assign(nf, 0);
  m.n_bf = 0;
  for (int c = 0; c < nv; ++c) {
    double acc[3] = {0, 0, 0};
    for (int k = m.v_vptr[c]; k < m.v_vptr[c + 1]; ++k)
      for (int d = 0; d < nd; ++d) acc[d] = acc[d] + m.vx[3 * (size_t)m.v_vrt[k] + d];
    for (int d = 0; d < nd; ++d) m.vol_x[3 * (size_t)c + d] = acc[d] / (double)(m.v_vptr[c + 1] - m.v_vptr[c]);
    // boundary type and ib in (cell, local face) order, grd_geom.f90:45-69
    for (int k = 0; k < m.v_fptr[c + 1] - m.v_fptr[c]; ++k) {
      int code = m.v_bnd[m.v_fptr[c] + k];
      if (code <= 0) continue;
      int lst[4], nl;
      cell_face_vertices(m, c, k, lst, &nl);
      for (int s = m.v_fptr[c]; s < m.v_fptr[c + 1]; ++s) {
        int g = m.fce_lst[s];
        if (same_face(&m.f_vrt[m.f_vptr[g]], m.f_vptr[g + 1] - m.f_vptr[g], lst, nl)) {
          m.f_bty[g] = code;
          m.f_ib[g] = m.n_bf++;
        }
      }
    }
  }
  for (int g = 0; g < nf; ++g)
    if (m.f_in[g] < 0 && m.f_ib[g] < 0) return "boundary face without a boundary code";
  for (int g = 0; g < nf; ++g) {
    const int ip = m.f_ip[g];
    double *xf = &m.f_x[3 * (size_t)g], *n = &m.f_n[3 * (size_t)g];
    const int *fv = &m.f_vrt[m.f_vptr[g]];
    const int nfv = m.f_vptr[g + 1] - m.f_vptr[g];
    if (nd == 2) {
      const double *x1 = &m.vx[3 * (size_t)fv[0]], *x2 = &m.vx[3 * (size_t)fv[1]];
      double e[3] = {x2[0] - x1[0], x2[1] - x1[1], 0.0}, z[3] = {0.0, 0.0, 1.0}, cr[3];
      double s0 = mag(e, 3);
      cross(e, z, cr);
      double mg = mag(cr, 2);
      n[0] = cr[0] / mg;
      n[1] = cr[1] / mg;
      m.f_s[g] = s0;
      xf[0] = 0.5 * (x1[0] + x2[0]);
      xf[1] = 0.5 * (x1[1] + x2[1]);
    } else {
      double vec[5][3] = {{0}}, s0 = 0.0;
      const double *x1 = &m.vx[3 * (size_t)fv[0]];
      for (int k = 2; k <= nfv; ++k) {
        const double *x2 = &m.vx[3 * (size_t)fv[k - 1]];
        for (int d = 0; d < 3; ++d) vec[k][d] = x2[d] - x1[d];
        if (k > 2) {
          double cr[3];
          cross(vec[k - 1], vec[k], cr);
          for (int d = 0; d < 3; ++d) vec[1][d] = 0.5 * cr[d];
          double s1 = mag(vec[1], 3);
          for (int d = 0; d < 3; ++d) vec[0][d] = vec[0][d] + vec[1][d] / s1;
          s0 = s0 + s1;
        }
      }
      for (int d = 0; d < 3; ++d) n[d] = vec[0][d] / (double)(nfv - 2);
      m.f_s[g] = s0;
      double acc[3] = {0, 0, 0};
      for (int k = 0; k < nfv; ++k)
        for (int d = 0; d < 3; ++d) acc[d] = acc[d] + m.vx[3 * (size_t)fv[k] + d];
      for (int d = 0; d < 3; ++d) xf[d] = acc[d] / (double)nfv;
    }
    double vp = xf[0] * m.f_s[g] * n[0];  // volume by x-flux, :152-153
    m.vol_v[ip] = m.vol_v[ip] + vp;
    const double *xP = &m.vol_x[3 * (size_t)ip];
    double t[3], xpa[3];
    for (int d = 0; d < nd; ++d) t[d] = xf[d] - xP[d];
    double dd = dot(t, n, nd);
    for (int d = 0; d < nd; ++d) xpa[d] = xf[d] - n[d] * dd;
    for (int d = 0; d < nd; ++d) m.f_dxp[3 * (size_t)g + d] = xpa[d] - xP[d];
    const int in = m.f_in[g];
    double dfp[3];
    for (int d = 0; d < nd; ++d) dfp[d] = xf[d] - xpa[d];
    if (in >= 0) {
      double vpn = xf[0] * m.f_s[g] * n[0];
      m.vol_v[in] = m.vol_v[in] - vpn;
      const double *xN = &m.vol_x[3 * (size_t)in];
      double xna[3], dpn[3];
      for (int d = 0; d < nd; ++d) t[d] = xf[d] - xN[d];
      double dn = dot(t, n, nd);
      for (int d = 0; d < nd; ++d) xna[d] = xf[d] - n[d] * dn;
      for (int d = 0; d < nd; ++d) m.f_dxn[3 * (size_t)g + d] = xna[d] - xN[d];
      for (int d = 0; d < nd; ++d) dpn[d] = xna[d] - xpa[d];
      m.f_lpn[g] = mag(dpn, nd);
      m.f_wfp[g] = mag(dfp, nd) / m.f_lpn[g];
    } else {
      m.f_lpn[g] = mag(dfp, nd);
    }
  }
  for (int c = 0; c < nv; ++c) m.vol_vr[c] = 1.0 / m.vol_v[c];
  return "";
}

// ---- grd_read ----
static int boundary_code(const char *name32) {
  char nm[64];
  int a = 0, n = 0;
  while (name32[a] == ' ') ++a;
  while (name32[a + n] && !std::strchr(" \r\n", name32[a + n]) && n < 63) { nm[n] = name32[a + n]; ++n; }
  nm[n] = 0;
  const char *us = std::strchr(nm, '_');
  int bt = (us && us[1] >= '0' && us[1] <= '9') ? us[1] - '0' : 0;
  if (!std::strncmp(nm, "inlet", 5)) bt += 10;
  if (!std::strncmp(nm, "outlet", 6)) bt += 20;
  if (!std::strncmp(nm, "wall", 4)) bt += 30;
  if (!std::strncmp(nm, "symmetry", 8)) bt += 40;
  if (!std::strncmp(nm, "pressure", 8)) bt += 50;
  return bt;
}

static void first_token(const char *line, char *tok) {
  int a = 0, n = 0;
  while (line[a] == ' ' || line[a] == '\t' || line[a] == ',') ++a;
  while (line[a + n] && !std::strchr(" \t,/\r\n", line[a + n]) && n < 31) { tok[n] = line[a + n]; ++n; }
  tok[n] = 0;
}

std::string read_neu(Mesh &m, const char *path, double l_sc) {
  FILE *fp = std::fopen(path, "r");
  if (!fp) return std::string("cannot open ") + path;
  char line[1024], tok[32];
  int n_vrt = 0, n_vol = 0, n_grp = 0, n_bs = 0, n_dx = 0, n_vc = 0, ibs = 0;
  std::vector<double> vx;
  std::vector<int> ty, nvc, vtmp, vptr, vrt, sptr, bnd;
  bool have_elem = false;
  std::string err;
  while (std::fgets(line, sizeof line, fp)) {
    first_token(line, tok);
    if (!tok[0]) continue;
    if (!std::strcmp(tok, "NUMNP")) {
      if (!std::fgets(line, sizeof line, fp)) break;
      if (std::sscanf(line, "%d %d %d %d %d %d", &n_vrt, &n_vol, &n_grp, &n_bs, &n_dx, &n_vc) != 6) {
        err = "bad NUMNP record";
        break;
      }
      vx.assign(3 * (size_t)n_vrt, 0.0);
      ty.assign(n_vol, 0);
      nvc.assign(n_vol, 0);
      vtmp.assign(8 * (size_t)n_vol, 0);
    } else if (!std::strcmp(tok, "NODAL")) {
      for (int k = 0; k < n_vrt; ++k) {
        int idx = 0;
        double x[3] = {0, 0, 0};
        if (!std::fgets(line, sizeof line, fp) || std::sscanf(line, "%d %lf %lf %lf", &idx, &x[0], &x[1], &x[2]) < 1 + n_dx ||
            idx < 1 || idx > n_vrt) {
          err = "bad NODAL record";
          break;
        }
        for (int d = 0; d < n_dx; ++d) vx[3 * (size_t)(idx - 1) + d] = x[d] * l_sc;
      }
    } else if (!std::strcmp(tok, "ELEMENTS")) {
      // idx type nvrt v1..v7 / continuation records of up to 7 ids
      std::vector<int> ints;
      const char *cur = "";
      auto next_int = [&](int &v) -> bool {
        for (;;) {
          while (*cur == ' ' || *cur == '\t') ++cur;
          if (*cur == 0 || *cur == '\n' || *cur == '\r') {
            if (!std::fgets(line, sizeof line, fp)) return false;
            cur = line;
            continue;
          }
          char *end;
          long x = std::strtol(cur, &end, 10);
          if (end == cur) return false;
          v = (int)x;
          cur = end;
          return true;
        }
      };
      for (int k = 0; k < n_vol && err.empty(); ++k) {
        int idx, t, nn;
        if (!next_int(idx) || !next_int(t) || !next_int(nn) || idx < 1 || idx > n_vol || nn < 1 || nn > 8) {
          err = "bad ELEMENTS record";
          break;
        }
        ty[idx - 1] = t;
        nvc[idx - 1] = nn;
        for (int a = 0; a < nn; ++a)
          if (!next_int(vtmp[8 * (size_t)(idx - 1) + a])) { err = "bad ELEMENTS record"; break; }
      }
      vptr.assign(n_vol + 1, 0);
      for (int c = 0; c < n_vol; ++c) vptr[c + 1] = vptr[c] + nvc[c];
      vrt.resize(vptr[n_vol]);
      for (int c = 0; c < n_vol; ++c)
        for (int a = 0; a < nvc[c]; ++a) vrt[vptr[c] + a] = vtmp[8 * (size_t)c + a];
      sptr.assign(n_vol + 1, 0);
      for (int c = 0; c < n_vol; ++c) sptr[c + 1] = sptr[c] + elem_nfce(ty[c]);
      bnd.assign(sptr[n_vol], 0);
      have_elem = true;
    } else if (!std::strcmp(tok, "BOUNDARY")) {
      ++ibs;
      if (!have_elem || !std::fgets(line, sizeof line, fp)) { err = "BOUNDARY before ELEMENTS"; break; }
      char name[33];
      std::memcpy(name, line, 32);
      name[32] = 0;
      int bty_file = 0, n = 0;
      if (std::strlen(line) < 33 || std::sscanf(line + 32, "%d %d", &bty_file, &n) != 2) { err = "bad BOUNDARY header"; break; }
      int code = boundary_code(name);
      for (int k = 0; k < n; ++k) {
        int iv, et, ivf;
        if (!std::fgets(line, sizeof line, fp) || std::sscanf(line, "%d %d %d", &iv, &et, &ivf) != 3 || iv < 1 ||
            iv > n_vol || ivf < 1 || ivf > sptr[iv] - sptr[iv - 1]) {
          err = "bad BOUNDARY record";
          break;
        }
        bnd[sptr[iv - 1] + ivf - 1] = code;
      }
    }
    if (!err.empty()) break;
    if (n_bs > 0 && ibs == n_bs) break;
  }
  std::fclose(fp);
  if (!err.empty()) return err;
  if (!have_elem || n_vrt == 0) return "incomplete neutral file";
  (void)n_grp;
  (void)n_vc;
  return init_cells(m, n_dx, n_vrt, vx.data(), n_vol, ty.data(), vptr.data(), vrt.data(), bnd.data());
}

static uint64_t splitmix64(uint64_t &s) {
  uint64_t z = (s += 0x9E3779B97F4A7C15ULL);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}

std::string make_box(Mesh &m, int nx, int ny, int nz, double lx, double ly, double lz, double jitter,
                     uint64_t seed, int lid, int wall) {
  if (nx < 1 || ny < 1 || nz < 0) return "bad box size";
  const bool three = nz > 0;
  const int nzv = three ? nz + 1 : 1, nzc = three ? nz : 1;
  const long n_vrt = (long)(nx + 1) * (ny + 1) * nzv, n_vol = (long)nx * ny * nzc;
  if (n_vrt > 0x7fffffffL / 3 || n_vol * 8 > 0x7fffffffL) return "box too large for 32-bit ids";
  std::vector<double> vx(3 * (size_t)n_vrt, 0.0);
  const double h[3] = {lx / nx, ly / ny, three ? lz / nz : 0.0};
  uint64_t st = seed;
  for (int k = 0; k < nzv; ++k)
    for (int j = 0; j <= ny; ++j)
      for (int i = 0; i <= nx; ++i) {
        double *x = &vx[3 * ((size_t)i + (size_t)(nx + 1) * (j + (size_t)(ny + 1) * k))];
        x[0] = lx * (double)i / (double)nx;
        x[1] = ly * (double)j / (double)ny;
        x[2] = three ? lz * (double)k / (double)nz : 0.0;
        bool interior = i > 0 && i < nx && j > 0 && j < ny && (!three || (k > 0 && k < nz));
        if (jitter > 0.0 && interior)
          for (int d = 0; d < (three ? 3 : 2); ++d) {
            double u = (double)(splitmix64(st) >> 11) * 0x1.0p-53;
            x[d] = x[d] + (2.0 * u - 1.0) * jitter * h[d];
          }
      }
  const int nv = three ? 8 : 4, nf = three ? 6 : 4;
  std::vector<int> ty(n_vol), vptr(n_vol + 1), vrt(nv * (size_t)n_vol), bnd(nf * (size_t)n_vol, 0);
  vptr[0] = 0;
  auto V = [&](int a, int b, int c) { return 1 + a + (nx + 1) * (b + (ny + 1) * c); };
  for (int k = 0; k < nzc; ++k)
    for (int j = 0; j < ny; ++j)
      for (int i = 0; i < nx; ++i) {
        size_t c = (size_t)i + (size_t)nx * (j + (size_t)ny * k);
        ty[c] = three ? 4 : 2;
        vptr[c + 1] = (int)((c + 1) * nv);
        int *w = &vrt[nv * c], *b = &bnd[nf * c];
        if (three) {
          w[0] = V(i, j, k); w[1] = V(i + 1, j, k); w[2] = V(i, j + 1, k); w[3] = V(i + 1, j + 1, k);
          w[4] = V(i, j, k + 1); w[5] = V(i + 1, j, k + 1); w[6] = V(i, j + 1, k + 1); w[7] = V(i + 1, j + 1, k + 1);
          if (j == 0) b[0] = wall;
          if (i == nx - 1) b[1] = wall;
          if (j == ny - 1) b[2] = wall;
          if (i == 0) b[3] = wall;
          if (k == 0) b[4] = wall;
          if (k == nz - 1) b[5] = lid;
        } else {
          w[0] = V(i, j, 0); w[1] = V(i + 1, j, 0); w[2] = V(i + 1, j + 1, 0); w[3] = V(i, j + 1, 0);
          if (j == 0) b[0] = wall;
          if (i == nx - 1) b[1] = wall;
          if (j == ny - 1) b[2] = lid;
          if (i == 0) b[3] = wall;
        }
      }
  return init_cells(m, three ? 3 : 2, (int)n_vrt, vx.data(), (int)n_vol, ty.data(), vptr.data(), vrt.data(), bnd.data());
}

std::string finish(Mesh &m) {
  std::string e = connect(m);
  if (!e.empty()) return e;
  return geometry(m);
}

}  // namespace fvk
