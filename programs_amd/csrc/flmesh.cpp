// flowlab-layout mesh from gmsh records (see flmesh.h).
//
// The element face table is flowlab's mesh format file
// (flowlab/etc/mshf/gmsh.fmt: Face.Vertex_Lists per element type), data kept
// alike: each face's local vertices in the order that makes the face
// normal of the geometry (flowlab/src/geometry_m.f90:174-204) point out of
// the element.
//
// Faces: every domain element's faces in element order; a face met a second
// time gets that element as its neighbour, a boundary element with the same
// vertex set makes it a boundary face (neighbour -k).  As flowlab lists them
// (flowlab/src/xgrph.cpp:206-221), boundary faces come first and interior
// faces follow in reverse order of discovery; an interior face keeps the
// vertex order of the element that met it first (its owner).  A face met
// three times, a boundary element without a domain face, or a domain face
// with neither neighbour nor boundary element is an error.
#include "flmesh.h"

#include <algorithm>
#include <cstdint>
#include <unordered_map>

namespace fvk {

namespace {

struct FaceDef {
  int n_fce;
  int nv[6];
  int v[6][4];
};

// gmsh element type -> faces (gmsh.fmt); types 1 (line) and 0 unused here
const FaceDef *face_def(int ty) {
  static const FaceDef tri = {3, {2, 2, 2}, {{0, 1}, {1, 2}, {2, 0}}};
  static const FaceDef quad = {4, {2, 2, 2, 2}, {{0, 1}, {1, 2}, {2, 3}, {3, 0}}};
  static const FaceDef tet = {4, {3, 3, 3, 3}, {{0, 3, 2}, {0, 1, 3}, {0, 2, 1}, {1, 2, 3}}};
  static const FaceDef hex = {6, {4, 4, 4, 4, 4, 4},
                              {{4, 5, 6, 7}, {4, 0, 1, 5}, {4, 7, 3, 0}, {2, 6, 5, 1}, {2, 1, 0, 3}, {2, 3, 7, 6}}};
  static const FaceDef prism = {5, {3, 4, 4, 4, 3}, {{4, 5, 3}, {4, 1, 2, 5}, {4, 3, 0, 1}, {2, 0, 3, 5}, {2, 1, 0}}};
  static const FaceDef pyr = {5, {3, 4, 3, 3, 3}, {{1, 2, 4}, {1, 0, 3, 2}, {1, 4, 0}, {3, 4, 2}, {3, 0, 4}}};
  switch (ty) {
    case 2: return &tri;
    case 3: return &quad;
    case 4: return &tet;
    case 5: return &hex;
    case 6: return &prism;
    case 7: return &pyr;
    default: return nullptr;
  }
}

struct Key {
  int v[4];
  int n;
  bool operator==(const Key &o) const {
    if (n != o.n) return false;
    for (int i = 0; i < n; ++i)
      if (v[i] != o.v[i]) return false;
    return true;
  }
};
struct KeyHash {
  size_t operator()(const Key &k) const {
    uint64_t h = 0x9E3779B97F4A7C15ULL ^ (uint64_t)k.n;
    for (int i = 0; i < k.n; ++i) {
      h ^= (uint64_t)(uint32_t)k.v[i] + 0x9E3779B97F4A7C15ULL + (h << 6) + (h >> 2);
      h *= 0xBF58476D1CE4E5B9ULL;
    }
    return (size_t)h;
  }
};

Key make_key(const int *v, int n) {
  Key k;
  k.n = n;
  for (int i = 0; i < n; ++i) k.v[i] = v[i];
  std::sort(k.v, k.v + n);
  return k;
}

}  // namespace

std::string fl_mesh_from_gmsh(FlMesh &m, const GmshRaw &g, int n_dim, double l_sc) {
  static const int kDim[8] = {0, 1, 2, 2, 3, 3, 3, 3};
  if (n_dim != 2 && n_dim != 3) return "n_dim must be 2 or 3";
  m = FlMesh();
  m.n_dim = n_dim;
  m.n_vrt = (int)g.node_id.size();
  std::unordered_map<int, int> pos;
  pos.reserve(g.node_id.size() * 2);
  for (size_t i = 0; i < g.node_id.size(); ++i)
    if (!pos.emplace(g.node_id[i], (int)i + 1).second) return "duplicate node id " + std::to_string(g.node_id[i]);
  m.x.resize(3 * (size_t)m.n_vrt, 0.0);
  for (int i = 0; i < m.n_vrt; ++i)
    for (int d = 0; d < 3; ++d) m.x[3 * (size_t)i + d] = g.node_x[3 * (size_t)i + d] * l_sc;
  // boundary and domain elements in file order
  std::vector<int> bnd_e, dmn_e;
  for (size_t e = 0; e < g.e_ty.size(); ++e) {
    const int d = kDim[g.e_ty[e]];
    if (d == n_dim) dmn_e.push_back((int)e);
    else if (d == n_dim - 1) bnd_e.push_back((int)e);
  }
  m.n_bnd = (int)bnd_e.size();
  m.n_dmn = (int)dmn_e.size();
  if (!m.n_dmn) return "no elements of dimension " + std::to_string(n_dim);
  auto vid = [&](int e, int k, int &out) -> bool {
    auto it = pos.find(g.e_vrt[g.e_vptr[e] + k]);
    if (it == pos.end()) return false;
    out = it->second;
    return true;
  };
  m.elm_vptr.assign(1, 1);
  for (int c = 0; c < m.n_dmn; ++c) {
    const int e = dmn_e[c];
    if (!face_def(g.e_ty[e])) return "unsupported cell type " + std::to_string(g.e_ty[e]);
    const int nv = g.e_vptr[e + 1] - g.e_vptr[e];
    for (int k = 0; k < nv; ++k) {
      int v;
      if (!vid(e, k, v)) return "element " + std::to_string(g.e_idx[e]) + " names an unknown node";
      m.elm_vrt.push_back(v);
    }
    m.elm_vptr.push_back((int)m.elm_vrt.size() + 1);
    m.elm_ty.push_back(g.e_ty[e]);
    m.elm_idx.push_back(g.e_idx[e]);
  }
  for (int b = 0; b < m.n_bnd; ++b) m.bnd_tag.push_back(g.e_tag0[bnd_e[b]]);
  // faces in order of discovery
  std::vector<int> f_own, f_ngb, f_vptr{0}, f_vrt;
  std::unordered_map<Key, int, KeyHash> fmap;
  fmap.reserve((size_t)m.n_dmn * 8);
  for (int c = 0; c < m.n_dmn; ++c) {
    const FaceDef *fd = face_def(m.elm_ty[c]);
    const int *ev = &m.elm_vrt[m.elm_vptr[c] - 1];
    const int glob = m.n_bnd + c + 1;
    for (int k = 0; k < fd->n_fce; ++k) {
      int fv[4];
      for (int a = 0; a < fd->nv[k]; ++a) fv[a] = ev[fd->v[k][a]];
      const Key key = make_key(fv, fd->nv[k]);
      auto it = fmap.find(key);
      if (it == fmap.end()) {
        fmap.emplace(key, (int)f_own.size());
        f_own.push_back(glob);
        f_ngb.push_back(0);
        for (int a = 0; a < fd->nv[k]; ++a) f_vrt.push_back(fv[a]);
        f_vptr.push_back((int)f_vrt.size());
      } else {
        const int f = it->second;
        if (f_ngb[f] != 0) return "face shared by more than two elements (element " + std::to_string(m.elm_idx[c]) + ")";
        f_ngb[f] = glob;
      }
    }
  }
  for (int b = 0; b < m.n_bnd; ++b) {
    const int e = bnd_e[b];
    const int nv = g.e_vptr[e + 1] - g.e_vptr[e];
    if (nv < 2 || nv > 4) return "unsupported boundary element type " + std::to_string(g.e_ty[e]);
    int fv[4];
    for (int k = 0; k < nv; ++k)
      if (!vid(e, k, fv[k])) return "element " + std::to_string(g.e_idx[e]) + " names an unknown node";
    auto it = fmap.find(make_key(fv, nv));
    if (it == fmap.end()) return "boundary element " + std::to_string(g.e_idx[e]) + " matches no cell face";
    const int f = it->second;
    if (f_ngb[f] != 0) return "boundary element " + std::to_string(g.e_idx[e]) + " lies on an interior face";
    f_ngb[f] = -(b + 1);
  }
  const int nf = (int)f_own.size();
  for (int f = 0; f < nf; ++f)
    if (f_ngb[f] == 0) return "a face of element " + std::to_string(m.elm_idx[f_own[f] - m.n_bnd - 1]) +
                              " has neither a neighbour nor a boundary element";
  // boundary faces first, then interior faces in reverse order of discovery
  std::vector<int> order;
  order.reserve(nf);
  for (int f = 0; f < nf; ++f)
    if (f_ngb[f] < 0) order.push_back(f);
  for (int f = nf - 1; f >= 0; --f)
    if (f_ngb[f] > 0) order.push_back(f);
  m.n_fce = nf;
  m.fce_vptr.assign(1, 1);
  for (int f : order) {
    m.fce_elm.push_back(f_own[f]);
    m.fce_elm.push_back(f_ngb[f]);
    for (int k = f_vptr[f]; k < f_vptr[f + 1]; ++k) m.fce_vrt.push_back(f_vrt[k]);
    m.fce_vptr.push_back((int)m.fce_vrt.size() + 1);
  }
  // cell matrix pattern: the diagonal and the interior neighbours, ascending
  std::vector<std::vector<int>> nb(m.n_dmn);
  for (int c = 0; c < m.n_dmn; ++c) nb[c].push_back(c + 1);
  for (int f = 0; f < nf; ++f)
    if (f_ngb[f] > 0) {
      const int a = f_own[f] - m.n_bnd, b = f_ngb[f] - m.n_bnd;
      nb[a - 1].push_back(b);
      nb[b - 1].push_back(a);
    }
  m.ng_ptr.assign(1, 1);
  for (int c = 0; c < m.n_dmn; ++c) {
    std::sort(nb[c].begin(), nb[c].end());
    for (int j : nb[c]) m.ng_col.push_back(j);
    m.ng_ptr.push_back((int)m.ng_col.size() + 1);
  }
  return "";
}

}  // namespace fvk
