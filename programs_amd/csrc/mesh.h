// Host-side mesh of the fvk library (0-based ids internally; the C ABI
// converts to the reference's 1-based numbering).
#pragma once
#include <cstdint>
#include <string>
#include <vector>

namespace fvk {

struct Mesh {
  int n_dx = 0, n_vrt = 0, n_vol = 0, n_fce = 0, n_bf = 0, n_slot = 0, max_fce = 0;
  std::vector<double> vx;                 // 3*n_vrt
  std::vector<int> v_ty, v_vptr, v_vrt;   // element type, vertex lists (0-based vertex ids)
  std::vector<int> v_fptr, v_bnd;         // slot ranges, boundary code per (cell, local face)
  std::vector<int> fce_lst, ng_lst;       // per slot: face (ascending), neighbour cell or -1
  std::vector<int> f_ip, f_in, f_ib, f_bty, f_vptr, f_vrt;  // f_in = -1 and f_ib >= 0 on boundary
  std::vector<double> vol_x, vol_v, vol_vr;                 // vol_x stride 3
  std::vector<double> f_x, f_n, f_dxp, f_dxn, f_s, f_lpn, f_wfp;  // vectors stride 3
  std::vector<int> elm_idx;  // element number of each cell in its source file (empty: 1..n_vol)
};

int elem_nfce(int ty);
std::string init_cells(Mesh &m, int n_dx, int n_vrt, const double *vx3, int n_vol, const int *v_ty,
                       const int *v_vptr, const int *v_vrt1, const int *bnd);
std::string read_neu(Mesh &m, const char *path, double l_sc);
// The records of a gmsh 2.2 ASCII file, as flowlab's read_gmsh_mesh takes
// them (flowlab/src/mesh.cpp:39-187): node ids and coordinates, and per
// element its number, type, first tag (the physical tag) and node ids.
struct GmshRaw {
  std::vector<int> node_id;
  std::vector<double> node_x;  // 3 per node
  std::vector<int> e_idx, e_ty, e_tag0, e_vptr, e_vrt;
};
std::string parse_gmsh(GmshRaw &g, const char *path);
// gmsh 2.2 ASCII mesh; n_dx <= 0: from the highest element dimension;
// code_map: n_map pairs (physical tag, boundary code)
std::string read_gmsh(Mesh &m, const char *path, int n_dx, double l_sc, const int *code_map, int n_map);
std::string make_box(Mesh &m, int nx, int ny, int nz, double lx, double ly, double lz, double jitter,
                     uint64_t seed, int lid, int wall);
std::string connect(Mesh &m);
std::string geometry(Mesh &m);
std::string finish(Mesh &m);  // connect + geometry
// the same on a GPU (meshdev.hip): bit-identical arrays
std::string finish_device(Mesh &m, int device);

}  // namespace fvk
