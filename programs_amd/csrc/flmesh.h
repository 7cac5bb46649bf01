// Host mesh in the layout flowlab's Fortran solver receives from its C++
// mesh pipeline (flowlab/src/main.cpp:146-480, navierstokes.f90:1-300), one
// serial block: boundary elements numbered 1..n_bnd, domain elements
// n_bnd+1 .. n_bnd+n_dmn, faces with their (owner, neighbour) element pair
// and vertex list, and the CSR pattern of the cell matrix (diagonal
// included, columns ascending).  All indices 1-based, as Fortran reads them.
#pragma once
#include <string>
#include <vector>

#include "mesh.h"

namespace fvk {

struct FlMesh {
  int n_dim = 0, n_vrt = 0, n_bnd = 0, n_dmn = 0, n_fce = 0;
  std::vector<double> x;             // 3 per vertex
  std::vector<int> elm_vptr, elm_vrt;  // domain element vertex lists (n_dmn+1; vertex ids 1..n_vrt)
  std::vector<int> elm_ty;           // gmsh element type of each domain element
  std::vector<int> elm_idx;          // gmsh element number of each domain element
  std::vector<int> bnd_tag;          // physical tag of each boundary element
  std::vector<int> fce_elm;          // 2 per face: owner element, neighbour element or -(boundary element)
  std::vector<int> fce_vptr, fce_vrt;  // face vertex lists (n_fce+1; 1-based offsets)
  std::vector<int> ng_ptr, ng_col;   // cell matrix pattern (n_dmn+1; 1-based offsets; local columns)
};

// From gmsh records: elements of dimension n_dim are the domain, elements of
// dimension n_dim - 1 the boundary, each in file order; coordinates scaled
// by l_sc.
std::string fl_mesh_from_gmsh(FlMesh &m, const GmshRaw &g, int n_dim, double l_sc);

}  // namespace fvk
