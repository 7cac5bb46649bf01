// Golden-output driver of flowlab's gmsh reader and field writers
// (flowlab/src/mesh.cpp: read_gmsh_mesh, write_gmsh_field_,
// write_gmsh_vector_field_).
//
// TEST INFRASTRUCTURE.  This program is the project's own code; it is
// linked with flowlab/src/mesh.cpp compiled where it lies (see
// oracle/build_ref.sh).
//
//   gmsh_ref read MSH OUT        read_gmsh_mesh(mode 's', section 'A') and
//                                dump the arrays it fills (int32 stream:
//                                n_vrt, vrt_idx, n_elm, elm_idx, tag_ptr,
//                                tag_lst, vrt_ptr, vrt_lst; then float64
//                                x_vrt)
//   gmsh_ref field IN OUT        write_gmsh_field_ of IN (int32 n, int32
//                                elm_idx[n], float64 phi[n]) to file OUT
//   gmsh_ref vector IN OUT       write_gmsh_vector_field_ of IN (int32 n,
//                                int32 n_dim, int32 elm_idx[n], float64
//                                phi[n*n_dim]) to file OUT
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "mesh.h"

static void put(FILE *f, const int *a, int n) { fwrite(a, sizeof(int), n, f); }

int main(int argc, char **argv) {
  if (argc != 4) { fprintf(stderr, "usage: gmsh_ref read|field|vector IN OUT\n"); return 2; }
  if (!strcmp(argv[1], "read")) {
    mesh_t m;
    memset(&m, 0, sizeof m);
    m.part = 0;
    m.n_part = 1;
    read_gmsh_mesh(&m, argv[2], 'A', 's');
    FILE *f = fopen(argv[3], "wb");
    if (!f) return 1;
    put(f, &m.n_vrt, 1);
    put(f, m.vrt_idx, m.n_vrt);
    put(f, &m.n_elm, 1);
    put(f, m.elm_idx, m.n_elm);
    put(f, m.elm_tag_ptr, m.n_elm + 1);
    put(f, m.elm_tag_lst, m.elm_tag_ptr[m.n_elm]);
    put(f, m.elm_vrt_ptr, m.n_elm + 1);
    put(f, m.elm_vrt_lst, m.elm_vrt_ptr[m.n_elm]);
    fwrite(m.x_vrt, sizeof(double), (size_t)3 * m.n_vrt, f);
    fclose(f);
    return 0;
  }
  FILE *f = fopen(argv[2], "rb");
  if (!f) return 1;
  int n = 0, nd = 1;
  if (fread(&n, sizeof n, 1, f) != 1) return 1;
  const bool vec = !strcmp(argv[1], "vector");
  if (vec && fread(&nd, sizeof nd, 1, f) != 1) return 1;
  std::vector<int> idx(n);
  std::vector<double> phi((size_t)n * nd);
  if (n && (fread(idx.data(), sizeof(int), n, f) != (size_t)n ||
            fread(phi.data(), sizeof(double), phi.size(), f) != phi.size()))
    return 1;
  fclose(f);
  int len = (int)strlen(argv[3]);
  if (vec) return write_gmsh_vector_field_(&len, argv[3], &n, &nd, idx.data(), phi.data());
  return write_gmsh_field_(&len, argv[3], &n, idx.data(), phi.data());
}
