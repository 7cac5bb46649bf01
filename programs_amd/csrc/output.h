// Output formats of the path's fields (host side): Tecplot FEBLOCK files as
// spray-simulation's eqn_outp / prp_outp write them, gmsh ElementData files
// as flowlab's write_gmsh_field_ / write_gmsh_vector_field_ write them.
#pragma once
#include <string>

namespace fvk {

// flang's list-directed editing of a REAL(8) value
std::string fmt_list_real(double x);

// One list-directed output record: a blank opens it, one blank separates
// items except two character items in a row.
class ListRecord {
 public:
  void chr(const std::string &s);
  void real(double x);
  void integer(long v);
  std::string take();  // the record with its newline; starts the next one

 private:
  void num(const std::string &s);
  std::string line_ = " ";
  bool first_ = true, last_chr_ = false;
};

struct TecplotIn {
  bool kind_eqn = true;  // eqn_outp (n_sc x n_vc values per cell) or prp_outp (one)
  int igrd = 1, iph = 1, ieq = 1, ict = 0, n_tl = 1;
  int n_dx = 3, n_vrt = 0, n_vol = 0, n_sc = 1, n_vc = 1;
  const double *vx = nullptr;       // 3 * n_vrt
  const int *v_vptr = nullptr;      // n_vol + 1
  const int *v_vrt = nullptr;       // 0-based vertex ids
  const double *val = nullptr;      // n_vol x (n_sc * n_vc), cell-major
};

// writes <dir><title>.dat; returns "" or an error message; path receives
// the file name
std::string tecplot_write(const TecplotIn &in, const char *dir, std::string &path);
std::string gmsh_write(const char *fname, int n, int n_cmp, int n_dim, const int *elm_idx, const double *phi);

}  // namespace fvk
