// Output formats of the path's fields (host side):
//
//   * Tecplot FEBLOCK files of cell-centred fields, as spray-simulation's
//     eqn_outp / prp_outp write them (spray-simulation/cfd_outp.f90:8-189):
//     a list-directed Fortran record per header line, per vertex coordinate,
//     per cell value and per cell vertex list (converted to Tecplot's brick
//     or quadrilateral by the table of spray-simulation/tecplot.vrt,
//     grd_constr.f90:59-87);
//   * gmsh ElementData files, as flowlab's write_gmsh_field_ and
//     write_gmsh_vector_field_ write them (flowlab/src/mesh.cpp:298-412).
//
// The Tecplot files reproduce the bytes of the flang build of the
// reference: list-directed output (a blank opens every record, one blank
// between items except between two character items) with flang's editing
// of REAL(8) values: the fewest significant digits p for which a p-digit
// decimal lies strictly between the midpoints to the neighbouring doubles,
// and of those the middle one (which is not always the nearest: the 17-digit
// form of 2.2719716328605812519...E-222 is printed ...812); in F form
// (leading zero dropped: ".1", "12.5", "100.") when the value rounded to one
// significant digit is 0.d x 10^e with 0 <= e <= 15, else in E form with one
// digit before the point ("3.125E-02", "1.E+16", "1.E-300"); zero is "0."
// (or "-0.").  Established on 14,575 values printed by a flang-built probe
// (tests/golden/flang_list_real.npz).
#include <cerrno>
#include <charconv>
#include <cstdint>
#include <cstdlib>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "output.h"

namespace fvk {

namespace {

// Minimal unsigned big integer (base 2^32 limbs, little endian) for the
// exact digit search of values at or above 2^52.
struct Big {
  std::vector<uint32_t> w;
  explicit Big(uint64_t v = 0) {
    while (v) { w.push_back((uint32_t)v); v >>= 32; }
  }
  void trim() { while (!w.empty() && !w.back()) w.pop_back(); }
  void mul(uint32_t m) {
    uint64_t c = 0;
    for (auto &x : w) { uint64_t t = (uint64_t)x * m + c; x = (uint32_t)t; c = t >> 32; }
    if (c) w.push_back((uint32_t)c);
  }
  void pow10(int k) { for (; k > 0; --k) mul(10); }
  void shl(int k) {
    if (w.empty() || k <= 0) return;
    const int limbs = k / 32, bits = k % 32;
    std::vector<uint32_t> r(limbs, 0);
    uint32_t c = 0;
    for (auto x : w) {
      r.push_back(bits ? (x << bits) | c : x);
      c = bits ? x >> (32 - bits) : 0;
    }
    if (c) r.push_back(c);
    w.swap(r);
  }
  static int cmp(const Big &a, const Big &b) {
    if (a.w.size() != b.w.size()) return a.w.size() < b.w.size() ? -1 : 1;
    for (size_t i = a.w.size(); i-- > 0;)
      if (a.w[i] != b.w[i]) return a.w[i] < b.w[i] ? -1 : 1;
    return 0;
  }
  void sub(const Big &b) {  // *this >= b
    int64_t br = 0;
    for (size_t i = 0; i < w.size(); ++i) {
      int64_t t = (int64_t)w[i] - (i < b.w.size() ? b.w[i] : 0) - br;
      br = t < 0;
      w[i] = (uint32_t)(t + (br ? ((int64_t)1 << 32) : 0));
    }
    trim();
  }
};

// floor(num / den) for a quotient below 2^63; rem_zero: exact division
uint64_t divq(Big num, const Big &den, bool &rem_zero) {
  uint64_t q = 0;
  for (int b = 62; b >= 0; --b) {
    Big t = den;
    t.shl(b);
    if (Big::cmp(num, t) >= 0) { num.sub(t); q |= (uint64_t)1 << b; }
  }
  num.trim();
  rem_zero = num.w.empty();
  return q;
}

// value (4 f + d) 2^(e-2) / 10^t as num/den
void ratio(uint64_t f4d, int e, int t, Big &num, Big &den) {
  num = Big(f4d);
  den = Big(1);
  if (e - 2 >= 0) num.shl(e - 2); else den.shl(2 - e);
  if (t >= 0) den.pow10(t); else num.pow10(-t);
}

// The digits flang's list-directed output gives a finite a > 0: among the
// fewest significant digits p for which some p-digit decimal lies strictly
// between the midpoints to the neighbouring doubles, the middle one
// (floor of the mean of the least and the greatest).  Returns the digit
// string D and k with a = 0.D x 10^k.
void flang_digits(double a, std::string &D, int &k) {
  char buf[64];
  auto r = std::to_chars(buf, buf + sizeof buf, a, std::chars_format::scientific);
  *r.ptr = '\0';
  const char *e = std::strchr(buf, 'e');
  const int ex = std::atoi(e + 1);
  uint64_t C = 0;
  int n = 0;
  for (const char *p = buf; p < e; ++p)
    if (*p != '.') { C = C * 10 + (uint64_t)(*p - '0'); ++n; }
  if (a < 4503599627370496.0) {
    // Below 2^52 no midpoint between doubles is a decimal of 17 or fewer
    // significant digits, so the shortest round-trip length n is flang's,
    // and a candidate D x 10^t lies strictly inside iff it reads back to a
    // (fewer than 10 candidates at the shortest length).
    const int t = ex - (n - 1);
    auto member = [&](uint64_t d) {
      char s[48];
      std::snprintf(s, sizeof s, "%llue%d", (unsigned long long)d, t);
      return std::strtod(s, nullptr) == a;
    };
    uint64_t lo = C, hi = C;
    for (int j = 0; j < 9 && lo > 0 && member(lo - 1); ++j) --lo;
    for (int j = 0; j < 9 && member(hi + 1); ++j) ++hi;
    const uint64_t v = lo + (hi - lo) / 2;
    D = std::to_string(v);
    k = t + (int)D.size();
  } else {
    // exact search on the big-integer boundaries
    int ee;
    const double m = std::frexp(a, &ee);
    const uint64_t f = (uint64_t)std::ldexp(m, 53);
    const int e2 = ee - 53;                       // a = f 2^e2
    const bool lower_half = f == ((uint64_t)1 << 52) && e2 > -1074;
    const uint64_t L4 = 4 * f - (lower_half ? 1 : 2), U4 = 4 * f + 2;
    // K: 10^(K-1) <= U < 10^K
    int K = ex + 1;
    {
      Big num, den;
      ratio(U4, e2, K, num, den);
      if (Big::cmp(num, den) >= 0) ++K;
    }
    for (int p = std::max(1, n); p <= 40; ++p) {
      const int t = K - p;
      Big num, den;
      bool ez;
      ratio(L4, e2, t, num, den);
      uint64_t lo = divq(num, den, ez) + 1;  // strictly above L
      ratio(U4, e2, t, num, den);
      uint64_t hi = divq(num, den, ez);
      if (ez) --hi;                           // strictly below U
      if (lo <= hi) {
        const uint64_t v = lo + (hi - lo) / 2;
        D = std::to_string(v);
        k = t + (int)D.size();
        break;
      }
    }
  }
  while (D.size() > 1 && D.back() == '0') D.pop_back();
}

}  // namespace

std::string fmt_list_real(double x) {
  if (x == 0.0) return std::signbit(x) ? "-0." : "0.";
  if (std::isnan(x)) return "NaN";
  if (std::isinf(x)) return x < 0 ? "-Inf" : "Inf";
  const double a = std::fabs(x);
  char buf[64];
  // decimal exponent of |x| rounded to one significant digit: 0.d x 10^expo1
  auto r1 = std::to_chars(buf, buf + sizeof buf, a, std::chars_format::scientific, 0);
  *r1.ptr = '\0';
  const int expo1 = std::atoi(std::strchr(buf, 'e') + 1) + 1;
  std::string digits;
  int k = 0;  // a = 0.digits x 10^k
  flang_digits(a, digits, k);
  std::string out = x < 0 ? "-" : "";
  const int nd = (int)digits.size();
  if (expo1 >= 0 && expo1 <= 15) {
    if (k >= nd) out += digits + std::string(k - nd, '0') + ".";
    else if (k > 0) out += digits.substr(0, k) + "." + digits.substr(k);
    else out += "." + std::string(-k, '0') + digits;
  } else {
    const int ex = k - 1;
    out += digits.substr(0, 1) + "." + digits.substr(1);
    char eb[16];
    std::snprintf(eb, sizeof eb, "E%c%02d", ex < 0 ? '-' : '+', ex < 0 ? -ex : ex);
    out += eb;
  }
  return out;
}

void ListRecord::chr(const std::string &s) {
  if (!first_ && !last_chr_) line_ += ' ';
  line_ += s;
  first_ = false;
  last_chr_ = true;
}

void ListRecord::num(const std::string &s) {
  if (!first_) line_ += ' ';
  line_ += s;
  first_ = false;
  last_chr_ = false;
}

void ListRecord::real(double x) { num(fmt_list_real(x)); }
void ListRecord::integer(long v) { num(std::to_string(v)); }

std::string ListRecord::take() {
  std::string s = line_ + "\n";
  line_ = " ";
  first_ = true;
  last_chr_ = false;
  return s;
}

namespace {

// Tecplot vertex arrangement of each element (1-based positions in the
// element's vertex list), the data of spray-simulation/tecplot.vrt:
// quadrilateral and triangle in 2D; brick, wedge, tetrahedron and pyramid
// in 3D.
const int kTec2d[2][4] = {{1, 2, 3, 4}, {1, 2, 3, 3}};
const int kTec3d[4][8] = {{5, 6, 2, 1, 7, 8, 4, 3},
                          {1, 3, 6, 4, 2, 2, 5, 5},
                          {3, 3, 4, 1, 3, 3, 2, 2},
                          {1, 2, 4, 3, 5, 5, 5, 5}};

// grd_postpr (grd_constr.f90:79-85): the element's vertex list in Tecplot's
// order; an element matching no row keeps zeros
void postpr(int n_dx, const int *vl, int nv, int *out, int nout) {
  for (int k = 0; k < nout; ++k) out[k] = 0;
  const int *row = nullptr;
  if (n_dx == 2 && nv == 4) row = kTec2d[0];
  if (n_dx == 2 && nv == 3) row = kTec2d[1];
  if (n_dx == 3 && nv == 8) row = kTec3d[0];
  if (n_dx == 3 && nv == 6) row = kTec3d[1];
  if (n_dx == 3 && nv == 4) row = kTec3d[2];
  if (n_dx == 3 && nv == 5) row = kTec3d[3];
  if (!row) return;
  for (int k = 0; k < nout; ++k) out[k] = vl[row[k] - 1];
}

// the file name of eqn_outp / prp_outp: grd<i>%phs<j>%<kind><k>, with
// @NNNN (the output counter, zero-padded to 4) in transient runs
std::string title_of(const char *kind, int igrd, int iph, int ieq, int ict, int n_tl) {
  // write(ch,'(i1)') of each index: one digit, '*' when it does not fit
  auto i1 = [](int v) { return v >= 0 && v <= 9 ? std::to_string(v) : std::string("*"); };
  std::string t = "grd" + i1(igrd) + "%phs" + i1(iph) + "%" + kind + i1(ieq);
  if (n_tl > 1 && ict > 0) {
    std::string c = std::to_string(ict);
    if (c.size() > 4) c = "****";
    t += "@" + std::string(4 - c.size(), '0') + c;
  }
  return t;
}

}  // namespace

std::string tecplot_write(const TecplotIn &in, const char *dir, std::string &path) {
  const bool eqn = in.kind_eqn;
  const std::string title = title_of(eqn ? "eqn" : "prp", in.igrd, in.iph, in.ieq, in.ict, in.n_tl);
  path = std::string(dir ? dir : "") + title + ".dat";
  FILE *fp = std::fopen(path.c_str(), "w");
  if (!fp) return "cannot create " + path + ": " + std::strerror(errno);
  std::string dx_var, et_kind;
  int n_vrt = 0;
  if (in.n_dx == 2) { dx_var = "\"x_[1]\",\"x_[2]\""; et_kind = "quadrilateral"; n_vrt = 4; }
  if (in.n_dx == 3) { dx_var = "\"x_[1]\",\"x_[2]\",\"x_[3]\""; et_kind = "brick"; n_vrt = 8; }
  std::string phi_var;
  if (eqn) {
    for (int isc = 1; isc <= in.n_sc; ++isc)
      for (int ivc = 1; ivc <= in.n_vc; ++ivc)
        phi_var += ",\"phi_[" + std::to_string(isc) + "," + std::to_string(ivc) + "]\"";
  } else {
    phi_var = ",\"rho_[1]\"";
  }
  ListRecord r;
  std::string buf;
  buf.reserve(1 << 20);
  auto flush = [&]() {
    if (buf.size() >= (1u << 20)) {
      std::fwrite(buf.data(), 1, buf.size(), fp);
      buf.clear();
    }
  };
  r.chr("title = \"");
  r.chr(title);
  r.chr("\"");
  buf += r.take();
  r.chr("variables = " + dx_var + phi_var);
  buf += r.take();
  r.chr("zone n = ");
  r.integer(in.n_vrt);
  r.chr(", e = ");
  r.integer(in.n_vol);
  r.chr(", f = feblock, et = " + et_kind);
  buf += r.take();
  r.chr("varlocation = ([" + std::to_string(in.n_dx + 1) + "-99]=cellcentered)");
  buf += r.take();
  for (int d = 0; d < in.n_dx; ++d)
    for (int v = 0; v < in.n_vrt; ++v) {
      r.real(in.vx[(size_t)3 * v + d]);
      buf += r.take();
      flush();
    }
  // values: phi(isc, ivc) of every cell, component by component
  const int ncomp = eqn ? in.n_sc * in.n_vc : 1;
  for (int k = 0; k < ncomp; ++k)
    for (int i = 0; i < in.n_vol; ++i) {
      r.real(in.val[(size_t)i * ncomp + k]);
      buf += r.take();
      flush();
    }
  std::vector<int> vl(n_vrt);
  for (int i = 0; i < in.n_vol; ++i) {
    const int a = in.v_vptr[i], b = in.v_vptr[i + 1];
    std::vector<int> one(b - a);
    for (int k = a; k < b; ++k) one[k - a] = in.v_vrt[k] + 1;
    postpr(in.n_dx, one.data(), b - a, vl.data(), n_vrt);
    for (int k = 0; k < n_vrt; ++k) r.integer(vl[k]);
    buf += r.take();
    flush();
  }
  std::fwrite(buf.data(), 1, buf.size(), fp);
  if (std::fclose(fp) != 0) return "write error on " + path;
  return "";
}

// flowlab/src/mesh.cpp:298-412: header "$MeshFormat" with the gmsh version
// 2.2 0 8 (floats printed with %g), one string tag (the file name), one
// real tag (time 0), four integer tags (time step 0, components, elements,
// partition 0), then "index value(s)" per element with %g
static const float kGmshVer[3] = {2.2f, 0.f, 8.f};

std::string gmsh_write(const char *fname, int n, int n_cmp, int n_dim, const int *elm_idx, const double *phi) {
  FILE *fp = std::fopen(fname, "w");
  if (!fp) return std::string("cannot create ") + fname + ": " + std::strerror(errno);
  std::fprintf(fp, "$MeshFormat\n");
  std::fprintf(fp, "%g %g %g\n", kGmshVer[0], kGmshVer[1], kGmshVer[2]);
  std::fprintf(fp, "$EndMeshFormat\n");
  std::fprintf(fp, "$ElementData\n");
  std::fprintf(fp, "%d\n", 1);
  std::fprintf(fp, "%s\n", fname);
  std::fprintf(fp, "%d\n", 1);
  std::fprintf(fp, "%g\n", 0.);
  std::fprintf(fp, "%d\n", 4);
  std::fprintf(fp, "%d\n%d\n%d\n%d\n", 0, n_cmp, n, 0);
  for (int i = 0; i < n; ++i) {
    if (n_cmp == 1) {
      std::fprintf(fp, "%d %g\n", elm_idx[i], phi[i]);
    } else {
      double v[3] = {0.0, 0.0, 0.0};
      for (int j = 0; j < n_dim && j < 3; ++j) v[j] = phi[(size_t)n_dim * i + j];
      std::fprintf(fp, "%d %g %g %g\n", elm_idx[i], v[0], v[1], v[2]);
    }
  }
  std::fprintf(fp, "$EndElementData\n");
  if (std::fclose(fp) != 0) return std::string("write error on ") + fname;
  return "";
}

}  // namespace fvk
