// C ABI of the fvk library (include/fvk.h): mesh, device context, solver,
// gradient and SIMPLE entry points.  Host arrays use the reference's
// numbering and layouts; everything on the device is in row order.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <numeric>
#include <string>
#include <vector>

#include "flow.h"
#include "mesh.h"
#include "output.h"

struct fvk_mesh {
  fvk::Mesh m;
};
struct fvk_flow {
  fvk::Flow f;
};

namespace fvk {

static thread_local std::string g_err;
void set_error(const std::string &msg) { g_err = msg; }


// ---------------------------------------------------------------- helpers
struct Topo {  // host topology + geometry, 0-based, vectors stride 3
  int n_dx = 0, n = 0, nf = 0, nbf = 0;
  std::vector<int> fip, fin, fib, fbty, cptr, clst;
  std::vector<double> vx, vv, vvr, fx, fn, fdp, fdn, fs, flpn, fwfp;
};

static Topo topo_from_mesh(const Mesh &m) {
  Topo t;
  t.n_dx = m.n_dx; t.n = m.n_vol; t.nf = m.n_fce; t.nbf = m.n_bf;
  t.fip = m.f_ip; t.fin = m.f_in; t.fib = m.f_ib; t.fbty = m.f_bty;
  t.cptr = m.v_fptr; t.clst = m.fce_lst;
  t.vx = m.vol_x; t.vv = m.vol_v; t.vvr = m.vol_vr;
  t.fx = m.f_x; t.fn = m.f_n; t.fdp = m.f_dxp; t.fdn = m.f_dxn;
  t.fs = m.f_s; t.flpn = m.f_lpn; t.fwfp = m.f_wfp;
  return t;
}

// Reverse Cuthill-McKee order of the cell graph (George & Liu): start from a
// pseudo-peripheral cell of minimum degree, visit neighbours by ascending
// (degree, id), reverse the visit order.  Deterministic.
static std::vector<int> rcm_order(const Topo &t) {
  const int n = t.n;
  std::vector<int> deg(n, 0), xadj(n + 1, 0), adj;
  for (int f = 0; f < t.nf; ++f)
    if (t.fin[f] >= 0) { deg[t.fip[f]]++; deg[t.fin[f]]++; }
  for (int i = 0; i < n; ++i) xadj[i + 1] = xadj[i] + deg[i];
  adj.resize(xadj[n]);
  std::vector<int> fill(n, 0);
  for (int f = 0; f < t.nf; ++f)
    if (t.fin[f] >= 0) {
      int a = t.fip[f], b = t.fin[f];
      adj[xadj[a] + fill[a]++] = b;
      adj[xadj[b] + fill[b]++] = a;
    }
  for (int i = 0; i < n; ++i)
    std::sort(adj.begin() + xadj[i], adj.begin() + xadj[i + 1],
              [&](int a, int b) { return deg[a] != deg[b] ? deg[a] < deg[b] : a < b; });
  std::vector<int> order, lvl(n, -1);
  order.reserve(n);
  std::vector<char> seen(n, 0);
  auto bfs_last = [&](int s, std::vector<int> &q) {
    q.clear();
    q.push_back(s);
    std::vector<int> mark;
    lvl[s] = 0;
    mark.push_back(s);
    for (size_t h = 0; h < q.size(); ++h) {
      int v = q[h];
      for (int k = xadj[v]; k < xadj[v + 1]; ++k)
        if (lvl[adj[k]] < 0) { lvl[adj[k]] = lvl[v] + 1; q.push_back(adj[k]); mark.push_back(adj[k]); }
    }
    int last = lvl[q.back()], best = q.back();
    for (int v : q)
      if (lvl[v] == last && (deg[v] < deg[best] || (deg[v] == deg[best] && v < best))) best = v;
    int ecc = last;
    for (int v : mark) lvl[v] = -1;
    return std::make_pair(best, ecc);
  };
  std::vector<int> q;
  for (int s0 = 0; s0 < n; ++s0) {
    if (seen[s0]) continue;
    // component of s0: pick its minimum-degree cell, then walk to a
    // pseudo-peripheral one
    bfs_last(s0, q);
    int s = s0;
    for (int v : q)
      if (deg[v] < deg[s] || (deg[v] == deg[s] && v < s)) s = v;
    int ecc = -1;
    for (int rep = 0; rep < 4; ++rep) {
      auto pr = bfs_last(s, q);
      if (pr.second <= ecc) break;
      ecc = pr.second;
      s = pr.first;
    }
    size_t base = order.size();
    order.push_back(s);
    seen[s] = 1;
    for (size_t h = base; h < order.size(); ++h) {
      int v = order[h];
      for (int k = xadj[v]; k < xadj[v + 1]; ++k)
        if (!seen[adj[k]]) { seen[adj[k]] = 1; order.push_back(adj[k]); }
    }
  }
  std::reverse(order.begin(), order.end());
  return order;
}

template <class T>
static int upload(Ctx &c, T **dst, const std::vector<T> &src) {
  int rc = dalloc(c, dst, src.size());
  if (rc) return rc;
  if (!src.empty()) FVK_HIP(hipMemcpy(*dst, src.data(), src.size() * sizeof(T), hipMemcpyHostToDevice));
  return 0;
}

static int build_levels(Ctx &c, const Topo &t, const std::vector<int> &iperm, bool fwd, Levels &L) {
  const int n = t.n;
  std::vector<int> lev(n, 0);
  int nl = 0;
  for (int s = 0; s < n; ++s) {
    const int o = fwd ? s : n - 1 - s;
    int l = 0;
    for (int k = t.cptr[o]; k < t.cptr[o + 1]; ++k) {
      const int f = t.clst[k];
      if (t.fin[f] < 0) continue;
      const int j = t.fip[f] == o ? t.fin[f] : t.fip[f];
      if (fwd ? j < o : j > o) l = std::max(l, lev[j] + 1);
    }
    lev[o] = l;
    nl = std::max(nl, l + 1);
  }
  L.ptr.assign(nl + 1, 0);
  for (int o = 0; o < n; ++o) L.ptr[lev[o] + 1]++;
  for (int l = 0; l < nl; ++l) L.ptr[l + 1] += L.ptr[l];
  std::vector<int> rows(n), pos(L.ptr.begin(), L.ptr.end() - 1);
  // rows of a level in ascending row order
  std::vector<int> byrow(n);
  for (int o = 0; o < n; ++o) byrow[iperm[o]] = o;
  for (int r = 0; r < n; ++r) {
    int o = byrow[r];
    rows[pos[lev[o]]++] = r;
  }
  return upload(c, &L.rows, rows);
}

// gk of DMesh: for the neighbour side of an interior face, the slot of the
// face in the owner row's list
__global__ __launch_bounds__(BS) void k_slot_owner(DMesh M, signed char *gk) {
  const size_t total = (size_t)M.W * M.n;
  for (size_t o = blockIdx.x * (size_t)BS + threadIdx.x; o < total; o += (size_t)gridDim.x * BS) {
    const int e = M.cf[o];
    signed char k = -1;
    if (e != PAD && e < 0) {
      const int f = ~e, r = M.fip[f];
      for (int kk = 0; kk < M.W; ++kk)
        if (M.cf[(size_t)kk * M.n + r] == f) k = (signed char)kk;
    }
    gk[o] = k;
  }
}

// Per-slot copies of the face geometry (the slot-major tables of DMesh);
// boundary slots take the face centre minus the cell centre in gp.
__global__ __launch_bounds__(BS) void k_slot_geom(DMesh M, double *gs, double *gnx, double *gny, double *gnz,
                                                  double *gw, double *gpx, double *gpy, double *gpz, double *gqx,
                                                  double *gqy, double *gqz, int *go) {
  const size_t total = (size_t)M.W * M.n;
  for (size_t o = blockIdx.x * (size_t)BS + threadIdx.x; o < total; o += (size_t)gridDim.x * BS) {
    const int e = M.cf[o], i = (int)(o % (size_t)M.n);
    double v[11] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
    int other = -1;
    if (e != PAD) {
      const bool own = e >= 0;
      const int f = own ? e : ~e;
      v[0] = M.fs[f];
      v[1] = M.fnx[f]; v[2] = M.fny[f]; v[3] = M.fnz[f];
      if (M.fin[f] >= 0) {
        other = own ? M.fin[f] : M.fip[f];
        v[4] = M.fwfp[f];
        v[5] = M.dpx[f]; v[6] = M.dpy[f]; v[7] = M.dpz[f];
        v[8] = M.dnx[f]; v[9] = M.dny[f]; v[10] = M.dnz[f];
      } else {
        other = ~M.fib[f];
        v[5] = M.fx[f] - M.cx[i]; v[6] = M.fy[f] - M.cy[i]; v[7] = M.fz[f] - M.cz[i];
      }
    }
    gs[o] = v[0]; gnx[o] = v[1]; gny[o] = v[2]; gnz[o] = v[3]; gw[o] = v[4];
    gpx[o] = v[5]; gpy[o] = v[6]; gpz[o] = v[7]; gqx[o] = v[8]; gqy[o] = v[9]; gqz[o] = v[10];
    go[o] = other;
  }
}

// Scratch of fvk_solv_res_dev, the entry point that allocates nothing: the
// per-block partial sums of the residual kernel (6 per block, at most RG
// blocks: 48 KiB).  Each GPU holds its own copy until the process ends.
// Host code never uses the array's name as an address: the device address
// of each GPU's copy is taken once per device, with that device current
// (scratch_init, from the context build), and kept in g_scratch.
static __device__ __align__(256) double scratch[6 * RG];
static std::mutex g_scratch_mu;
static const int kMaxDevices = 64;
static double *g_scratch[kMaxDevices] = {};

static int scratch_init(int device) {
  if (device < 0 || device >= kMaxDevices) {
    set_error("device index beyond the scratch table");
    return FVK_ERR_ARG;
  }
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (g_scratch[device]) return 0;
  double *ptr = nullptr;
  hipError_t e = hipGetSymbolAddress((void **)&ptr, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !ptr) {
    set_error(std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    return FVK_ERR_HIP;
  }
  g_scratch[device] = ptr;
  return 0;
}

static int ctx_build(Ctx &c, const Topo &t, int device, int flags) {
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= device) {
    set_error("no HIP device");
    return FVK_ERR_NODEV;
  }
  c.device = device;
  FVK_HIP(hipSetDevice(device));
  {
    int rc = scratch_init(device);
    if (rc) return rc;
  }
  FVK_HIP(hipStreamCreateWithFlags(&c.stream, hipStreamNonBlocking));
  FVK_HIP(hipEventCreate(&c.ev0));
  FVK_HIP(hipEventCreate(&c.ev1));
  FVK_HIP(hipHostMalloc((void **)&c.poll_h, 2 * sizeof(KS), hipHostMallocDefault));
  for (int k = 0; k < 2; ++k) FVK_HIP(hipEventCreateWithFlags(&c.poll_ev[k], hipEventDisableTiming));
  const int n = t.n, nf = t.nf;
  c.n = n; c.nf = nf; c.nbf = t.nbf; c.ndx = t.n_dx;
  // row order
  if (flags & FVK_CTX_RCM) c.perm = rcm_order(t);
  else { c.perm.resize(n); std::iota(c.perm.begin(), c.perm.end(), 0); }
  c.iperm.assign(n, 0);
  for (int r = 0; r < n; ++r) c.iperm[c.perm[r]] = r;
  // cell-face table and interior-slot ELL
  int W = 0, Ws = 0;
  for (int o = 0; o < n; ++o) {
    int a = t.cptr[o + 1] - t.cptr[o], b = 0;
    for (int k = t.cptr[o]; k < t.cptr[o + 1]; ++k) b += t.fin[t.clst[k]] >= 0;
    W = std::max(W, a);
    Ws = std::max(Ws, b);
  }
  if (Ws == 0) Ws = 1;
  c.W = W; c.Ws = Ws;
  const int ld = slot_ld(n);
  std::vector<int> cf((size_t)W * n, PAD), spc((size_t)Ws * ld, -1), spf((size_t)Ws * ld, -1);
  for (int r = 0; r < n; ++r) {
    const int o = c.perm[r];
    int k = 0, ks = 0;
    for (int s = t.cptr[o]; s < t.cptr[o + 1]; ++s, ++k) {
      const int f = t.clst[s];
      const bool own = t.fip[f] == o;
      cf[(size_t)k * n + r] = own ? f : ~f;
      if (t.fin[f] >= 0) {
        const int j = own ? t.fin[f] : t.fip[f];
        spc[(size_t)ks * ld + r] = c.iperm[j];
        spf[(size_t)ks * ld + r] = f;
        ++ks;
      }
    }
  }
  std::vector<int> fip(nf), fin(nf);
  for (int f = 0; f < nf; ++f) {
    fip[f] = c.iperm[t.fip[f]];
    fin[f] = t.fin[f] >= 0 ? c.iperm[t.fin[f]] : -1;
  }
  c.h_fip = fip; c.h_fin = fin; c.h_fib = t.fib; c.h_fbty = t.fbty;
  std::vector<double> cx(n), cy(n), cz(n), v(n), vr(n);
  for (int r = 0; r < n; ++r) {
    const int o = c.perm[r];
    cx[r] = t.vx[3 * (size_t)o]; cy[r] = t.vx[3 * (size_t)o + 1]; cz[r] = t.vx[3 * (size_t)o + 2];
    v[r] = t.vv[o]; vr[r] = t.vvr[o];
  }
  auto comp = [&](const std::vector<double> &a, int d) {
    std::vector<double> o(nf);
    for (int f = 0; f < nf; ++f) o[f] = a[3 * (size_t)f + d];
    return o;
  };
  DMesh &M = c.M;
  M.n = n; M.nf = nf; M.nbf = t.nbf; M.W = W; M.Ws = Ws; M.ndx = t.n_dx; M.ld = ld;
  double *p;
  int *ip;
  int rc = 0;
#define UP(field, vec) do { if ((rc = upload(c, &p, vec))) return rc; M.field = p; } while (0)
#define UPI(field, vec) do { if ((rc = upload(c, &ip, vec))) return rc; M.field = ip; } while (0)
  UP(cx, cx); UP(cy, cy); UP(cz, cz); UP(v, v); UP(vr, vr);
  UPI(rank, c.perm); UPI(fip, fip); UPI(fin, fin); UPI(fib, t.fib); UPI(fbty, t.fbty);
  UP(fs, t.fs); UP(flpn, t.flpn); UP(fwfp, t.fwfp);
  UP(fx, comp(t.fx, 0)); UP(fy, comp(t.fx, 1)); UP(fz, comp(t.fx, 2));
  UP(fnx, comp(t.fn, 0)); UP(fny, comp(t.fn, 1)); UP(fnz, comp(t.fn, 2));
  UP(dpx, comp(t.fdp, 0)); UP(dpy, comp(t.fdp, 1)); UP(dpz, comp(t.fdp, 2));
  UP(dnx, comp(t.fdn, 0)); UP(dny, comp(t.fdn, 1)); UP(dnz, comp(t.fdn, 2));
  UPI(cf, cf); UPI(spc, spc); UPI(spf, spf);
#undef UP
#undef UPI
  {
    const size_t ns = (size_t)W * n;
    double *g[11];
    int *go = nullptr;
    for (int k = 0; k < 11; ++k)
      if ((rc = dalloc(c, &g[k], ns))) return rc;
    if ((rc = dalloc(c, &go, ns))) return rc;
    const int gb = (int)std::max<size_t>(1, std::min<size_t>((ns + BS - 1) / BS, 65536));
    k_slot_geom<<<gb, BS, 0, c.stream>>>(M, g[0], g[1], g[2], g[3], g[4], g[5], g[6], g[7], g[8], g[9], g[10], go);
    FVK_HIP(hipGetLastError());
    M.gs = g[0]; M.gnx = g[1]; M.gny = g[2]; M.gnz = g[3]; M.gw = g[4];
    M.gpx = g[5]; M.gpy = g[6]; M.gpz = g[7]; M.gqx = g[8]; M.gqy = g[9]; M.gqz = g[10];
    M.go = go;
    signed char *gk = nullptr;
    if ((rc = dalloc(c, &gk, ns))) return rc;
    k_slot_owner<<<gb, BS, 0, c.stream>>>(M, gk);
    FVK_HIP(hipGetLastError());
    M.gk = gk;
  }
  if ((rc = build_levels(c, t, c.iperm, true, c.lf))) return rc;
  if ((rc = build_levels(c, t, c.iperm, false, c.lb))) return rc;
  if ((rc = dalloc(c, &c.diag, n))) return rc;
  if ((rc = dalloc(c, &c.spv, (size_t)Ws * ld))) return rc;
  if ((rc = dalloc(c, &c.ngp, nf))) return rc;
  if ((rc = dalloc(c, &c.ngn, nf))) return rc;
  if ((rc = dalloc(c, &c.pc_c, n))) return rc;
  if ((rc = dalloc(c, &c.pc_cr, n))) return rc;
  if ((rc = dalloc(c, &c.sweep_t, (size_t)3 * n))) return rc;
  if ((rc = dalloc(c, &c.sweep_int, 4))) return rc;
  if ((rc = dalloc(c, &c.sweep_lmask, n))) return rc;
  if ((rc = dalloc(c, &c.sweep_umask, n))) return rc;
  FVK_HIP(hipDeviceGetAttribute(&c.n_cu, hipDeviceAttributeMultiprocessorCount, c.device));
  if (Ws > 8) c.sweep_one_launch = false;  // slot masks hold 8 slots
  if ((rc = sweep_masks_build(c))) return rc;
  if ((rc = sweep_parts_build(c))) return rc;
  if ((rc = dalloc(c, &c.wk, (size_t)8 * wk_ld(n)))) return rc;
  if ((rc = dalloc(c, &c.ks, 3))) return rc;
  if ((rc = dalloc(c, &c.ksr, 6))) return rc;
  if ((rc = dalloc(c, &c.part, (size_t)4 * RG * 8))) return rc;
  if ((rc = dalloc(c, &c.stage, (size_t)3 * std::max(n, nf)))) return rc;
  FVK_HIP(hipStreamSynchronize(c.stream));
  return 0;
}

static void ctx_free(Ctx &c) {
  if (c.stream) hipStreamSynchronize(c.stream);
  for (void *p : c.owned) hipFree(p);
  c.owned.clear();
  if (c.ev0) hipEventDestroy(c.ev0);
  if (c.ev1) hipEventDestroy(c.ev1);
  for (int k = 0; k < 2; ++k)
    if (c.poll_ev[k]) hipEventDestroy(c.poll_ev[k]);
  if (c.poll_h) hipHostFree(c.poll_h);
  c.poll_h = nullptr;
  if (c.stream) hipStreamDestroy(c.stream);
  c.stream = nullptr;
}

// reference entity-major host array (stride ld, ncomp components) <-> device
// rows (component-major)
static int to_dev(Ctx &c, const double *h, int ncomp, int ld, int count, const std::vector<int> *perm, double *d) {
  std::vector<double> tmp((size_t)ncomp * count);
  for (int r = 0; r < count; ++r) {
    const int o = perm ? (*perm)[r] : r;
    for (int q = 0; q < ncomp; ++q) tmp[(size_t)q * count + r] = h[(size_t)o * ld + q];
  }
  FVK_HIP(hipStreamSynchronize(c.stream));
  FVK_HIP(hipMemcpy(d, tmp.data(), tmp.size() * sizeof(double), hipMemcpyHostToDevice));
  return 0;
}
static int to_host(Ctx &c, const double *d, int ncomp, int ld, int count, const std::vector<int> *perm, double *h) {
  std::vector<double> tmp((size_t)ncomp * count);
  FVK_HIP(hipStreamSynchronize(c.stream));
  FVK_HIP(hipMemcpy(tmp.data(), d, tmp.size() * sizeof(double), hipMemcpyDeviceToHost));
  for (int r = 0; r < count; ++r) {
    const int o = perm ? (*perm)[r] : r;
    for (int q = 0; q < ncomp; ++q) h[(size_t)o * ld + q] = tmp[(size_t)q * count + r];
  }
  return 0;
}

__global__ __launch_bounds__(BS) void k_spv_from_faces(DMesh M, const double *ngp, const double *ngn, double *spv) {
  const int i = blockIdx.x * BS + threadIdx.x;
  if (i >= M.n) return;
  int ks = 0;
  for (int k = 0; k < M.W; ++k) {
    const int e = M.cf[k * M.n + i];
    if (e == PAD) break;
    const int f = e >= 0 ? e : ~e;
    if (M.fin[f] < 0) continue;
    spv[(size_t)ks * M.ld + i] = e >= 0 ? ngp[f] : ngn[f];
    ++ks;
  }
}

}  // namespace fvk

using namespace fvk;

#define GUARD(x) \
  do {           \
    int rc_ = (x); \
    if (rc_) return rc_; \
  } while (0)

extern "C" {

const char *fvk_last_error(void) { return g_err.c_str(); }
int fvk_version(void) { return 1; }

// ---------------------------------------------------------------- mesh
// Device on which mesh constructors build connectivity and geometry (-1:
// the host)
static int g_mesh_device = -1;

static std::string finish_sel(fvk::Mesh &m) {
  return g_mesh_device >= 0 ? finish_device(m, g_mesh_device) : finish(m);
}

static int mesh_done(fvk_mesh *m, const std::string &e, fvk_mesh **out) {
  if (!e.empty()) {
    set_error(e);
    delete m;
    return FVK_ERR_ARG;
  }
  *out = m;
  return FVK_OK;
}

int fvk_mesh_read_neu(const char *path, double l_sc, fvk_mesh **out) {
  if (!path || !out) { set_error("null argument"); return FVK_ERR_ARG; }
  fvk_mesh *m = new fvk_mesh;
  std::string e = read_neu(m->m, path, l_sc);
  if (!e.empty()) {
    set_error(e);
    delete m;
    return e.rfind("cannot open", 0) == 0 ? FVK_ERR_IO : FVK_ERR_ARG;
  }
  return mesh_done(m, finish_sel(m->m), out);
}

int fvk_mesh_read_gmsh(const char *path, int n_dx, double l_sc, const int *code_map, int n_map, fvk_mesh **out) {
  if (!path || !out || n_map < 0 || (n_map > 0 && !code_map)) { set_error("bad argument"); return FVK_ERR_ARG; }
  fvk_mesh *m = new fvk_mesh;
  std::string e = read_gmsh(m->m, path, n_dx, l_sc, code_map, n_map);
  if (!e.empty()) {
    set_error(e);
    delete m;
    return e.rfind("cannot open", 0) == 0 ? FVK_ERR_IO : FVK_ERR_ARG;
  }
  return mesh_done(m, finish_sel(m->m), out);
}

int fvk_mesh_set_build_device(int device) {
  if (device < -1) { set_error("bad device"); return FVK_ERR_ARG; }
  g_mesh_device = device;
  return FVK_OK;
}

int fvk_mesh_get_elm_idx(const fvk_mesh *m, int *elm_idx) {
  if (!m || !elm_idx) { set_error("null argument"); return FVK_ERR_ARG; }
  for (int c = 0; c < m->m.n_vol; ++c) elm_idx[c] = m->m.elm_idx.empty() ? c + 1 : m->m.elm_idx[c];
  return FVK_OK;
}

int fvk_mesh_get_cells(const fvk_mesh *m, double *vx3, int *v_ty, int *cell_vptr, int *cell_vrt) {
  if (!m) { set_error("null argument"); return FVK_ERR_ARG; }
  const fvk::Mesh &M = m->m;
  if (vx3) std::memcpy(vx3, M.vx.data(), sizeof(double) * M.vx.size());
  if (v_ty) std::memcpy(v_ty, M.v_ty.data(), sizeof(int) * M.v_ty.size());
  if (cell_vptr) std::memcpy(cell_vptr, M.v_vptr.data(), sizeof(int) * M.v_vptr.size());
  if (cell_vrt)
    for (size_t k = 0; k < M.v_vrt.size(); ++k) cell_vrt[k] = M.v_vrt[k] + 1;
  return FVK_OK;
}

int fvk_mesh_box(int nx, int ny, int nz, double lx, double ly, double lz, double jitter, uint64_t seed, int lid_code,
                 int wall_code, fvk_mesh **out) {
  if (!out) { set_error("null argument"); return FVK_ERR_ARG; }
  fvk_mesh *m = new fvk_mesh;
  std::string e = make_box(m->m, nx, ny, nz, lx, ly, lz, jitter, seed, lid_code, wall_code);
  if (e.empty()) e = finish_sel(m->m);
  return mesh_done(m, e, out);
}

int fvk_mesh_from_cells(int n_dx, int n_vrt, const double *vx3, int n_vol, const int *v_ty, const int *v_vptr,
                        const int *v_vrt, const int *bnd, fvk_mesh **out) {
  if (!vx3 || !v_ty || !v_vptr || !v_vrt || !bnd || !out || n_vol < 1 || (n_dx != 2 && n_dx != 3)) {
    set_error("bad argument");
    return FVK_ERR_ARG;
  }
  fvk_mesh *m = new fvk_mesh;
  std::string e = init_cells(m->m, n_dx, n_vrt, vx3, n_vol, v_ty, v_vptr, v_vrt, bnd);
  if (e.empty()) e = finish_sel(m->m);
  return mesh_done(m, e, out);
}

int fvk_mesh_get_info(const fvk_mesh *mm, fvk_mesh_info *info) {
  if (!mm || !info) { set_error("null argument"); return FVK_ERR_ARG; }
  const Mesh &m = mm->m;
  info->n_dx = m.n_dx; info->n_vrt = m.n_vrt; info->n_vol = m.n_vol; info->n_fce = m.n_fce;
  info->n_bf = m.n_bf; info->n_slot = m.n_slot; info->max_fce = m.max_fce;
  return FVK_OK;
}

int fvk_mesh_get_topology(const fvk_mesh *mm, int *fce_ip, int *fce_in, int *fce_ib, int *fce_bty, int *cell_ptr,
                          int *fce_lst, int *ng_lst) {
  if (!mm) { set_error("null argument"); return FVK_ERR_ARG; }
  const Mesh &m = mm->m;
  for (int f = 0; f < m.n_fce; ++f) {
    if (fce_ip) fce_ip[f] = m.f_ip[f] + 1;
    if (fce_in) fce_in[f] = m.f_in[f] + 1;
    if (fce_ib) fce_ib[f] = m.f_ib[f] >= 0 ? m.n_vol + 1 + m.f_ib[f] : 0;
    if (fce_bty) fce_bty[f] = m.f_bty[f];
  }
  if (cell_ptr) for (int c = 0; c <= m.n_vol; ++c) cell_ptr[c] = m.v_fptr[c];
  for (int s = 0; s < m.n_slot; ++s) {
    if (fce_lst) fce_lst[s] = m.fce_lst[s] + 1;
    if (ng_lst) ng_lst[s] = m.ng_lst[s] + 1;
  }
  return FVK_OK;
}

int fvk_mesh_get_geometry(const fvk_mesh *mm, double *vol_x, double *vol_v, double *fce_x, double *fce_n,
                          double *fce_dxp, double *fce_dxn, double *fce_s, double *fce_lpn, double *fce_wfp) {
  if (!mm) { set_error("null argument"); return FVK_ERR_ARG; }
  const Mesh &m = mm->m;
  const int nd = m.n_dx;
  for (int c = 0; c < m.n_vol; ++c) {
    for (int d = 0; d < nd; ++d)
      if (vol_x) vol_x[(size_t)c * nd + d] = m.vol_x[3 * (size_t)c + d];
    if (vol_v) vol_v[c] = m.vol_v[c];
  }
  for (int f = 0; f < m.n_fce; ++f) {
    for (int d = 0; d < nd; ++d) {
      const size_t o = (size_t)f * nd + d, i = 3 * (size_t)f + d;
      if (fce_x) fce_x[o] = m.f_x[i];
      if (fce_n) fce_n[o] = m.f_n[i];
      if (fce_dxp) fce_dxp[o] = m.f_dxp[i];
      if (fce_dxn) fce_dxn[o] = m.f_dxn[i];
    }
    if (fce_s) fce_s[f] = m.f_s[f];
    if (fce_lpn) fce_lpn[f] = m.f_lpn[f];
    if (fce_wfp) fce_wfp[f] = m.f_wfp[f];
  }
  return FVK_OK;
}

void fvk_mesh_free(fvk_mesh *m) { delete m; }

// ---------------------------------------------------------------- context
int fvk_ctx_create(const fvk_mesh *m, int device, int flags, fvk_ctx **out) {
  if (!m || !out) { set_error("null argument"); return FVK_ERR_ARG; }
  fvk_ctx *c = new fvk_ctx;
  int rc = ctx_build(c->c, topo_from_mesh(m->m), device, flags);
  if (rc) {
    ctx_free(c->c);
    delete c;
    return rc;
  }
  *out = c;
  return FVK_OK;
}

int fvk_ctx_create_arrays(int n_dx, int n_vol, int n_fce, int n_bf, const int *fce_ip, const int *fce_in,
                          const int *fce_ib, const int *fce_bty, const int *cell_ptr, const int *fce_lst,
                          const double *vol_x, const double *vol_v, const double *fce_x, const double *fce_n,
                          const double *fce_dxp, const double *fce_dxn, const double *fce_s, const double *fce_lpn,
                          const double *fce_wfp, int device, int flags, fvk_ctx **out) {
  if (!out || !fce_ip || !fce_in || !fce_ib || !fce_bty || !cell_ptr || !fce_lst || !vol_x || !vol_v || !fce_x ||
      !fce_n || !fce_dxp || !fce_dxn || !fce_s || !fce_lpn || !fce_wfp || (n_dx != 2 && n_dx != 3) || n_vol < 1) {
    set_error("bad argument");
    return FVK_ERR_ARG;
  }
  Topo t;
  t.n_dx = n_dx; t.n = n_vol; t.nf = n_fce; t.nbf = n_bf;
  t.fip.resize(n_fce); t.fin.resize(n_fce); t.fib.resize(n_fce); t.fbty.assign(fce_bty, fce_bty + n_fce);
  for (int f = 0; f < n_fce; ++f) {
    t.fip[f] = fce_ip[f] - 1;
    t.fin[f] = fce_in[f] - 1;
    t.fib[f] = fce_ib[f] > 0 ? fce_ib[f] - n_vol - 1 : -1;
    if (t.fip[f] < 0 || t.fip[f] >= n_vol || t.fin[f] >= n_vol || (t.fin[f] < 0 && (t.fib[f] < 0 || t.fib[f] >= n_bf))) {
      set_error("face index out of range");
      return FVK_ERR_ARG;
    }
  }
  t.cptr.assign(cell_ptr, cell_ptr + n_vol + 1);
  t.clst.resize(cell_ptr[n_vol]);
  for (int s = 0; s < cell_ptr[n_vol]; ++s) {
    t.clst[s] = fce_lst[s] - 1;
    if (t.clst[s] < 0 || t.clst[s] >= n_fce) { set_error("fce_lst entry out of range"); return FVK_ERR_ARG; }
  }
  auto v3 = [&](const double *a, int cnt) {
    std::vector<double> o(3 * (size_t)cnt, 0.0);
    for (int k = 0; k < cnt; ++k)
      for (int d = 0; d < n_dx; ++d) o[3 * (size_t)k + d] = a[(size_t)k * n_dx + d];
    return o;
  };
  t.vx = v3(vol_x, n_vol);
  t.vv.assign(vol_v, vol_v + n_vol);
  t.vvr.resize(n_vol);
  for (int c = 0; c < n_vol; ++c) t.vvr[c] = 1.0 / t.vv[c];  // grd_geom.f90:216
  t.fx = v3(fce_x, n_fce); t.fn = v3(fce_n, n_fce); t.fdp = v3(fce_dxp, n_fce); t.fdn = v3(fce_dxn, n_fce);
  t.fs.assign(fce_s, fce_s + n_fce); t.flpn.assign(fce_lpn, fce_lpn + n_fce); t.fwfp.assign(fce_wfp, fce_wfp + n_fce);
  fvk_ctx *c = new fvk_ctx;
  int rc = ctx_build(c->c, t, device, flags);
  if (rc) {
    ctx_free(c->c);
    delete c;
    return rc;
  }
  *out = c;
  return FVK_OK;
}

void fvk_ctx_destroy(fvk_ctx *c) {
  if (!c) return;
  ctx_free(c->c);
  delete c;
}

int fvk_ctx_sync(fvk_ctx *c) {
  if (!c) return FVK_ERR_ARG;
  FVK_HIP(hipStreamSynchronize(c->c.stream));
  return sweep_status(c->c);
}

int fvk_mesh_rcm(const fvk_mesh *m, int *perm) {
  if (!m || !perm) { set_error("null argument"); return FVK_ERR_ARG; }
  const std::vector<int> o = rcm_order(topo_from_mesh(m->m));
  for (size_t r = 0; r < o.size(); ++r) perm[r] = o[r] + 1;
  return FVK_OK;
}

int fvk_ctx_get_perm(const fvk_ctx *c, int *perm) {
  if (!c || !perm) return FVK_ERR_ARG;
  for (int r = 0; r < c->c.n; ++r) perm[r] = c->c.perm[r] + 1;
  return FVK_OK;
}

int fvk_ctx_levels(const fvk_ctx *c, int *n_fwd, int *n_bwd) {
  if (!c) return FVK_ERR_ARG;
  if (n_fwd) *n_fwd = c->c.lf.count();
  if (n_bwd) *n_bwd = c->c.lb.count();
  return FVK_OK;
}

// ---------------------------------------------------------------- solvers
int fvk_set_matrix(fvk_ctx *cc, const double *pl, const double *ng_p, const double *ng_n) {
  if (!cc || !pl || !ng_p || !ng_n) { set_error("null argument"); return FVK_ERR_ARG; }
  Ctx &c = cc->c;
  GUARD(to_dev(c, pl, 1, 1, c.n, &c.perm, c.diag));
  FVK_HIP(hipStreamSynchronize(c.stream));
  FVK_HIP(hipMemcpy(c.ngp, ng_p, sizeof(double) * c.nf, hipMemcpyHostToDevice));
  FVK_HIP(hipMemcpy(c.ngn, ng_n, sizeof(double) * c.nf, hipMemcpyHostToDevice));
  k_spv_from_faces<<<grid_for(c.n), BS, 0, c.stream>>>(c.M, c.ngp, c.ngn, c.spv);
  FVK_HIP(hipGetLastError());
  c.pc_valid = false;
  return FVK_OK;
}

int fvk_solv_precond(fvk_ctx *cc, int imat, double *cv, double *c_r) {
  if (!cc || imat < 0 || imat > 2) { set_error("bad argument"); return FVK_ERR_ARG; }
  Ctx &c = cc->c;
  GUARD(precond_build(c, imat));
  GUARD(sweep_status(c));
  if (cv) GUARD(to_host(c, c.pc_c, 1, 1, c.n, &c.perm, cv));
  if (c_r) GUARD(to_host(c, c.pc_cr, 1, 1, c.n, &c.perm, c_r));
  return FVK_OK;
}

int fvk_solv_direct(fvk_ctx *cc, int imat, const double *cv, const double *c_r, const double *b, double *x) {
  if (!cc || !b || !x || imat < 0 || imat > 2 || (!cv) != (!c_r)) { set_error("bad argument"); return FVK_ERR_ARG; }
  Ctx &c = cc->c;
  if (cv) {
    GUARD(to_dev(c, cv, 1, 1, c.n, &c.perm, c.pc_c));
    GUARD(to_dev(c, c_r, 1, 1, c.n, &c.perm, c.pc_cr));
    c.pc_imat = imat;
    c.pc_valid = true;
  } else if (!c.pc_valid || c.pc_imat != imat) {
    GUARD(precond_build(c, imat));
  }
  double *db = wk_slot(c, 5), *dx = wk_slot(c, 6);
  GUARD(to_dev(c, b, 1, 1, c.n, &c.perm, db));
  GUARD(precond_apply(c, imat, 1, db, dx, nullptr));
  GUARD(sweep_status(c));
  GUARD(to_host(c, dx, 1, 1, c.n, &c.perm, x));
  return FVK_OK;
}

static int col_to_dev(Ctx &c, const double *res0, int nvc, double *d) {
  // res0(n_vol, n_vc) column-major -> component-major rows
  std::vector<double> tmp((size_t)nvc * c.n);
  for (int q = 0; q < nvc; ++q)
    for (int r = 0; r < c.n; ++r) tmp[(size_t)q * c.n + r] = res0[(size_t)q * c.n + c.perm[r]];
  FVK_HIP(hipStreamSynchronize(c.stream));
  FVK_HIP(hipMemcpy(d, tmp.data(), tmp.size() * sizeof(double), hipMemcpyHostToDevice));
  return 0;
}
static int col_to_host(Ctx &c, const double *d, int nvc, double *res0) {
  std::vector<double> tmp((size_t)nvc * c.n);
  FVK_HIP(hipStreamSynchronize(c.stream));
  FVK_HIP(hipMemcpy(tmp.data(), d, tmp.size() * sizeof(double), hipMemcpyDeviceToHost));
  for (int q = 0; q < nvc; ++q)
    for (int r = 0; r < c.n; ++r) res0[(size_t)q * c.n + c.perm[r]] = tmp[(size_t)q * c.n + r];
  return 0;
}

int fvk_solv_res(fvk_ctx *cc, int n_vc, const double *phi, double *res0, double *res0_n, int not_r0) {
  if (!cc || !res0 || !res0_n || n_vc < 1 || n_vc > 3 || (!not_r0 && !phi)) { set_error("bad argument"); return FVK_ERR_ARG; }
  Ctx &c = cc->c;
  double *dx = wk_slot(c, 5), *dr = wk_slot(c, 6);
  if (!not_r0) GUARD(to_dev(c, phi, n_vc, n_vc, c.n, &c.perm, dx));
  GUARD(col_to_dev(c, res0, n_vc, dr));
  GUARD(residual(c, n_vc, dx, dr, not_r0 != 0));
  GUARD(col_to_host(c, dr, n_vc, res0));
  KS h[3];
  FVK_HIP(hipMemcpy(h, c.ks, sizeof h, hipMemcpyDeviceToHost));
  for (int q = 0; q < n_vc; ++q) res0_n[q] = h[q].res0_n;
  return FVK_OK;
}

int fvk_solv_res_dev(fvk_ctx *cc, int n_vc, const double *phi_dev, double *res0_dev, double *res0_n_dev,
                     int not_r0) {
  if (!cc || !res0_dev || !res0_n_dev || (!not_r0 && !phi_dev)) { set_error("null argument"); return FVK_ERR_ARG; }
  if (n_vc < 1 || n_vc > 3) { set_error("n_vc must be 1, 2 or 3 (the scratch holds 3 components)"); return FVK_ERR_ARG; }
  Ctx &c = cc->c;
  double *part = (c.device >= 0 && c.device < kMaxDevices) ? g_scratch[c.device] : nullptr;
  if (!part) { set_error("scratch of this device not initialised"); return FVK_ERR_STATE; }
  FVK_HIP(hipSetDevice(c.device));
  return residual_to(c, n_vc, phi_dev, res0_dev, not_r0 != 0, part, res0_n_dev);
}

static int solve_common(fvk_ctx *cc, int kind, int imat, double reduc, int n_iter, int n_vc, double *phi,
                        double *res0, double *res0_n, int not_r0, fvk_stats *st) {
  if (!cc || !phi || !res0 || !res0_n || n_vc < 1 || n_vc > 3 || imat < 0 || imat > 2 || n_iter < 0) {
    set_error("bad argument");
    return FVK_ERR_ARG;
  }
  Ctx &c = cc->c;
  double *dx = wk_slot(c, 5), *dr = wk_slot(c, 6);
  GUARD(to_dev(c, phi, n_vc, n_vc, c.n, &c.perm, dx));
  GUARD(col_to_dev(c, res0, n_vc, dr));
  FVK_HIP(hipEventRecord(c.ev0, c.stream));
  GUARD(krylov(c, kind, imat, reduc, n_iter, n_vc, dx, dr, not_r0 != 0));
  FVK_HIP(hipEventRecord(c.ev1, c.stream));
  GUARD(to_host(c, dx, n_vc, n_vc, c.n, &c.perm, phi));
  GUARD(col_to_host(c, dr, n_vc, res0));
  KS h[3];
  FVK_HIP(hipMemcpy(h, c.ks, sizeof h, hipMemcpyDeviceToHost));
  for (int q = 0; q < n_vc; ++q) res0_n[q] = h[q].res0_n;
  if (st) {
    std::memset(st, 0, sizeof *st);
    for (int q = 0; q < n_vc; ++q) { st->iters[q] = h[q].iter; st->res_rel[q] = h[q].res_n; }
    float ms = 0;
    hipEventElapsedTime(&ms, c.ev0, c.ev1);
    st->seconds = ms * 1e-3;
  }
  return FVK_OK;
}

int fvk_solv_cg(fvk_ctx *c, int imat, double reduc, int n_iter, int n_vc, double *phi, double *res0, double *res0_n,
                int not_r0, fvk_stats *st) {
  const int rc = solve_common(c, 0, imat, reduc, n_iter, n_vc, phi, res0, res0_n, not_r0, st);
  return rc ? rc : sweep_status(c->c);
}
int fvk_solv_bicg(fvk_ctx *c, int imat, double reduc, int n_iter, int n_vc, double *phi, double *res0,
                  double *res0_n, int not_r0, fvk_stats *st) {
  const int rc = solve_common(c, 1, imat, reduc, n_iter, n_vc, phi, res0, res0_n, not_r0, st);
  return rc ? rc : sweep_status(c->c);
}

int fvk_solv_grad(fvk_ctx *cc, int n_vc, const double *phi, double *gout, int n_iter, int neu) {
  if (!cc || !phi || !gout || n_vc < 1 || n_vc > 3 || n_iter < 1) { set_error("bad argument"); return FVK_ERR_ARG; }
  Ctx &c = cc->c;
  const int n = c.n, nb = c.nbf, nd = c.ndx;
  if (!c.bnd_buf) GUARD(dalloc(c, &c.bnd_buf, (size_t)3 * std::max(nb, 1)));
  // g and g0 hold 3 n_vc x n doubles: three slots each
  double *dphi = wk_slot(c, 0), *dphib = c.bnd_buf, *g = wk_slot(c, 2), *g0 = wk_slot(c, 5);
  GUARD(to_dev(c, phi, n_vc, n_vc, n, &c.perm, dphi));
  if (nb) GUARD(to_dev(c, phi + (size_t)n * n_vc, n_vc, n_vc, nb, nullptr, dphib));
  GUARD(grad(c, n_vc, dphi, dphib, g, n_iter, neu, nullptr, g0));
  std::vector<double> tmp((size_t)9 * n);
  FVK_HIP(hipStreamSynchronize(c.stream));
  FVK_HIP(hipMemcpy(tmp.data(), g, sizeof(double) * 3 * n_vc * n, hipMemcpyDeviceToHost));
  for (int r = 0; r < n; ++r) {
    const int o = c.perm[r];
    for (int q = 0; q < n_vc; ++q)
      for (int d = 0; d < nd; ++d) gout[((size_t)o * n_vc + q) * nd + d] = tmp[(size_t)(q * 3 + d) * n + r];
  }
  return FVK_OK;
}

// ---------------------------------------------------------------- systems
static int sys_alloc(Ctx &c) {
  if (c.sys_b) return 0;
  GUARD(dalloc(c, &c.sys_b, c.n));
  GUARD(dalloc(c, &c.sys_x, c.n));
  GUARD(dalloc(c, &c.sys_r, (size_t)c.n + 8));
  return 0;
}

static uint64_t sm64(uint64_t x) {
  x += 0x9E3779B97F4A7C15ULL;
  x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ULL;
  x = (x ^ (x >> 27)) * 0x94D049BB133111EBULL;
  return x ^ (x >> 31);
}
// standard normal from two counter-based uniforms (Box-Muller)
static double normal_at(uint64_t seed, uint64_t k) {
  const double u1 = ((double)(sm64(seed ^ (2 * k)) >> 11) + 0.5) * 0x1.0p-53;
  const double u2 = (double)(sm64(seed ^ (2 * k + 1)) >> 11) * 0x1.0p-53;
  return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
}

int fvk_sys_set(fvk_ctx *cc, const double *pl, const double *a, const double *b) {
  if (!cc || !pl || !a || !b) { set_error("null argument"); return FVK_ERR_ARG; }
  Ctx &c = cc->c;
  GUARD(sys_alloc(c));
  std::vector<double> ng(c.nf);
  for (int f = 0; f < c.nf; ++f) ng[f] = -a[f];
  GUARD(fvk_set_matrix(cc, pl, ng.data(), ng.data()));
  GUARD(to_dev(c, b, 1, 1, c.n, &c.perm, c.sys_b));
  return FVK_OK;
}

int fvk_sys_random(fvk_ctx *cc, uint64_t seed, double sigma, double *b_out) {
  if (!cc) { set_error("null argument"); return FVK_ERR_ARG; }
  Ctx &c = cc->c;
  // geometry factor s/l_pn per face, on the host
  std::vector<double> s(c.nf), l(c.nf);
  FVK_HIP(hipMemcpy(s.data(), c.M.fs, sizeof(double) * c.nf, hipMemcpyDeviceToHost));
  FVK_HIP(hipMemcpy(l.data(), c.M.flpn, sizeof(double) * c.nf, hipMemcpyDeviceToHost));
  std::vector<double> af(c.nf, 0.0), pl(c.n, 0.0), b(c.n);
  std::vector<int> orig_ip(c.nf), orig_in(c.nf);
  for (int f = 0; f < c.nf; ++f) {
    if (c.h_fin[f] < 0) continue;
    const double a = s[f] / l[f] * std::exp(sigma * normal_at(seed, (uint64_t)f));
    af[f] = a;
    pl[c.perm[c.h_fip[f]]] += a;
    pl[c.perm[c.h_fin[f]]] += a;
  }
  double mean = 0.0;
  for (int i = 0; i < c.n; ++i) { b[i] = normal_at(seed ^ 0x5DEECE66DULL, (uint64_t)i + 0x100000000ULL); mean += b[i]; }
  mean /= c.n;
  for (int i = 0; i < c.n; ++i) b[i] -= mean;
  if (b_out) std::memcpy(b_out, b.data(), sizeof(double) * c.n);
  return fvk_sys_set(cc, pl.data(), af.data(), b.data());
}

int fvk_sys_get_x(fvk_ctx *cc, double *x) {
  if (!cc || !x || !cc->c.sys_x) { set_error("no system"); return FVK_ERR_STATE; }
  return to_host(cc->c, cc->c.sys_x, 1, 1, cc->c.n, &cc->c.perm, x);
}

int fvk_sys_get(fvk_ctx *cc, double *pl, double *a, double *b) {
  if (!cc || !cc->c.sys_b) { set_error("no system"); return FVK_ERR_STATE; }
  Ctx &c = cc->c;
  if (pl) GUARD(to_host(c, c.diag, 1, 1, c.n, &c.perm, pl));
  if (a) {
    FVK_HIP(hipStreamSynchronize(c.stream));
    FVK_HIP(hipMemcpy(a, c.ngp, sizeof(double) * c.nf, hipMemcpyDeviceToHost));
    for (int f = 0; f < c.nf; ++f) a[f] = -a[f];
  }
  if (b) GUARD(to_host(c, c.sys_b, 1, 1, c.n, &c.perm, b));
  return FVK_OK;
}

static long nnz_offdiag(const Ctx &c) {
  long k = 0;
  for (int f = 0; f < c.nf; ++f) k += c.h_fin[f] >= 0 ? 2 : 0;
  return k;
}

static int solve_pcg64(Ctx &c, int imat, double tol, int max_iter, KTimer *T, KS &h) {
  return pcg_tol(c, imat, c.sys_b, c.sys_x, tol, max_iter, 64, T, h);
}

int fvk_sys_solve(fvk_ctx *cc, int method, int imat, double tol, int max_iter, int time_kernel,
                  fvk_solve_info *info) {
  if (!cc || !info || imat < 0 || imat > 2 || max_iter < 1 || method < FVK_SOLVE_PCG64 ||
      method > FVK_SOLVE_DC32_FACTOR || (method != FVK_SOLVE_PCG64 && imat != 0)) {
    set_error("bad argument");
    return FVK_ERR_ARG;
  }
  Ctx &c = cc->c;
  if (!c.sys_b) { set_error("no system loaded"); return FVK_ERR_STATE; }
  if (method == FVK_SOLVE_DC32_FACTOR && !c.mx.fac_valid) {
    set_error("no preconditioner factor: call fvk_factor_set or fvk_factor_build first");
    return FVK_ERR_STATE;
  }
  std::memset(info, 0, sizeof *info);
  const int n = c.n;
  KTimer T;
  if (time_kernel) T.init(64);
  FVK_HIP(hipEventRecord(c.ev0, c.stream));
  int status = FVK_OK;
  KS h;
  std::memset(&h, 0, sizeof h);
  if (method == FVK_SOLVE_PCG64) {
    GUARD(solve_pcg64(c, imat, tol, max_iter, time_kernel ? &T : nullptr, h));
    info->iters = h.iter;
    info->res_rec = h.res_n;
    status = h.done && h.res_n < tol ? FVK_OK : FVK_NOT_CONVERGED;
  } else {
    DcResult R;
    GUARD(dc_solve(c, method == FVK_SOLVE_DC32_FACTOR, c.sys_b, c.sys_x, tol, max_iter, R,
                   time_kernel ? &T : nullptr));
    info->iters = R.iters;
    info->res_rec = R.res_l1;
    info->outer = R.outer;
    info->breakdowns = R.breakdowns;
    for (int k = 0; k < 8; ++k) {
      info->inner_iters[k] = R.inner[k];
      info->outer_res[k] = R.res_before[k];
    }
    status = R.converged ? FVK_OK : FVK_NOT_CONVERGED;
    if (method == FVK_SOLVE_DC32_FACTOR && !R.converged) {
      // the plain FP64 Jacobi PCG takes over from x = 0 with a budget of
      // its own (max_iter: the learned solve was allowed what the plain
      // solver is allowed)
      KS hp;
      GUARD(solve_pcg64(c, 0, tol, max_iter, nullptr, hp));
      info->fallback = 1;
      info->iters += hp.iter;
      info->res_rec = hp.res_n;
      status = hp.done && hp.res_n < tol ? FVK_FALLBACK : FVK_NOT_CONVERGED;
    }
  }
  FVK_HIP(hipEventRecord(c.ev1, c.stream));
  FVK_HIP(hipEventSynchronize(c.ev1));
  float ms = 0.f;
  FVK_HIP(hipEventElapsedTime(&ms, c.ev0, c.ev1));
  double tr[4];
  GUARD(true_residual(c, c.sys_x, c.sys_b, c.sys_r, tr));
  info->seconds = ms * 1e-3;
  info->res_true_l1 = tr[0] / tr[2];
  info->res_true_l2 = std::sqrt(tr[1] / tr[3]);
  if (time_kernel && T.launches) {
    const double nnz_off = (double)nnz_offdiag(c);
    info->kern_seconds = T.total_ms * 1e-3 / T.launches;
    if (method == FVK_SOLVE_PCG64)
      // k_cg_pq: z, p_old, diag read; p_new, q written; column + value per off-diagonal
      info->kern_bytes = 40.0 * n + 12.0 * nnz_off;
    else if (method == FVK_SOLVE_DC32)
      // k_mx_pq on the Jacobi-scaled FP32 system (unit diagonal): z, p_old
      // read; p_new, q written; 16/32-bit column + FP32 value per off-diagonal
      info->kern_bytes = 16.0 * n + (c.mx.i16 ? 6.0 : 8.0) * nnz_off;
    else if (c.mx.qfull)
      // k_mx_qf (factor path, full rows): p read, q written; a column and
      // an FP32 value per off-diagonal
      info->kern_bytes = 8.0 * n + (c.mx.i16 ? 6.0 : 8.0) * nnz_off;
    else
      // k_mx_q (factor path): p read, q written; per upper off-diagonal a
      // column and an FP32 value, per lower one a column and a one-byte slot
      // index (the value is the upper entry of the neighbour, read from cache)
      info->kern_bytes = 8.0 * n + (c.mx.i16 ? 4.5 : 6.5) * nnz_off;
    info->kern_launches = T.launches;
  }
  return status;
}

int fvk_ctx_set_param(fvk_ctx *cc, const char *name, double value) {
  if (!cc || !name) { set_error("null argument"); return FVK_ERR_ARG; }
  Mix &X = cc->c.mx;
  if (!std::strcmp(name, "dc_inner_reduc") && value > 0 && value < 1) X.inner_reduc = value;
  else if (!std::strcmp(name, "dc_max_outer") && value >= 1 && value <= 1000) X.max_outer = (int)value;
  else if (!std::strcmp(name, "dc_carry") && (value == 0 || value == 1)) X.carry = value != 0;
  else if (!std::strcmp(name, "dc_max_breakdowns") && value >= 0 && value <= 100 && value == (int)value)
    X.max_breakdowns = (int)value;
  else if (!std::strcmp(name, "dc_eta_margin") && value > 0 && value <= 1) X.eta_margin = value;
  else if (!std::strcmp(name, "dc_replace_every") && value >= 0 && value <= 1e6 && value == (int)value)
    X.replace_every = (int)value;
  else if (!std::strcmp(name, "sweep_one_launch") && (value == 0 || value == 1)) {
    const bool was = cc->c.sweep_one_launch;
    cc->c.sweep_one_launch = value != 0 && cc->c.Ws <= 8;
    if (cc->c.sweep_one_launch && !was) {
      const int rc = sweep_parts_build(cc->c);
      if (rc) return rc;
    }
  } else if (!std::strcmp(name, "sweep_parts") && (value == 0 || value == 1))
    cc->c.sweep_parts = value != 0;
  else if (!std::strcmp(name, "sweep_part_rows") && value >= 256 && value <= 6400 && value == (int)value) {
    cc->c.sweep_part_rows = (int)value;
    const int rc = sweep_parts_build(cc->c);
    if (rc) return rc;
  }
  else if (!std::strcmp(name, "sweep_blocks_per_cu") && value >= 1 && value <= 8 && value == (int)value)
    cc->c.sweep_blocks_per_cu = (int)value;
  else if (!std::strcmp(name, "sweep_fail") && value == 1 && cc->c.sweep_int) {
    const int one = 1;
    FVK_HIP(hipMemcpyAsync(cc->c.sweep_int + 3, &one, sizeof one, hipMemcpyHostToDevice, cc->c.stream));
    FVK_HIP(hipStreamSynchronize(cc->c.stream));
  }
  else { set_error(std::string("unknown parameter or value out of range: ") + name); return FVK_ERR_ARG; }
  return FVK_OK;
}

int fvk_factor_set(fvk_ctx *cc, const float *l, const float *sig) {
  if (!cc || !l || !sig) { set_error("null argument"); return FVK_ERR_ARG; }
  Ctx &c = cc->c;
  GUARD(mix_build(c));
  std::vector<float> fl, fs;
  fl.reserve(c.mx.nfi);
  fs.reserve(c.mx.nfi);
  for (int f = 0; f < c.nf; ++f)
    if (c.h_fin[f] >= 0) { fl.push_back(l[f]); fs.push_back(sig[f]); }
  FVK_HIP(hipStreamSynchronize(c.stream));
  if (!fl.empty()) {
    FVK_HIP(hipMemcpy(c.mx.fl, fl.data(), fl.size() * sizeof(float), hipMemcpyHostToDevice));
    FVK_HIP(hipMemcpy(c.mx.fsig, fs.data(), fs.size() * sizeof(float), hipMemcpyHostToDevice));
  }
  return factor_assemble(c);
}

int fvk_factor_apply(fvk_ctx *cc, const double *r, double *z) {
  if (!cc || !r || !z) { set_error("null argument"); return FVK_ERR_ARG; }
  Ctx &c = cc->c;
  GUARD(mix_build(c));
  double *dr = c.stage, *dz = c.stage + c.n;
  GUARD(to_dev(c, r, 1, 1, c.n, &c.perm, dr));
  GUARD(factor_apply(c, dr, dz));
  return to_host(c, dz, 1, 1, c.n, &c.perm, z);
}

int fvk_bench_spmv(fvk_ctx *cc, int reps, double *sec) {
  if (!cc || !sec || reps < 1) { set_error("bad argument"); return FVK_ERR_ARG; }
  Ctx &c = cc->c;
  double *x = wk_slot(c, 0), *y = wk_slot(c, 1);
  for (int k = 0; k < 3; ++k) GUARD(spmv(c, 1, x, y));
  FVK_HIP(hipEventRecord(c.ev0, c.stream));
  for (int k = 0; k < reps; ++k) GUARD(spmv(c, 1, x, y));
  FVK_HIP(hipEventRecord(c.ev1, c.stream));
  FVK_HIP(hipEventSynchronize(c.ev1));
  float ms = 0.f;
  FVK_HIP(hipEventElapsedTime(&ms, c.ev0, c.ev1));
  *sec = ms * 1e-3 / reps;
  return FVK_OK;
}

int fvk_bench_sweep(fvk_ctx *cc, int imat, int nvc, int reps, double *sec) {
  if (!cc || !sec || reps < 1 || imat < 1 || imat > 2 || nvc < 1 || nvc > 3) {
    set_error("bad argument");
    return FVK_ERR_ARG;
  }
  Ctx &c = cc->c;
  double *b = wk_slot(c, 0), *x = wk_slot(c, 1);
  GUARD(precond_build(c, imat));
  for (int k = 0; k < 2; ++k) GUARD(precond_apply(c, imat, nvc, b, x, nullptr));
  FVK_HIP(hipEventRecord(c.ev0, c.stream));
  for (int k = 0; k < reps; ++k) GUARD(precond_apply(c, imat, nvc, b, x, nullptr));
  FVK_HIP(hipEventRecord(c.ev1, c.stream));
  FVK_HIP(hipEventSynchronize(c.ev1));
  float ms = 0.f;
  FVK_HIP(hipEventElapsedTime(&ms, c.ev0, c.ev1));
  *sec = ms * 1e-3 / reps;
  return sweep_status(c);
}

__global__ __launch_bounds__(BS) void k_copy(long n4, const double4 *__restrict__ a, double4 *__restrict__ b) {
  for (long i = blockIdx.x * (long)BS + threadIdx.x; i < n4; i += (long)gridDim.x * BS) b[i] = a[i];
}

int fvk_bench_copy(int device, long nbytes, int reps, double *sec) {
  if (!sec || reps < 1 || nbytes < 32) { set_error("bad argument"); return FVK_ERR_ARG; }
  FVK_HIP(hipSetDevice(device));
  const long n4 = nbytes / 32;
  void *a = nullptr, *b = nullptr;
  FVK_HIP(hipMalloc(&a, n4 * 32));
  if (hipMalloc(&b, n4 * 32) != hipSuccess) { hipFree(a); set_error("hipMalloc failed"); return FVK_ERR_HIP; }
  hipMemset(a, 0, n4 * 32);
  hipEvent_t e0, e1;
  hipEventCreate(&e0);
  hipEventCreate(&e1);
  const int G = 256 * 8 * 4;
  for (int k = 0; k < 3; ++k) k_copy<<<G, BS>>>(n4, (const double4 *)a, (double4 *)b);
  hipEventRecord(e0, 0);
  for (int k = 0; k < reps; ++k) k_copy<<<G, BS>>>(n4, (const double4 *)a, (double4 *)b);
  hipEventRecord(e1, 0);
  hipEventSynchronize(e1);
  float ms = 0.f;
  hipEventElapsedTime(&ms, e0, e1);
  hipError_t err = hipGetLastError();
  hipEventDestroy(e0);
  hipEventDestroy(e1);
  hipFree(a);
  hipFree(b);
  if (err != hipSuccess) { set_error(hipGetErrorString(err)); return FVK_ERR_HIP; }
  *sec = ms * 1e-3 / reps;
  return FVK_OK;
}

// Streaming probes of the bandwidth ceiling: every thread keeps U
// independent 16-byte loads in flight (grid-stride over float4 blocks of U).
constexpr int kStreamU = 4;
__global__ __launch_bounds__(BS) void k_stream_copy(long n16, const float4 *__restrict__ a, float4 *__restrict__ b) {
  constexpr int U = kStreamU;
  const long stride = (long)gridDim.x * BS;
  for (long i = blockIdx.x * (long)BS + threadIdx.x; i < n16; i += stride * U) {
    float4 v[U];
#pragma unroll
    for (int u = 0; u < U; ++u) if (i + u * stride < n16) v[u] = a[i + u * stride];
#pragma unroll
    for (int u = 0; u < U; ++u) if (i + u * stride < n16) b[i + u * stride] = v[u];
  }
}

__global__ __launch_bounds__(BS) void k_stream_read(long n16, const float4 *__restrict__ a, float *__restrict__ sink) {
  constexpr int U = kStreamU;
  const long stride = (long)gridDim.x * BS;
  float acc = 0.f;
  for (long i = blockIdx.x * (long)BS + threadIdx.x; i < n16; i += stride * U) {
    float4 v[U];
#pragma unroll
    for (int u = 0; u < U; ++u) v[u] = i + u * stride < n16 ? a[i + u * stride] : make_float4(0.f, 0.f, 0.f, 0.f);
#pragma unroll
    for (int u = 0; u < U; ++u) acc += v[u].x + v[u].y + v[u].z + v[u].w;
  }
  if (acc == 1234.5f) sink[0] = acc;  // keeps the loads; never true for the zero-filled input
}

__global__ __launch_bounds__(BS) void k_stream_fill(long n16, float4 *__restrict__ b) {
  for (long i = blockIdx.x * (long)BS + threadIdx.x; i < n16; i += (long)gridDim.x * BS)
    b[i] = make_float4(0.f, 0.f, 0.f, 0.f);
}

// Eight 16-byte non-temporal loads in flight per thread, one grid of 8
// blocks per CU (full occupancy, no tail wave).
typedef float nt4 __attribute__((ext_vector_type(4)));
__global__ __launch_bounds__(BS) void k_stream_copy_nt(long n16, const nt4 *__restrict__ a, nt4 *__restrict__ b) {
  constexpr int U = 8;
  const long stride = (long)gridDim.x * BS;
  for (long i = blockIdx.x * (long)BS + threadIdx.x; i < n16; i += stride * U) {
    nt4 v[U];
#pragma unroll
    for (int u = 0; u < U; ++u)
      if (i + u * stride < n16) v[u] = __builtin_nontemporal_load(a + i + u * stride);
#pragma unroll
    for (int u = 0; u < U; ++u)
      if (i + u * stride < n16) __builtin_nontemporal_store(v[u], b + i + u * stride);
  }
}

__global__ __launch_bounds__(BS) void k_stream_read_nt(long n16, const nt4 *__restrict__ a, float *__restrict__ sink) {
  constexpr int U = 8;
  const long stride = (long)gridDim.x * BS;
  float acc = 0.f;
  for (long i = blockIdx.x * (long)BS + threadIdx.x; i < n16; i += stride * U) {
    nt4 v[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      v[u] = nt4{0.f, 0.f, 0.f, 0.f};
      if (i + u * stride < n16) v[u] = __builtin_nontemporal_load(a + i + u * stride);
    }
#pragma unroll
    for (int u = 0; u < U; ++u) acc += v[u].x + v[u].y + v[u].z + v[u].w;
  }
  if (acc == 1234.5f) sink[0] = acc;  // keeps the loads; never true for the zero-filled input
}

int fvk_bench_stream(int device, long nbytes, int reps, int kind, double *sec) {
  if (!sec || reps < 1 || nbytes < 4096 || kind < 0 || kind > 4) { set_error("bad argument"); return FVK_ERR_ARG; }
  FVK_HIP(hipSetDevice(device));
  const long n16 = nbytes / 16;
  void *a = nullptr, *b = nullptr;
  FVK_HIP(hipMalloc(&a, n16 * 16));
  if (hipMalloc(&b, n16 * 16) != hipSuccess) { hipFree(a); set_error("hipMalloc failed"); return FVK_ERR_HIP; }
  hipMemset(a, 0, n16 * 16);
  hipEvent_t e0, e1;
  hipEventCreate(&e0);
  hipEventCreate(&e1);
  int ncu = 256;
  hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, device);
  const int G = 256 * 16, Gnt = ncu * 8;
  auto launch = [&]() {
    if (kind == 0) k_stream_copy<<<G, BS>>>(n16, (const float4 *)a, (float4 *)b);
    else if (kind == 1) k_stream_read<<<G, BS>>>(n16, (const float4 *)a, (float *)b);
    else if (kind == 2) k_stream_fill<<<G, BS>>>(n16, (float4 *)b);
    else if (kind == 3) k_stream_copy_nt<<<Gnt, BS>>>(n16, (const nt4 *)a, (nt4 *)b);
    else k_stream_read_nt<<<Gnt, BS>>>(n16, (const nt4 *)a, (float *)b);
  };
  for (int k = 0; k < 3; ++k) launch();
  hipEventRecord(e0, 0);
  for (int k = 0; k < reps; ++k) launch();
  hipEventRecord(e1, 0);
  hipEventSynchronize(e1);
  float ms = 0.f;
  hipEventElapsedTime(&ms, e0, e1);
  hipError_t err = hipGetLastError();
  hipEventDestroy(e0);
  hipEventDestroy(e1);
  hipFree(a);
  hipFree(b);
  if (err != hipSuccess) { set_error(hipGetErrorString(err)); return FVK_ERR_HIP; }
  *sec = ms * 1e-3 / reps;
  return FVK_OK;
}

// ---------------------------------------------------------------- flow
int fvk_flow_create(fvk_ctx *cc, const fvk_flow_cfg *cfg, fvk_flow **out) {
  if (!cc || !cfg || !out) { set_error("null argument"); return FVK_ERR_ARG; }
  if (cfg->imat_u < 0 || cfg->imat_u > 2 || cfg->imat_p < 0 || cfg->imat_p > 2 || cfg->n_grad < 1 ||
      cfg->urf_u <= 0.0 || cfg->n_bicg < 0 || cfg->n_cg < 0 || cfg->p_solver < 0 || cfg->p_solver > 3 ||
      (cfg->p_solver > 0 && !(cfg->p_tol > 0.0 && cfg->p_tol < 1.0))) {
    set_error("bad flow configuration");
    return FVK_ERR_ARG;
  }
  Ctx &c = cc->c;
  fvk_flow *ff = new fvk_flow;
  Flow &F = ff->f;
  F.c = &c;
  F.cfg = *cfg;
  F.nvc = c.ndx;
  const size_t n = c.n, nf = c.nf, nb = std::max(c.nbf, 1);
  int rc = 0;
  auto A = [&](double **p, size_t cnt) { if (!rc) rc = dalloc(c, p, cnt); };
  A(&F.u, 3 * n); A(&F.ub, 3 * nb); A(&F.p, n); A(&F.dpdx, 3 * n); A(&F.spd, nf); A(&F.frc_dns, nf);
  A(&F.cs, n); A(&F.frc_vol, n); A(&F.ddx, 9 * n); A(&F.g0, 9 * n); A(&F.rhs, 3 * n); A(&F.rhs_p, n);
  A(&F.cor, 3 * nf); A(&F.cfl, nf); A(&F.cpp, nf); A(&F.cinout, nf); A(&F.pl_r, n); A(&F.pc, n);
  A(&F.ddxp, 3 * n); A(&F.red_scratch, 8);
  if (!rc) rc = dalloc(c, &F.sc, 1);
  // inlet / outlet face lists and outlet cells
  std::vector<int> fin_l, fout_l;
  for (int f = 0; f < c.nf; ++f) {
    if (c.h_fin[f] >= 0) continue;
    int cls = c.h_fbty[f] - c.h_fbty[f] % 10;
    if (cls == 10) fin_l.push_back(f);
    if (cls == 20) fout_l.push_back(f);
  }
  F.n_in = (int)fin_l.size();
  F.n_out = (int)fout_l.size();
  std::vector<int> io(fin_l);
  io.insert(io.end(), fout_l.begin(), fout_l.end());
  if (!rc && !io.empty()) rc = upload(c, &F.io_faces, io);
  if (!rc && !fout_l.empty()) {
    std::vector<int> cell_oc(c.n, -1), oc_cell, oc_face;
    for (int f : fout_l) {  // ascending face order
      int r = c.h_fip[f];
      if (cell_oc[r] < 0) {
        cell_oc[r] = (int)oc_cell.size();
        oc_cell.push_back(r);
        oc_face.insert(oc_face.end(), {-1, -1, -1, -1});
      }
      int m = cell_oc[r], k = 0;
      while (k < 4 && oc_face[4 * m + k] >= 0) ++k;
      if (k == 4) { rc = FVK_ERR_ARG; set_error("cell with more than 4 outlet faces"); break; }
      oc_face[4 * m + k] = f;
    }
    F.n_oc = (int)oc_cell.size();
    if (!rc) rc = upload(c, &F.cell_oc, cell_oc);
    if (!rc) rc = upload(c, &F.oc_cell, oc_cell);
    if (!rc) rc = upload(c, &F.oc_face, oc_face);
    if (!rc) rc = dalloc(c, &F.oc_u, (size_t)F.n_oc * 12);
  }
  F.row_of_cell1 = c.iperm[0];
  if (!rc) {
    hipError_t e = hipStreamSynchronize(c.stream);
    if (e != hipSuccess) { set_error(hipGetErrorString(e)); rc = FVK_ERR_HIP; }
  }
  if (rc) {
    delete ff;
    return rc;
  }
  *out = ff;
  return FVK_OK;
}

void fvk_flow_destroy(fvk_flow *f) {
  if (!f) return;
  for (auto g : f->f.graphs) hipGraphExecDestroy(g);
  delete f;  // device memory is owned by the context
}

int fvk_flow_configure(fvk_flow *ff, const fvk_flow_cfg *cfg) {
  if (!ff || !cfg) { set_error("null argument"); return FVK_ERR_ARG; }
  if (cfg->imat_u < 0 || cfg->imat_u > 2 || cfg->imat_p < 0 || cfg->imat_p > 2 || cfg->n_grad < 1 ||
      cfg->urf_u <= 0.0 || cfg->n_bicg < 0 || cfg->n_cg < 0 || cfg->p_solver < 0 || cfg->p_solver > 3 ||
      (cfg->p_solver > 0 && !(cfg->p_tol > 0.0 && cfg->p_tol < 1.0))) {
    set_error("bad flow configuration");
    return FVK_ERR_ARG;
  }
  ff->f.cfg = *cfg;
  return FVK_OK;
}

int fvk_flow_scalars(fvk_flow *ff, int set, double *c_in, double *c_out) {
  if (!ff || !c_in || !c_out) { set_error("null argument"); return FVK_ERR_ARG; }
  Flow &F = ff->f;
  FVK_HIP(hipStreamSynchronize(F.c->stream));
  FlowSc s;
  FVK_HIP(hipMemcpy(&s, F.sc, sizeof s, hipMemcpyDeviceToHost));
  if (set) {
    s.c_in = *c_in;
    s.c_out = *c_out;
    FVK_HIP(hipMemcpy(F.sc, &s, sizeof s, hipMemcpyHostToDevice));
  } else {
    *c_in = s.c_in;
    *c_out = s.c_out;
  }
  return FVK_OK;
}

int fvk_flow_set_bnd_u(fvk_flow *ff, int code, const double *u) {
  if (!ff || !u) { set_error("null argument"); return FVK_ERR_ARG; }
  Flow &F = ff->f;
  Ctx &c = *F.c;
  const int nb = c.nbf, nvc = F.nvc;
  std::vector<double> tmp((size_t)3 * std::max(nb, 1));
  FVK_HIP(hipStreamSynchronize(c.stream));
  FVK_HIP(hipMemcpy(tmp.data(), F.ub, sizeof(double) * 3 * std::max(nb, 1), hipMemcpyDeviceToHost));
  for (int f = 0; f < c.nf; ++f)
    if (c.h_fin[f] < 0 && c.h_fbty[f] == code)
      for (int q = 0; q < nvc; ++q) tmp[(size_t)q * nb + c.h_fib[f]] = u[q];
  FVK_HIP(hipMemcpy(F.ub, tmp.data(), sizeof(double) * 3 * std::max(nb, 1), hipMemcpyHostToDevice));
  return FVK_OK;
}

int fvk_flow_set_state(fvk_flow *ff, const double *u_cell, const double *u_bnd, const double *p, const double *dpdx,
                       const double *spd) {
  if (!ff) return FVK_ERR_ARG;
  Flow &F = ff->f;
  Ctx &c = *F.c;
  FVK_HIP(hipStreamSynchronize(c.stream));
  if (u_cell) GUARD(to_dev(c, u_cell, F.nvc, F.nvc, c.n, &c.perm, F.u));
  if (u_bnd && c.nbf) GUARD(to_dev(c, u_bnd, F.nvc, F.nvc, c.nbf, nullptr, F.ub));
  if (p) GUARD(to_dev(c, p, 1, 1, c.n, &c.perm, F.p));
  if (dpdx) GUARD(to_dev(c, dpdx, c.ndx, c.ndx, c.n, &c.perm, F.dpdx));
  if (spd) FVK_HIP(hipMemcpy(F.spd, spd, sizeof(double) * c.nf, hipMemcpyHostToDevice));
  return FVK_OK;
}

int fvk_flow_get_state(fvk_flow *ff, double *u_cell, double *u_bnd, double *p, double *dpdx, double *spd, double *cs,
                       double *frc_dns) {
  if (!ff) return FVK_ERR_ARG;
  Flow &F = ff->f;
  Ctx &c = *F.c;
  if (u_cell) GUARD(to_host(c, F.u, F.nvc, F.nvc, c.n, &c.perm, u_cell));
  if (u_bnd && c.nbf) GUARD(to_host(c, F.ub, F.nvc, F.nvc, c.nbf, nullptr, u_bnd));
  if (p) GUARD(to_host(c, F.p, 1, 1, c.n, &c.perm, p));
  if (dpdx) GUARD(to_host(c, F.dpdx, c.ndx, c.ndx, c.n, &c.perm, dpdx));
  if (cs) GUARD(to_host(c, F.cs, 1, 1, c.n, &c.perm, cs));
  FVK_HIP(hipStreamSynchronize(c.stream));
  if (spd) FVK_HIP(hipMemcpy(spd, F.spd, sizeof(double) * c.nf, hipMemcpyDeviceToHost));
  if (frc_dns) FVK_HIP(hipMemcpy(frc_dns, F.frc_dns, sizeof(double) * c.nf, hipMemcpyDeviceToHost));
  return FVK_OK;
}

int fvk_flow_get_matrix(fvk_flow *ff, double *pl, double *ng_p, double *ng_n) {
  if (!ff) return FVK_ERR_ARG;
  Flow &F = ff->f;
  Ctx &c = *F.c;
  if (pl) GUARD(to_host(c, c.diag, 1, 1, c.n, &c.perm, pl));
  FVK_HIP(hipStreamSynchronize(c.stream));
  if (ng_p) FVK_HIP(hipMemcpy(ng_p, c.ngp, sizeof(double) * c.nf, hipMemcpyDeviceToHost));
  if (ng_n) FVK_HIP(hipMemcpy(ng_n, c.ngn, sizeof(double) * c.nf, hipMemcpyDeviceToHost));
  return FVK_OK;
}

int fvk_flow_get_rhs(fvk_flow *ff, double *rhs) {
  if (!ff || !rhs) { set_error("null argument"); return FVK_ERR_ARG; }
  Flow &F = ff->f;
  return to_host(*F.c, F.rhs, 1, 1, F.c->n, &F.c->perm, rhs);
}

static int read_hist(Flow &F, fvk_outer_hist *h) {
  FlowSc s;
  FVK_HIP(hipStreamSynchronize(F.c->stream));
  FVK_HIP(hipMemcpy(&s, F.sc, sizeof s, hipMemcpyDeviceToHost));
  if (h) {
    h->reduc = s.h_reduc; h->redpp = s.h_redpp; h->redpp_sum = s.h_redpp_sum;
    h->red_p = s.h_red_p; h->red_p_sum = s.h_red_p_sum; h->mx_res_n = s.mx_res_n;
    for (int q = 0; q < 3; ++q) h->it_u[q] = s.it_u[q];
    h->it_pp = s.it_pp; h->it_p = s.it_p;
    for (int q = 0; q < 3; ++q) h->res0_n_u[q] = s.res0_n_u[q];
    h->res0_n_pp = s.res0_n_pp;
    h->res0_n_p = s.res0_n_p;
  }
  if (s.diverged) { set_error("momentum residual exceeded large"); return FVK_ERR_DIVERGED; }
  return sweep_status(*F.c);
}

int fvk_flow_setup_eqn(fvk_flow *ff, int it, fvk_outer_hist *h) {
  if (!ff || it < 1) return FVK_ERR_ARG;
  bool skipped;
  GUARD(mom_setup(ff->f, it, skipped));
  return read_hist(ff->f, h);
}

int fvk_flow_setup_pp(fvk_flow *ff, int it, fvk_outer_hist *h) {
  if (!ff || it < 1) return FVK_ERR_ARG;
  GUARD(pp_setup(ff->f, it));
  return read_hist(ff->f, h);
}

int fvk_flow_outer(fvk_flow *ff, int it, fvk_outer_hist *h) {
  if (!ff || it < 1) return FVK_ERR_ARG;
  Flow &F = ff->f;
  GUARD(outer_begin(F));
  bool skipped;
  GUARD(mom_setup(F, it, skipped));
  GUARD(pp_setup(F, it));
  return read_hist(F, h);
}

int fvk_flow_set_model(fvk_flow *ff, fvk_model *m, int train_steps, double lr) {
  if (!ff || train_steps < 0 || (train_steps > 0 && !(lr > 0))) { set_error("bad argument"); return FVK_ERR_ARG; }
  ff->f.model = m;
  ff->f.train_steps = m ? train_steps : 0;
  ff->f.train_lr = lr;
  return FVK_OK;
}

int fvk_flow_run_graph(fvk_flow *ff, int it0, int n, fvk_outer_hist *h) {
  if (!ff || it0 < 1 || n < 0) { set_error("bad argument"); return FVK_ERR_ARG; }
  int replayed = 0;
  return flow_run_graph(ff->f, it0, n, h, replayed);
}

int fvk_flow_run(fvk_flow *ff, int it0, int n, fvk_outer_hist *h) {
  if (!ff || it0 < 1 || n < 0) return FVK_ERR_ARG;
  for (int k = 0; k < n; ++k) GUARD(fvk_flow_outer(ff, it0 + k, h ? h + k : nullptr));
  return FVK_OK;
}

}  // extern "C"

// ---------------------------------------------------------------- output

extern "C" {

int fvk_fmt_list_real(double x, char *buf, int len) {
  if (!buf || len < 1) { set_error("null argument"); return FVK_ERR_ARG; }
  const std::string t = fvk::fmt_list_real(x);
  if ((int)t.size() + 1 > len) { set_error("buffer too small"); return FVK_ERR_ARG; }
  std::memcpy(buf, t.c_str(), t.size() + 1);
  return FVK_OK;
}

static int tecplot_common(const fvk_mesh *mm, fvk_flow *ff, bool eqn, int igrd, int iph, int ieq, int ict,
                          int n_tl, const char *dir) {
  if (!mm || !ff || !dir) { set_error("null argument"); return FVK_ERR_ARG; }
  const fvk::Mesh &m = mm->m;
  Flow &F = ff->f;
  if (m.n_vol != F.c->n || m.n_dx != F.c->ndx) {
    set_error("mesh and flow differ in cells or dimension");
    return FVK_ERR_ARG;
  }
  const int nvc = eqn ? F.nvc : 1;
  std::vector<double> val((size_t)m.n_vol * nvc);
  if (eqn) GUARD(fvk_flow_get_state(ff, val.data(), nullptr, nullptr, nullptr, nullptr, nullptr, nullptr));
  else GUARD(fvk_flow_get_state(ff, nullptr, nullptr, val.data(), nullptr, nullptr, nullptr, nullptr));
  fvk::TecplotIn in;
  in.kind_eqn = eqn;
  in.igrd = igrd; in.iph = iph; in.ieq = ieq; in.ict = ict; in.n_tl = n_tl;
  in.n_dx = m.n_dx; in.n_vrt = m.n_vrt; in.n_vol = m.n_vol; in.n_sc = 1; in.n_vc = nvc;
  in.vx = m.vx.data(); in.v_vptr = m.v_vptr.data(); in.v_vrt = m.v_vrt.data(); in.val = val.data();
  std::string path;
  const std::string err = fvk::tecplot_write(in, dir, path);
  if (!err.empty()) { set_error(err); return FVK_ERR_IO; }
  return FVK_OK;
}

int fvk_eqn_outp(const fvk_mesh *m, fvk_flow *f, int igrd, int iph, int ieq, int ict, int n_tl, const char *dir) {
  return tecplot_common(m, f, true, igrd, iph, ieq, ict, n_tl, dir);
}

int fvk_prp_outp_p(const fvk_mesh *m, fvk_flow *f, int igrd, int iph, int ipr, int ict, int n_tl,
                   const char *dir) {
  return tecplot_common(m, f, false, igrd, iph, ipr, ict, n_tl, dir);
}

int fvk_write_gmsh_field(const char *fname, int n, const int *elm_idx, const double *phi) {
  if (!fname || n < 0 || (n > 0 && (!elm_idx || !phi))) { set_error("bad argument"); return FVK_ERR_ARG; }
  const std::string err = fvk::gmsh_write(fname, n, 1, 1, elm_idx, phi);
  if (!err.empty()) { set_error(err); return FVK_ERR_IO; }
  return FVK_OK;
}

int fvk_write_gmsh_vector_field(const char *fname, int n, int n_dim, const int *elm_idx, const double *phi) {
  if (!fname || n < 0 || n_dim < 1 || n_dim > 3 || (n > 0 && (!elm_idx || !phi))) {
    set_error("bad argument");
    return FVK_ERR_ARG;
  }
  const std::string err = fvk::gmsh_write(fname, n, 3, n_dim, elm_idx, phi);
  if (!err.empty()) { set_error(err); return FVK_ERR_IO; }
  return FVK_OK;
}

}  // extern "C"
