"""Python host mirror of the spray-simulation solver interface over libfvk.so.

The functions keep the reference's names and argument meaning
(spray-simulation/solv_*.f90, cfd_setup.f90, grd_solv.f90) on numpy arrays:

    ctx = grd_solv(mesh)                          # grd_solv.f90:7-24
    solv_precond(ctx, imat)  -> (c, c_r)          # solv_precond.f90:5-30
    solv_direct(ctx, imat, b) -> x                # solv_direct.f90:5-43
    solv_res(ctx, phi, res0, not_r0) -> res0_n    # solv_res.f90:5-30
    solv_cg(ctx, imat, reduc, n_iter, phi, res0, not_r0)    # solv_cg.f90:7-87
    solv_bicg(ctx, imat, reduc, n_iter, phi, res0, not_r0)  # solv_bicg.f90:7-121
    solv_grad(ctx, phi, n_iter, neu) -> grad      # cfd_setup.f90:765-866
    Flow(ctx, cfg).cfd_setup_eqn / cfd_setup_pp / outer      # cfd_setup.f90, cfd_solve.f90

phi(i)%vc is row i of a (n_vol, n_vc) array; res0 is (n_vol, n_vc) and is
overwritten with the residual like the reference's intent(inout) argument.
Errors raise FvkError carrying the library's status and message.  There is
no fallback: without libfvk.so (or without a GPU for device calls) the calls
fail.
"""
import atexit
import ctypes as C
import os
import sys
import weakref

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libfvk.so")

FVK_OK, FVK_FALLBACK, FVK_NOT_CONVERGED = 0, 1, 2
FVK_ERR_ARG, FVK_ERR_HIP, FVK_ERR_STATE, FVK_ERR_DIVERGED, FVK_ERR_IO, FVK_ERR_NODEV = -1, -2, -3, -4, -5, -6
CTX_RCM = 1

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)
_fp = C.POINTER(C.c_float)
_vp = C.c_void_p


class FvkError(RuntimeError):
    def __init__(self, status, msg):
        super().__init__("fvk status %d: %s" % (status, msg))
        self.status = status


class MeshInfo(C.Structure):
    _fields_ = [(k, C.c_int) for k in ("n_dx", "n_vrt", "n_vol", "n_fce", "n_bf", "n_slot", "max_fce")]


class Stats(C.Structure):
    _fields_ = [("iters", C.c_int * 3), ("res_rel", C.c_double * 3), ("fallback", C.c_int),
                ("seconds", C.c_double)]


class SolveInfo(C.Structure):
    _fields_ = [("iters", C.c_int), ("fallback", C.c_int), ("seconds", C.c_double), ("res_rec", C.c_double),
                ("res_true_l1", C.c_double), ("res_true_l2", C.c_double), ("kern_seconds", C.c_double),
                ("kern_bytes", C.c_double), ("kern_launches", C.c_long), ("outer", C.c_int),
                ("inner_iters", C.c_int * 8), ("outer_res", C.c_double * 8), ("breakdowns", C.c_int)]


SOLVE_PCG64 = 0         # reference PCG recurrence, FP64
SOLVE_DC32 = 1          # FP32-inner Jacobi PCG in FP64 defect correction
SOLVE_DC32_FACTOR = 2   # same with the factorised preconditioner G G^T


class FlowCfg(C.Structure):
    _fields_ = [("simpler", C.c_int), ("imat_u", C.c_int), ("imat_p", C.c_int), ("n_bicg", C.c_int),
                ("n_cg", C.c_int), ("n_grad", C.c_int), ("c_sch", C.c_int),
                ("urf_u", C.c_double), ("urf_p", C.c_double), ("reduc_u", C.c_double),
                ("reduc_p", C.c_double), ("c_bl", C.c_double), ("rho", C.c_double), ("mu", C.c_double),
                ("p_solver", C.c_int), ("p_tol", C.c_double)]


class OuterHist(C.Structure):
    _fields_ = [("reduc", C.c_double), ("redpp", C.c_double), ("redpp_sum", C.c_double),
                ("red_p", C.c_double), ("red_p_sum", C.c_double), ("mx_res_n", C.c_double),
                ("it_u", C.c_int * 3), ("it_pp", C.c_int), ("it_p", C.c_int),
                ("res0_n_u", C.c_double * 3), ("res0_n_pp", C.c_double), ("res0_n_p", C.c_double)]


_lib = None

# Live handles by kind.  At interpreter exit they are destroyed by an atexit
# hook, dependants first (flows, models, contexts, meshes), while the HIP
# runtime is still up; finalizers that run later (module teardown, garbage
# collection during shutdown) make no library call at all.
_LIVE = {k: weakref.WeakSet() for k in ("flow", "model", "ctx", "mesh")}


def _track(kind, obj):
    _LIVE[kind].add(obj)


def _release_all():
    for kind in ("flow", "model", "ctx", "mesh"):
        for obj in list(_LIVE[kind]):
            try:
                obj._release()
            except Exception:
                pass


def _finalizing():
    return sys.is_finalizing()


def lib():
    """Load libfvk.so (built in place by programs_amd/build.py)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise FvkError(-6, "libfvk.so is not built (run python -m programs_amd.build)")
    L = C.CDLL(LIB_PATH)
    L.fvk_last_error.restype = C.c_char_p
    sig = {
        "fvk_mesh_read_neu": [C.c_char_p, C.c_double, C.POINTER(_vp)],
        "fvk_mesh_box": [C.c_int, C.c_int, C.c_int, C.c_double, C.c_double, C.c_double, C.c_double, C.c_uint64,
                         C.c_int, C.c_int, C.POINTER(_vp)],
        "fvk_mesh_from_cells": [C.c_int, C.c_int, _dp, C.c_int, _ip, _ip, _ip, _ip, C.POINTER(_vp)],
        "fvk_mesh_get_info": [_vp, C.POINTER(MeshInfo)],
        "fvk_mesh_read_gmsh": [C.c_char_p, C.c_int, C.c_double, _ip, C.c_int, C.POINTER(_vp)],
        "fvk_mesh_get_elm_idx": [_vp, _ip],
        "fvk_mesh_set_build_device": [C.c_int],
        "fvk_mesh_get_cells": [_vp, _dp, _ip, _ip, _ip],
        "fvk_mesh_get_topology": [_vp] + [_ip] * 7,
        "fvk_mesh_get_geometry": [_vp] + [_dp] * 9,
        "fvk_ctx_create": [_vp, C.c_int, C.c_int, C.POINTER(_vp)],
        "fvk_ctx_create_arrays": [C.c_int] * 4 + [_ip] * 6 + [_dp] * 9 + [C.c_int, C.c_int, C.POINTER(_vp)],
        "fvk_ctx_sync": [_vp],
        "fvk_ctx_get_perm": [_vp, _ip],
        "fvk_mesh_rcm": [_vp, _ip],
        "fvk_ctx_levels": [_vp, _ip, _ip],
        "fvk_set_matrix": [_vp, _dp, _dp, _dp],
        "fvk_solv_precond": [_vp, C.c_int, _dp, _dp],
        "fvk_solv_direct": [_vp, C.c_int, _dp, _dp, _dp, _dp],
        "fvk_flow_configure": [_vp, C.POINTER(FlowCfg)],
        "fvk_flow_scalars": [_vp, C.c_int, _dp, _dp],
        "fvk_solv_res": [_vp, C.c_int, _dp, _dp, _dp, C.c_int],
        "fvk_solv_res_dev": [_vp, C.c_int, _vp, _vp, _vp, C.c_int],
        "fvk_solv_cg": [_vp, C.c_int, C.c_double, C.c_int, C.c_int, _dp, _dp, _dp, C.c_int, C.POINTER(Stats)],
        "fvk_solv_bicg": [_vp, C.c_int, C.c_double, C.c_int, C.c_int, _dp, _dp, _dp, C.c_int, C.POINTER(Stats)],
        "fvk_solv_grad": [_vp, C.c_int, _dp, _dp, C.c_int, C.c_int],
        "fvk_sys_random": [_vp, C.c_uint64, C.c_double, _dp],
        "fvk_sys_set": [_vp, _dp, _dp, _dp],
        "fvk_sys_solve": [_vp, C.c_int, C.c_int, C.c_double, C.c_int, C.c_int, C.POINTER(SolveInfo)],
        "fvk_sys_get_x": [_vp, _dp],
        "fvk_ctx_set_param": [_vp, C.c_char_p, C.c_double],
        "fvk_factor_set": [_vp, _fp, _fp],
        "fvk_factor_apply": [_vp, _dp, _dp],
        "fvk_factor_get": [_vp, _fp, _fp],
        "fvk_factor_features": [_vp, _fp],
        "fvk_factor_fsai": [_vp],
        "fvk_model_nparam": [],
        "fvk_model_create": [C.c_int, C.c_uint64, C.POINTER(_vp)],
        "fvk_model_get": [_vp, _fp],
        "fvk_model_set": [_vp, _fp],
        "fvk_factor_build": [_vp, _vp],
        "fvk_model_loss": [_vp, _vp, _dp, _fp, _fp, _fp],
        "fvk_model_train": [_vp, _vp, C.c_int, C.c_double, _dp],
        "fvk_sys_get": [_vp, _dp, _dp, _dp],
        "fvk_bench_spmv": [_vp, C.c_int, _dp],
        "fvk_bench_sweep": [_vp, C.c_int, C.c_int, C.c_int, _dp],
        "fvk_bench_copy": [C.c_int, C.c_long, C.c_int, _dp],
        "fvk_bench_stream": [C.c_int, C.c_long, C.c_int, C.c_int, _dp],
        "fvk_flow_create": [_vp, C.POINTER(FlowCfg), C.POINTER(_vp)],
        "fvk_flow_set_bnd_u": [_vp, C.c_int, _dp],
        "fvk_flow_set_state": [_vp] + [_dp] * 5,
        "fvk_flow_get_state": [_vp] + [_dp] * 7,
        "fvk_flow_get_matrix": [_vp, _dp, _dp, _dp],
        "fvk_flow_setup_eqn": [_vp, C.c_int, C.POINTER(OuterHist)],
        "fvk_flow_setup_pp": [_vp, C.c_int, C.POINTER(OuterHist)],
        "fvk_flow_outer": [_vp, C.c_int, C.POINTER(OuterHist)],
        "fvk_flow_run": [_vp, C.c_int, C.c_int, C.POINTER(OuterHist)],
        "fvk_flow_run_graph": [_vp, C.c_int, C.c_int, C.POINTER(OuterHist)],
        "fvk_flow_set_model": [_vp, _vp, C.c_int, C.c_double],
        "fvk_flow_get_rhs": [_vp, _dp],
        "fvk_eqn_outp": [_vp, _vp] + [C.c_int] * 5 + [C.c_char_p],
        "fvk_prp_outp_p": [_vp, _vp] + [C.c_int] * 5 + [C.c_char_p],
        "fvk_write_gmsh_field": [C.c_char_p, C.c_int, _ip, _dp],
        "fvk_write_gmsh_vector_field": [C.c_char_p, C.c_int, C.c_int, _ip, _dp],
        "fvk_fmt_list_real": [C.c_double, C.c_char_p, C.c_int],
        "fvk_fl_mesh_read_gmsh": [C.c_char_p, C.c_int, C.c_double, C.POINTER(_vp)],
        "fvk_fl_mesh_get_info": [_vp, _ip],
        "fvk_fl_mesh_from_records": [C.c_int, C.c_int, _dp, C.c_int, _ip, _ip, _ip, _ip, C.POINTER(_vp)],
        "fvk_fl_mesh_get_arrays": [_vp, _dp] + [_ip] * 9,
        "fvk_fl_ctx_create": [C.c_int, C.c_int, _dp, C.c_int, _ip, C.c_int, _ip, _ip, C.c_int, _ip, _ip, _ip,
                              _ip, _ip, C.c_int, C.POINTER(_vp)],
        "fvk_fl_geometry": [_vp] + [_dp] * 10,
        "fvk_fl_ctx_set_param": [_vp, C.c_char_p, C.c_double],
        "fvk_fl_ls_gradient": [_vp, C.c_int, _dp, _dp, _dp, C.c_int],
        "fvk_fl_lin_eqns_solver": [_vp, _dp, _dp, _dp, _dp, C.c_int, C.c_double, C.c_int],
        "fvk_fl_flow_create": [_vp, _dp, C.POINTER(_vp)],
        "fvk_fl_flow_set": [_vp, _dp, _dp, _dp, _dp],
        "fvk_fl_flow_get": [_vp] + [_dp] * 6,
        "fvk_fl_flow_get_system": [_vp] + [_dp] * 7,
        "fvk_fl_flow_step": [_vp, C.c_int, _dp],
        "fvk_fl_flow_run": [_vp, C.c_int, C.c_double, _dp, _dp, _ip],
    }
    for name, args in sig.items():
        # a library older than this binding lacks some entry points: those
        # fail when called, the rest keep working
        f = getattr(L, name, None)
        if f is None:
            continue
        f.argtypes = args
        f.restype = C.c_int
    for name in ("fvk_mesh_free", "fvk_ctx_destroy", "fvk_flow_destroy", "fvk_model_destroy", "fvk_fl_mesh_free",
                 "fvk_fl_ctx_destroy", "fvk_fl_flow_destroy"):
        f = getattr(L, name, None)
        if f is None:
            continue
        f.argtypes = [_vp]
        f.restype = None
    _lib = L
    atexit.register(_release_all)
    return L


def _check(rc):
    if rc < 0:
        raise FvkError(rc, lib().fvk_last_error().decode())
    return rc


def _d(a):
    return a.ctypes.data_as(_dp)


def _i(a):
    return a.ctypes.data_as(_ip)


def _f64(a, shape=None):
    a = np.ascontiguousarray(a, dtype=np.float64)
    return a.reshape(shape) if shape is not None else a


# ---------------------------------------------------------------- mesh
def mesh_build_device(device):
    """Mesh constructors build connectivity and geometry on this GPU (-1:
    on the host, the default); the arrays are the same bit for bit."""
    _check(lib().fvk_mesh_set_build_device(int(device)))


class Mesh:
    """The grd structure after grd_read / grd_conn / grd_geom."""

    def __init__(self, handle):
        self.h = handle
        _track("mesh", self)
        info = MeshInfo()
        _check(lib().fvk_mesh_get_info(self.h, C.byref(info)))
        self.info = info
        for k, _ in MeshInfo._fields_:
            setattr(self, k, getattr(info, k))

    @classmethod
    def read_neu(cls, path, l_sc=1.0):
        h = _vp()
        _check(lib().fvk_mesh_read_neu(os.fsencode(path), l_sc, C.byref(h)))
        return cls(h)

    @classmethod
    def read_gmsh(cls, path, n_dx=0, l_sc=1.0, code_map=None):
        """gmsh 2.2 ASCII mesh; code_map {physical tag: boundary code}."""
        pairs = np.array([[k, v] for k, v in (code_map or {}).items()], np.int32).reshape(-1)
        h = _vp()
        _check(lib().fvk_mesh_read_gmsh(os.fsencode(path), n_dx, l_sc, _i(pairs) if len(pairs) else None,
                                        len(pairs) // 2, C.byref(h)))
        return cls(h)

    def elm_idx(self):
        e = np.zeros(self.n_vol, np.int32)
        _check(lib().fvk_mesh_get_elm_idx(self.h, _i(e)))
        return e

    def cells(self):
        """Vertex coordinates (n_vrt, 3), element types, cell_vptr and the
        1-based cell vertex lists (Gambit order)."""
        vx = np.zeros((self.n_vrt, 3))
        ty = np.zeros(self.n_vol, np.int32)
        vp = np.zeros(self.n_vol + 1, np.int32)
        _check(lib().fvk_mesh_get_cells(self.h, None, None, _i(vp), None))
        vr = np.zeros(int(vp[-1]), np.int32)
        _check(lib().fvk_mesh_get_cells(self.h, _d(vx), _i(ty), _i(vp), _i(vr)))
        return vx, ty, vp, vr

    @classmethod
    def box(cls, nx, ny, nz, jitter=0.0, seed=0, lid=31, wall=32, L=(1.0, 1.0, 1.0)):
        h = _vp()
        _check(lib().fvk_mesh_box(nx, ny, nz, L[0], L[1], L[2], jitter, seed, lid, wall, C.byref(h)))
        return cls(h)

    @classmethod
    def from_cells(cls, n_dx, vx, v_ty, v_vptr, v_vrt, bnd):
        vx3 = np.zeros((len(vx), 3))
        vx3[:, : np.shape(vx)[1]] = vx
        arrs = [np.ascontiguousarray(a, np.int32) for a in (v_ty, v_vptr, v_vrt, bnd)]
        h = _vp()
        _check(lib().fvk_mesh_from_cells(n_dx, len(vx3), _d(vx3), len(arrs[0]), *[_i(a) for a in arrs], C.byref(h)))
        return cls(h)

    def topology(self):
        nf, nv, ns = self.n_fce, self.n_vol, self.n_slot
        o = dict(fce_ip=np.zeros(nf, np.int32), fce_in=np.zeros(nf, np.int32), fce_ib=np.zeros(nf, np.int32),
                 fce_bty=np.zeros(nf, np.int32), cell_ptr=np.zeros(nv + 1, np.int32),
                 fce_lst=np.zeros(ns, np.int32), ng_lst=np.zeros(ns, np.int32))
        _check(lib().fvk_mesh_get_topology(self.h, *[_i(o[k]) for k in ("fce_ip", "fce_in", "fce_ib", "fce_bty",
                                                                         "cell_ptr", "fce_lst", "ng_lst")]))
        return o

    def geometry(self):
        nd, nf, nv = self.n_dx, self.n_fce, self.n_vol
        o = dict(vol_x=np.zeros((nv, nd)), vol_v=np.zeros(nv), fce_x=np.zeros((nf, nd)), fce_n=np.zeros((nf, nd)),
                 fce_dxp=np.zeros((nf, nd)), fce_dxn=np.zeros((nf, nd)), fce_s=np.zeros(nf),
                 fce_lpn=np.zeros(nf), fce_wfp=np.zeros(nf))
        _check(lib().fvk_mesh_get_geometry(self.h, *[_d(o[k]) for k in ("vol_x", "vol_v", "fce_x", "fce_n",
                                                                         "fce_dxp", "fce_dxn", "fce_s", "fce_lpn",
                                                                         "fce_wfp")]))
        return o

    def rcm(self):
        """Reverse Cuthill-McKee cell order (1-based reference cell ids by
        row), as a context built with rcm=True uses it; host only."""
        p = np.zeros(self.n_vol, np.int32)
        _check(lib().fvk_mesh_rcm(self.h, _i(p)))
        return p

    def _release(self):
        if self.h:
            lib().fvk_mesh_free(self.h)
            self.h = None

    free = _release

    def __del__(self):
        if not _finalizing():
            self._release()


# ---------------------------------------------------------------- context
class Ctx:
    """Topology and geometry resident on the GPU (the product of grd_solv)."""

    def __init__(self, handle, n_vol, n_fce, n_dx, n_bf):
        self.h = handle
        self._flows = weakref.WeakSet()  # flows on this context: released first
        _track("ctx", self)
        self.n_vol, self.n_fce, self.n_dx, self.n_bf = n_vol, n_fce, n_dx, n_bf

    def perm(self):
        p = np.zeros(self.n_vol, np.int32)
        _check(lib().fvk_ctx_get_perm(self.h, _i(p)))
        return p

    def levels(self):
        a, b = C.c_int(), C.c_int()
        _check(lib().fvk_ctx_levels(self.h, C.byref(a), C.byref(b)))
        return a.value, b.value

    def sync(self):
        _check(lib().fvk_ctx_sync(self.h))

    def _release(self):
        if self.h:
            for f in list(self._flows):
                f._release()
            lib().fvk_ctx_destroy(self.h)
            self.h = None

    destroy = _release

    def __del__(self):
        if not _finalizing():
            self._release()


def grd_solv(mesh: Mesh, device=0, rcm=False):
    """Build the device context (grd_solv.f90:7-24 builds v_lst/ng)."""
    h = _vp()
    _check(lib().fvk_ctx_create(mesh.h, device, CTX_RCM if rcm else 0, C.byref(h)))
    return Ctx(h, mesh.n_vol, mesh.n_fce, mesh.n_dx, mesh.n_bf)


def grd_solv_arrays(topo, geom, n_dx, n_bf, device=0, rcm=False):
    """Context from reference-layout arrays (the Fortran drop-in path)."""
    t = {k: np.ascontiguousarray(v, np.int32) for k, v in topo.items() if k in
         ("fce_ip", "fce_in", "fce_ib", "fce_bty", "cell_ptr", "fce_lst")}
    g = {k: np.ascontiguousarray(v, np.float64) for k, v in geom.items()}
    n_vol, n_fce = len(t["cell_ptr"]) - 1, len(t["fce_ip"])
    h = _vp()
    _check(lib().fvk_ctx_create_arrays(
        n_dx, n_vol, n_fce, n_bf, *[_i(t[k]) for k in ("fce_ip", "fce_in", "fce_ib", "fce_bty", "cell_ptr", "fce_lst")],
        *[_d(g[k]) for k in ("vol_x", "vol_v", "fce_x", "fce_n", "fce_dxp", "fce_dxn", "fce_s", "fce_lpn",
                             "fce_wfp")], device, CTX_RCM if rcm else 0, C.byref(h)))
    return Ctx(h, n_vol, n_fce, n_dx, n_bf)


# ---------------------------------------------------------------- solvers
def set_matrix(ctx: Ctx, pl, ng_p, ng_n):
    """pl(n_vol), ng(:)%p, ng(:)%n (solv_data.f90:14-22)."""
    _check(lib().fvk_set_matrix(ctx.h, _d(_f64(pl)), _d(_f64(ng_p)), _d(_f64(ng_n))))


def solv_precond(ctx: Ctx, imat):
    c = np.zeros(ctx.n_vol)
    cr = np.zeros(ctx.n_vol)
    _check(lib().fvk_solv_precond(ctx.h, imat, _d(c), _d(cr)))
    return c, cr


def solv_direct(ctx: Ctx, imat, b, c=None, c_r=None):
    """x = M^-1 b; c, c_r as solv_precond returns them (None: factor of the
    current matrix)."""
    x = np.zeros(ctx.n_vol)
    bb = _f64(b)
    if c is None:
        _check(lib().fvk_solv_direct(ctx.h, imat, None, None, _d(bb), _d(x)))
    else:
        cc, cr = _f64(c), _f64(c_r)
        _check(lib().fvk_solv_direct(ctx.h, imat, _d(cc), _d(cr), _d(bb), _d(x)))
    return x


def _as2(a, n):
    a = np.asarray(a, np.float64)
    return a.reshape(n, -1) if a.ndim == 1 else a


def solv_res(ctx: Ctx, phi, res0, not_r0=False):
    """res0 <- res0 - A phi (unless not_r0); returns res0_n (L1 per component)."""
    res0 = _as2(res0, ctx.n_vol)
    nvc = res0.shape[1]
    col = np.asfortranarray(res0).ravel(order="F").copy()
    ph = _f64(phi if phi is not None else np.zeros((ctx.n_vol, nvc))).reshape(ctx.n_vol, nvc).copy()
    rn = np.zeros(nvc)
    _check(lib().fvk_solv_res(ctx.h, nvc, _d(ph), _d(col), _d(rn), int(bool(not_r0))))
    res0[...] = col.reshape(nvc, ctx.n_vol).T
    return rn


def solv_res_dev(ctx: Ctx, n_vc, phi_ptr, res0_ptr, res0_n_ptr, not_r0=False):
    """fvk_solv_res_dev on device pointers (integers): component-major
    arrays in the context's row order; queued on the context's stream."""
    _check(lib().fvk_solv_res_dev(ctx.h, n_vc, phi_ptr, res0_ptr, res0_n_ptr, int(bool(not_r0))))


def _solve(fn, ctx, imat, reduc, n_iter, phi, res0, not_r0):
    phi = _as2(phi, ctx.n_vol)
    res0 = _as2(res0, ctx.n_vol)
    nvc = res0.shape[1]
    ph = np.ascontiguousarray(phi).copy()
    col = res0.T.copy().ravel()
    rn = np.zeros(nvc)
    st = Stats()
    rc = _check(fn(ctx.h, imat, reduc, n_iter, nvc, _d(ph), _d(col), _d(rn), int(bool(not_r0)), C.byref(st)))
    phi[...] = ph
    res0[...] = col.reshape(nvc, ctx.n_vol).T
    return rn, st, rc


def solv_cg(ctx: Ctx, imat, reduc, n_iter, phi, res0, not_r0=False):
    """PCG on the current matrix; phi (n_vol, n_vc) and res0 are updated in place.
    Returns (res0_n, stats)."""
    rn, st, _ = _solve(lib().fvk_solv_cg, ctx, imat, reduc, n_iter, phi, res0, not_r0)
    return rn, st


def solv_bicg(ctx: Ctx, imat, reduc, n_iter, phi, res0, not_r0=False):
    rn, st, _ = _solve(lib().fvk_solv_bicg, ctx, imat, reduc, n_iter, phi, res0, not_r0)
    return rn, st


def solv_grad(ctx: Ctx, phi, n_iter, neu):
    """phi: (n_vol + n_bf, n_vc) cells then boundary slots ib.  Returns
    grad (n_vol, n_vc, n_dx)."""
    phi = _as2(phi, ctx.n_vol + ctx.n_bf)
    nvc = phi.shape[1]
    g = np.zeros((ctx.n_vol, nvc, ctx.n_dx))
    _check(lib().fvk_solv_grad(ctx.h, nvc, _d(_f64(phi)), _d(g), n_iter, neu))
    return g


# ---------------------------------------------------------------- systems
def sys_random(ctx: Ctx, seed, sigma=1.0):
    """Stand-alone pressure system: a_f = s/l_pn*exp(sigma N(0,1)), zero-mean
    N(0,1) RHS.  Returns b (reference order)."""
    b = np.zeros(ctx.n_vol)
    _check(lib().fvk_sys_random(ctx.h, seed, sigma, _d(b)))
    return b


def sys_set(ctx: Ctx, pl, a, b):
    keep = [_f64(pl), _f64(a), _f64(b)]
    _check(lib().fvk_sys_set(ctx.h, *[_d(k) for k in keep]))


def sys_get(ctx: Ctx):
    pl, a, b = np.zeros(ctx.n_vol), np.zeros(ctx.n_fce), np.zeros(ctx.n_vol)
    _check(lib().fvk_sys_get(ctx.h, _d(pl), _d(a), _d(b)))
    return pl, a, b


def sys_solve(ctx: Ctx, method=SOLVE_PCG64, imat=0, tol=1e-8, max_iter=100000, time_kernel=False):
    info = SolveInfo()
    rc = _check(lib().fvk_sys_solve(ctx.h, method, imat, tol, max_iter, int(bool(time_kernel)), C.byref(info)))
    out = {k: getattr(info, k) for k, _ in SolveInfo._fields_}
    out["inner_iters"] = list(info.inner_iters)[:max(info.outer, 0)]
    out["outer_res"] = list(info.outer_res)[:max(info.outer, 0)]
    out["status"] = rc
    return out


def set_param(ctx: Ctx, name, value):
    _check(lib().fvk_ctx_set_param(ctx.h, name.encode(), float(value)))


def factor_set(ctx: Ctx, l, sig):
    """Factor G = D^-1/2 (I + L) Delta from per-face l_f and sig_f (n_fce
    entries, reference face order; boundary faces ignored)."""
    l = np.ascontiguousarray(l, np.float32)
    sig = np.ascontiguousarray(sig, np.float32)
    _check(lib().fvk_factor_set(ctx.h, l.ctypes.data_as(_fp), sig.ctypes.data_as(_fp)))


def factor_apply(ctx: Ctx, r):
    r = np.ascontiguousarray(r, np.float64)
    z = np.zeros(ctx.n_vol)
    _check(lib().fvk_factor_apply(ctx.h, _d(r), _d(z)))
    return z


def factor_get(ctx: Ctx):
    l = np.zeros(ctx.n_fce, np.float32)
    sig = np.zeros(ctx.n_fce, np.float32)
    _check(lib().fvk_factor_get(ctx.h, l.ctypes.data_as(_fp), sig.ctypes.data_as(_fp)))
    return l, sig


def factor_fsai(ctx: Ctx):
    """Factor of the current matrix by FSAI in closed form (on the GPU)."""
    _check(lib().fvk_factor_fsai(ctx.h))


def factor_features(ctx: Ctx):
    """The network's twelve inputs per face as it reads them (bfloat16
    values), (n_fce, 12), reference face order, zero rows on boundary faces."""
    X = np.zeros((ctx.n_fce, 12), np.float32)
    _check(lib().fvk_factor_features(ctx.h, X.ctypes.data_as(_fp)))
    return X


class Model:
    """The per-face MLP of the learned factor, resident on one GPU."""

    def __init__(self, seed=0, device=0):
        self.h = _vp()
        _check(lib().fvk_model_create(device, seed, C.byref(self.h)))
        _track("model", self)
        self.nparam = lib().fvk_model_nparam()

    def get(self):
        th = np.zeros(self.nparam, np.float32)
        _check(lib().fvk_model_get(self.h, th.ctypes.data_as(_fp)))
        return th

    def set(self, theta):
        th = np.ascontiguousarray(theta, np.float32)
        assert th.size == self.nparam
        _check(lib().fvk_model_set(self.h, th.ctypes.data_as(_fp)))

    def build_factor(self, ctx: Ctx):
        _check(lib().fvk_factor_build(ctx.h, self.h))

    def loss(self, ctx: Ctx, grads=False):
        J = C.c_double()
        if not grads:
            _check(lib().fvk_model_loss(self.h, ctx.h, C.byref(J), None, None, None))
            return J.value
        g = np.zeros(self.nparam, np.float32)
        dl = np.zeros(ctx.n_fce, np.float32)
        ds = np.zeros(ctx.n_fce, np.float32)
        _check(lib().fvk_model_loss(self.h, ctx.h, C.byref(J), g.ctypes.data_as(_fp), dl.ctypes.data_as(_fp),
                                    ds.ctypes.data_as(_fp)))
        return J.value, g, dl, ds

    def train(self, ctx: Ctx, steps, lr=1e-3):
        losses = np.zeros(steps)
        _check(lib().fvk_model_train(self.h, ctx.h, steps, lr, _d(losses)))
        return losses

    def _release(self):
        if self.h:
            lib().fvk_model_destroy(self.h)
            self.h = None

    destroy = _release

    def __del__(self):
        if not _finalizing():
            self._release()


def sys_x(ctx: Ctx):
    x = np.zeros(ctx.n_vol)
    _check(lib().fvk_sys_get_x(ctx.h, _d(x)))
    return x


def bench_spmv(ctx: Ctx, reps=100):
    s = C.c_double()
    _check(lib().fvk_bench_spmv(ctx.h, reps, C.byref(s)))
    return s.value


def bench_sweep(ctx: Ctx, imat, nvc=1, reps=20):
    """Seconds per application of the DIC / DILU factor (both sweeps)."""
    s = C.c_double()
    _check(lib().fvk_bench_sweep(ctx.h, imat, nvc, reps, C.byref(s)))
    return s.value


def bench_copy(nbytes=1 << 30, reps=20, device=0):
    """Seconds per device copy of nbytes (read nbytes + write nbytes)."""
    s = C.c_double()
    _check(lib().fvk_bench_copy(device, int(nbytes), reps, C.byref(s)))
    return s.value


def bench_stream(kind, nbytes=1 << 30, reps=20, device=0):
    """Seconds per pass of a streaming probe: kind "copy" (nbytes read and
    written), "read" or "write" with four 16-byte loads in flight per thread,
    "copy_nt" or "read_nt" with eight non-temporal ones and a resident grid."""
    s = C.c_double()
    kinds = {"copy": 0, "read": 1, "write": 2, "copy_nt": 3, "read_nt": 4}
    _check(lib().fvk_bench_stream(device, int(nbytes), reps, kinds[kind], C.byref(s)))
    return s.value


# ---------------------------------------------------------------- flow
def flow_cfg(**kw):
    """Shipped settings (cfd_setup.f90:61,528,580-581,662,742; cfd_solve.f90:71-72)
    with the Re=100 cavity properties; keywords override."""
    c = dict(simpler=1, imat_u=2, imat_p=1, n_bicg=5, n_cg=20, n_grad=2, c_sch=1, urf_u=0.7, urf_p=0.1,
             reduc_u=0.1, reduc_p=0.1, c_bl=0.9, rho=1.0, mu=0.01, p_solver=0, p_tol=1e-8)
    c.update(kw)
    return FlowCfg(**c)


def _hist(h):
    return dict(reduc=h.reduc, redpp=h.redpp, redpp_sum=h.redpp_sum, red_p=h.red_p, red_p_sum=h.red_p_sum,
                mx_res_n=h.mx_res_n, it_u=list(h.it_u), it_pp=h.it_pp, it_p=h.it_p,
                res0_n_u=list(h.res0_n_u), res0_n_pp=h.res0_n_pp, res0_n_p=h.res0_n_p)


class Flow:
    """Single-phase steady SIMPLE state on a context (cfd_setup_eqn /
    cfd_setup_pp / the cfd_solve outer loop)."""

    def __init__(self, ctx: Ctx, cfg: FlowCfg):
        self.ctx = ctx
        self.cfg = cfg
        self.h = _vp()
        _check(lib().fvk_flow_create(ctx.h, C.byref(cfg), C.byref(self.h)))
        _track("flow", self)
        ctx._flows.add(self)

    def set_bnd_u(self, code, u):
        u3 = np.zeros(3)
        u3[: len(u)] = u
        _check(lib().fvk_flow_set_bnd_u(self.h, code, _d(u3)))

    def cfd_setup_eqn(self, it):
        h = OuterHist()
        _check(lib().fvk_flow_setup_eqn(self.h, it, C.byref(h)))
        return _hist(h)

    def cfd_setup_pp(self, it):
        h = OuterHist()
        _check(lib().fvk_flow_setup_pp(self.h, it, C.byref(h)))
        return _hist(h)

    def outer(self, it):
        h = OuterHist()
        _check(lib().fvk_flow_outer(self.h, it, C.byref(h)))
        return _hist(h)

    def run(self, it0, n):
        hs = (OuterHist * max(n, 1))()
        _check(lib().fvk_flow_run(self.h, it0, n, hs))
        return [_hist(hs[k]) for k in range(n)]

    def run_graph(self, it0, n):
        """n outer iterations, it0 eagerly and the rest as HIP-graph replays."""
        H = (OuterHist * max(n, 1))()
        _check(lib().fvk_flow_run_graph(self.h, it0, n, H))
        return [_hist(H[k]) for k in range(n)]

    def rhs(self):
        r = np.zeros(self.ctx.n_vol)
        _check(lib().fvk_flow_get_rhs(self.h, _d(r)))
        return r

    def set_model(self, model, train_steps=0, lr=1e-3):
        _check(lib().fvk_flow_set_model(self.h, model.h if model is not None else None, train_steps, lr))

    def state(self):
        c = self.ctx
        nd = c.n_dx
        o = dict(u_cell=np.zeros((c.n_vol, nd)), u_bnd=np.zeros((max(c.n_bf, 1), nd)), p=np.zeros(c.n_vol),
                 dpdx=np.zeros((c.n_vol, nd)), spd=np.zeros(c.n_fce), cs=np.zeros(c.n_vol),
                 frc_dns=np.zeros(c.n_fce))
        _check(lib().fvk_flow_get_state(self.h, *[_d(o[k]) for k in ("u_cell", "u_bnd", "p", "dpdx", "spd", "cs",
                                                                     "frc_dns")]))
        o["u_bnd"] = o["u_bnd"][: c.n_bf]
        return o

    def set_state(self, u_cell=None, u_bnd=None, p=None, dpdx=None, spd=None):
        keep = [None if a is None else _f64(a) for a in (u_cell, u_bnd, p, dpdx, spd)]
        _check(lib().fvk_flow_set_state(self.h, *[None if a is None else _d(a) for a in keep]))

    def matrix(self):
        c = self.ctx
        pl, gp, gn = np.zeros(c.n_vol), np.zeros(c.n_fce), np.zeros(c.n_fce)
        _check(lib().fvk_flow_get_matrix(self.h, _d(pl), _d(gp), _d(gn)))
        return pl, gp, gn

    def _release(self):
        if self.h:
            lib().fvk_flow_destroy(self.h)
            self.h = None

    destroy = _release

    def __del__(self):
        if not _finalizing():
            self._release()


# ---------------------------------------------------------------- output
def fmt_list_real(x):
    """flang's list-directed editing of one REAL(8) value (no leading blank)."""
    buf = C.create_string_buffer(64)
    _check(lib().fvk_fmt_list_real(float(x), buf, 64))
    return buf.value.decode()


def _title(kind, igrd, iph, k, ict, n_tl):
    one = lambda v: str(v) if 0 <= v <= 9 else "*"  # noqa: E731
    t = "grd%s%%phs%s%%%s%s" % (one(igrd), one(iph), kind, one(k))
    if n_tl > 1 and ict > 0:
        t += "@" + ("%04d" % ict if ict <= 9999 else "****")
    return t


def _dir(d):
    d = os.fspath(d)
    return d if d.endswith("/") or d == "" else d + "/"


def eqn_outp(mesh: Mesh, flow: Flow, dir, igrd=1, iph=1, ieq=1, ict=0, n_tl=1):
    """Tecplot file of the velocity (cfd_outp.f90 eqn_outp); returns its path."""
    d = _dir(dir)
    _check(lib().fvk_eqn_outp(mesh.h, flow.h, igrd, iph, ieq, ict, n_tl, os.fsencode(d)))
    return d + _title("eqn", igrd, iph, ieq, ict, n_tl) + ".dat"


def prp_outp_p(mesh: Mesh, flow: Flow, dir, igrd=1, iph=1, ipr=1, ict=0, n_tl=1):
    """Tecplot file of the pressure (cfd_outp.f90 prp_outp with p=.true.)."""
    d = _dir(dir)
    _check(lib().fvk_prp_outp_p(mesh.h, flow.h, igrd, iph, ipr, ict, n_tl, os.fsencode(d)))
    return d + _title("prp", igrd, iph, ipr, ict, n_tl) + ".dat"


def write_gmsh_field(fname, elm_idx, phi):
    """gmsh ElementData file of one value per element (flowlab's
    write_gmsh_field_); phi (n,) or (n, n_dim) for the vector form."""
    idx = np.ascontiguousarray(elm_idx, np.int32)
    ph = np.ascontiguousarray(phi, np.float64)
    if ph.ndim == 1:
        _check(lib().fvk_write_gmsh_field(os.fsencode(fname), len(idx), _i(idx), _d(ph)))
    else:
        _check(lib().fvk_write_gmsh_vector_field(os.fsencode(fname), len(idx), ph.shape[1], _i(idx), _d(ph)))
