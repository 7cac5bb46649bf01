"""ctypes binding of the CPU oracle (liboracle.so).

TEST INFRASTRUCTURE: imported only by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, as the checker.  The product (programs_amd/)
never imports it.

Arrays come back as numpy views converted to 0-based entity order: element
k of a per-cell array is cell k+1 of the reference, and index values
(ip, in, fce_lst, ng_lst, ib) keep the reference's 1-based numbering.
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")

_P = C.POINTER
dp = _P(C.c_double)
ip = _P(C.c_int)


class Mesh(C.Structure):
    _fields_ = [("n_dx", C.c_int), ("n_vrt", C.c_int), ("n_vol", C.c_int), ("n_fce", C.c_int),
                ("n_bf", C.c_int), ("n_vbf", C.c_int), ("n_slot", C.c_int),
                ("vx", dp),
                ("v_ty", ip), ("v_nvrt", ip), ("v_nfce", ip),
                ("v_vptr", ip), ("v_vrt", ip), ("v_fptr", ip),
                ("v_bnd", ip), ("v_fce", ip), ("v_ng", ip),
                ("v_x", dp), ("v_v", dp), ("v_vr", dp), ("v_sf", dp),
                ("f_ip", ip), ("f_in", ip), ("f_ib", ip), ("f_bty", ip), ("f_nvrt", ip),
                ("f_vptr", ip), ("f_vrt", ip),
                ("f_x", dp), ("f_n", dp), ("f_dxp", dp), ("f_dxn", dp),
                ("f_s", dp), ("f_lpn", dp), ("f_wfp", dp)]


class Sys(C.Structure):
    _fields_ = [("m", _P(Mesh)), ("pl", dp), ("ng_p", dp), ("ng_n", dp)]


class Cfg(C.Structure):
    _fields_ = [("simpler", C.c_int), ("imat_u", C.c_int), ("imat_p", C.c_int),
                ("n_bicg", C.c_int), ("n_cg", C.c_int), ("n_grad", C.c_int), ("c_sch", C.c_int),
                ("urf_u", C.c_double), ("urf_p", C.c_double), ("reduc_u", C.c_double),
                ("reduc_p", C.c_double), ("c_bl", C.c_double), ("rho", C.c_double),
                ("mu", C.c_double), ("p_solver", C.c_int), ("p_tol", C.c_double), ("p_cap", C.c_int)]


class Flow(C.Structure):
    _fields_ = [("m", _P(Mesh)), ("cfg", Cfg), ("n_vc", C.c_int),
                ("u", dp), ("p", dp), ("dpdx", dp), ("spd", dp), ("frc_dns", dp),
                ("cs", dp), ("frc_vol", dp),
                ("frc_c", dp), ("dns_c", dp), ("vis_c", dp),
                ("frc_b", dp), ("dns_b", dp), ("vis_b", dp), ("co_num", dp),
                ("A", Sys),
                ("res_n_u", C.c_double), ("pp_res_n", C.c_double), ("mx_res_n", C.c_double),
                ("c_in", C.c_double), ("c_out", C.c_double),
                ("h_reduc", C.c_double), ("h_redpp", C.c_double), ("h_redpp_sum", C.c_double),
                ("h_red_p", C.c_double), ("h_red_p_sum", C.c_double),
                ("res0_n_u", C.c_double * 3), ("res0_n_pp", C.c_double), ("res0_n_p", C.c_double),
                ("it_u", C.c_int * 3), ("it_pp", C.c_int), ("it_p", C.c_int),
                ("pc_last", dp), ("rhs_last", dp)]


def build():
    """Compile liboracle.so with the Makefile next to this file."""
    subprocess.run(["make", "-s", "-C", HERE], check=True)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        L = C.CDLL(LIB)
        L.fvo_small.restype = C.c_double
        L.fvo_large.restype = C.c_double
        L.fvo_read_neu.restype = _P(Mesh)
        L.fvo_read_neu.argtypes = [C.c_char_p, C.c_double]
        L.fvo_gen_box.restype = _P(Mesh)
        L.fvo_gen_box.argtypes = [C.c_int, C.c_int, C.c_int, C.c_double, C.c_double, C.c_double,
                                  C.c_double, C.c_uint64, C.c_int, C.c_int]
        L.fvo_mesh_raw.restype = _P(Mesh)
        L.fvo_mesh_raw.argtypes = [C.c_int, C.c_int, dp, C.c_int, ip, ip, ip, ip]
        for f in ("fvo_conn_literal", "fvo_conn_hash", "fvo_geom"):
            getattr(L, f).argtypes = [_P(Mesh)]
        L.fvo_mesh_free.argtypes = [_P(Mesh)]
        L.fvo_flow_new.restype = _P(Flow)
        L.fvo_flow_new.argtypes = [_P(Mesh), _P(Cfg)]
        L.fvo_flow_free.argtypes = [_P(Flow)]
        L.fvo_flow_set_bnd_u.argtypes = [_P(Flow), C.c_int, dp]
        L.fvo_grad.argtypes = [_P(Mesh), C.c_int, dp, dp, C.c_int, C.c_int]
        L.fvo_outer.argtypes = [_P(Flow), C.c_int, dp]
        L.fvo_setup_eqn.argtypes = [_P(Flow), C.c_int]
        L.fvo_setup_pp.argtypes = [_P(Flow), C.c_int]
        L.fvo_solv_precond.argtypes = [_P(Sys), C.c_int, dp, dp]
        L.fvo_solv_direct.argtypes = [_P(Sys), C.c_int, dp, dp, dp, dp]
        L.fvo_spmv.argtypes = [_P(Sys), dp, dp]
        for f in ("fvo_solv_cg", "fvo_solv_bicg"):
            getattr(L, f).argtypes = [_P(Sys), C.c_int, C.c_double, C.c_int, C.c_int, dp, C.c_int,
                                      dp, dp, C.c_int, ip, dp]
        L.fvo_solv_res.argtypes = [_P(Sys), C.c_int, dp, C.c_int, dp, dp, C.c_int]
        L.fvo_set_sum_order.argtypes = [C.c_int]
        L.fvo_is_singular.argtypes = [_P(Sys)]
        L.fvo_pcg_omp.argtypes = [_P(Sys), dp, dp, C.c_int, C.c_int]
        L.fvo_pcg_omp.restype = C.c_double
        L.fvo_pcg_tol.argtypes = [_P(Sys), dp, dp, C.c_double, C.c_int, ip, dp]
        _lib = L
    return _lib


def small():
    return lib().fvo_small()


class sum_order:
    """Context manager: run the oracle with pairwise reductions (1) instead of
    the reference's sequential ones (0)."""

    def __init__(self, order):
        self.order = order

    def __enter__(self):
        lib().fvo_set_sum_order(self.order)

    def __exit__(self, *a):
        lib().fvo_set_sum_order(0)


def _d(a):
    return a.ctypes.data_as(dp)


def _i(a):
    return a.ctypes.data_as(ip)


def arr(ptr, n, dtype=np.float64):
    """Copy n elements of a C array into numpy."""
    ct = C.c_double if dtype == np.float64 else C.c_int
    return np.ctypeslib.as_array(C.cast(ptr, _P(ct)), shape=(n,)).copy()


class OMesh:
    """Owning wrapper of an oracle mesh."""

    def __init__(self, ptr):
        if not ptr:
            raise RuntimeError("oracle mesh construction failed")
        self.ptr = ptr
        self.m = ptr.contents

    @classmethod
    def read_neu(cls, path, conn="hash"):
        o = cls(lib().fvo_read_neu(path.encode(), 1.0))
        o.finish(conn)
        return o

    @classmethod
    def box(cls, nx, ny, nz, jitter=0.0, seed=0, lid=31, wall=32, L=(1.0, 1.0, 1.0), conn="hash"):
        o = cls(lib().fvo_gen_box(nx, ny, nz, L[0], L[1], L[2], jitter, seed, lid, wall))
        o.finish(conn)
        return o

    @classmethod
    def raw(cls, n_dx, vx, v_ty, v_vptr, v_vrt, bnd, conn="hash"):
        vx = np.ascontiguousarray(vx, np.float64).reshape(-1, 3)
        keep = [np.ascontiguousarray(a, np.int32) for a in (v_ty, v_vptr, v_vrt, bnd)]
        o = cls(lib().fvo_mesh_raw(n_dx, vx.shape[0], _d(vx), len(keep[0]), *[_i(a) for a in keep]))
        o.finish(conn)
        return o

    def finish(self, conn):
        L = lib()
        (L.fvo_conn_literal if conn == "literal" else L.fvo_conn_hash)(self.ptr)
        L.fvo_geom(self.ptr)
        self.m = self.ptr.contents

    def free(self):
        if self.ptr:
            lib().fvo_mesh_free(self.ptr)
            self.ptr = None

    # 0-based numpy copies (entity k+1 at index k)
    def ints(self, name, n, stride=1):
        return arr(getattr(self.m, name), (n + 1) * stride, np.int32)[stride:]

    def reals(self, name, n, stride=1):
        return arr(getattr(self.m, name), (n + 1) * stride)[stride:]

    def vec(self, name, n):
        a = self.reals(name, n, 3).reshape(n, 3)
        return a[:, : self.m.n_dx].copy()

    def topology(self):
        m = self.m
        nf, nv, ns = m.n_fce, m.n_vol, m.n_slot
        return dict(
            n_dx=m.n_dx, n_vol=nv, n_fce=nf, n_bf=m.n_bf,
            fce_ip=self.ints("f_ip", nf), fce_in=self.ints("f_in", nf),
            fce_ib=self.ints("f_ib", nf), fce_bty=self.ints("f_bty", nf),
            fce_lst=self.ints("v_fce", ns), ng_lst=self.ints("v_ng", ns),
            cell_ptr=self.ints("v_fptr", nv + 1) - 1,
        )

    def geometry(self):
        m = self.m
        nf, nv = m.n_fce, m.n_vol
        return dict(
            vol_x=self.vec("v_x", nv), vol_v=self.reals("v_v", nv), vol_vr=self.reals("v_vr", nv),
            fce_x=self.vec("f_x", nf), fce_n=self.vec("f_n", nf), fce_dxp=self.vec("f_dxp", nf),
            fce_dxn=self.vec("f_dxn", nf), fce_s=self.reals("f_s", nf),
            fce_lpn=self.reals("f_lpn", nf), fce_wfp=self.reals("f_wfp", nf),
        )

    def grad(self, phi_cells, phi_bnd, n_iter, neu):
        """solv_grad of a scalar; phi_bnd indexed by ib-n_vol-1."""
        m = self.m
        phi = np.zeros(m.n_vbf + 1)
        phi[1:m.n_vol + 1] = phi_cells
        phi[m.n_vol + 1:] = phi_bnd
        g = np.zeros((m.n_vol + 1) * 3)
        lib().fvo_grad(self.ptr, 1, _d(phi), _d(g), n_iter, neu)
        return g.reshape(-1, 3)[1:, : m.n_dx].copy()


def default_cfg(**kw):
    """Shipped settings: SIMPLER, DIC pressure (imat 1), DILU momentum (imat 2),
    n_cg=20, n_bicg=5, reduc 0.1 (cfd_setup.f90:61,528,580-581,662,742), momentum
    urf 0.7, c_bl 0.9, minmod (cfd_solve.f90:71-72), Re=100 cavity."""
    c = dict(simpler=1, imat_u=2, imat_p=1, n_bicg=5, n_cg=20, n_grad=2, c_sch=1,
             urf_u=0.7, urf_p=0.1, reduc_u=0.1, reduc_p=0.1, c_bl=0.9, rho=1.0, mu=0.01,
             p_solver=0, p_tol=1e-8, p_cap=20000)
    c.update(kw)
    return Cfg(**c)


class OFlow:
    def __init__(self, mesh: OMesh, cfg: Cfg):
        self.mesh = mesh
        self.cfg = cfg
        self.ptr = lib().fvo_flow_new(mesh.ptr, C.byref(cfg))
        self.f = self.ptr.contents

    def set_bnd_u(self, code, u):
        u3 = np.zeros(3)
        u3[: len(u)] = u
        lib().fvo_flow_set_bnd_u(self.ptr, code, _d(u3))

    def outer(self, it):
        h = np.zeros(5)
        rc = lib().fvo_outer(self.ptr, it, _d(h))
        if rc < 0:
            raise RuntimeError("oracle SIMPLE iteration diverged")
        return h

    def last_pp_system(self):
        """The last p' solve: right-hand side and solution (cell order)."""
        nv = self.mesh.m.n_vol
        return arr(self.f.rhs_last, nv + 1)[1:], arr(self.f.pc_last, nv + 1)[1:]

    def fields(self):
        m = self.mesh.m
        nv, nb, nf, nvc = m.n_vol, m.n_vbf, m.n_fce, self.f.n_vc
        u = arr(self.f.u, (nb + 1) * nvc).reshape(nb + 1, nvc)
        return dict(
            u_cell=u[1:nv + 1].copy(), u_bnd=u[nv + 1:].copy(),
            p=arr(self.f.p, nv + 1)[1:], dpdx=arr(self.f.dpdx, 3 * (nv + 1)).reshape(-1, 3)[1:, : m.n_dx].copy(),
            spd=arr(self.f.spd, nf + 1)[1:], frc_dns=arr(self.f.frc_dns, nf + 1)[1:],
            cs=arr(self.f.cs, nv + 1)[1:],
            pl=arr(self.f.A.pl, nv + 1)[1:], ng_p=arr(self.f.A.ng_p, nf + 1)[1:],
            ng_n=arr(self.f.A.ng_n, nf + 1)[1:],
        )

    def free(self):
        if self.ptr:
            lib().fvo_flow_free(self.ptr)
            self.ptr = None


class OSys:
    """A face-based matrix (pl, ng_p, ng_n) on an oracle mesh, for solver tests."""

    def __init__(self, mesh: OMesh, pl, ngp, ngn):
        m = mesh.m
        self.mesh = mesh
        self.pl = np.concatenate([[0.0], np.asarray(pl, np.float64)])
        self.ngp = np.concatenate([[0.0], np.asarray(ngp, np.float64)])
        self.ngn = np.concatenate([[0.0], np.asarray(ngn, np.float64)])
        self.s = Sys(mesh.ptr, _d(self.pl), _d(self.ngp), _d(self.ngn))
        self.n = m.n_vol

    def precond(self, imat):
        c = np.zeros(self.n + 1)
        cr = np.zeros(self.n + 1)
        lib().fvo_solv_precond(C.byref(self.s), imat, _d(c), _d(cr))
        return c[1:], cr[1:]

    def direct(self, imat, b):
        c, cr = self.precond(imat)
        cc = np.concatenate([[0.0], c])
        crr = np.concatenate([[0.0], cr])
        x = np.zeros(self.n + 1)
        bb = np.concatenate([[0.0], b])
        lib().fvo_solv_direct(C.byref(self.s), imat, _d(cc), _d(crr), _d(x), _d(bb))
        return x[1:]

    def spmv(self, x):
        xx = np.concatenate([[0.0], x])
        y = np.zeros(self.n + 1)
        lib().fvo_spmv(C.byref(self.s), _d(xx), _d(y))
        return y[1:]

    def singular(self):
        return bool(lib().fvo_is_singular(C.byref(self.s)))

    def solve_tol(self, b, tol, max_iter=20000):
        """Solve to a relative L1 residual tol from x = 0 (fvo_pcg_tol).
        Returns x, iterations, |b|_1 after the projection, singular."""
        bb = np.concatenate([[0.0], np.asarray(b, np.float64)])
        x = np.zeros(self.n + 1)
        its = np.zeros(1, np.int32)
        r0 = np.zeros(1)
        sing = lib().fvo_pcg_tol(C.byref(self.s), _d(bb), _d(x), tol, max_iter, _i(its), _d(r0))
        return x[1:].copy(), int(its[0]), float(r0[0]), bool(sing)

    def pcg_omp(self, b, n_iter, threads=0):
        """CPU baseline timing on all cores: seconds for n_iter Jacobi-PCG
        iterations from x = 0 (OpenMP; not bitwise the sequential oracle).
        Returns (seconds, x)."""
        bb = np.concatenate([[0.0], np.asarray(b, np.float64)])
        x = np.zeros(self.n + 1)
        sec = lib().fvo_pcg_omp(C.byref(self.s), _d(bb), _d(x), int(n_iter), int(threads))
        return sec, x[1:].copy()

    def solve(self, kind, imat, reduc, n_iter, b, x0=None, not_r0=False):
        """b: (n, n_vc).  Returns x (n, n_vc), residual (n, n_vc), res0_n, iters, hist."""
        b = np.atleast_2d(np.asarray(b, np.float64).T).T if np.ndim(b) == 1 else np.asarray(b, np.float64)
        n, nvc = b.shape
        phi = np.zeros((n + 1, nvc))
        if x0 is not None:
            phi[1:] = np.asarray(x0).reshape(n, nvc)
        res = np.ascontiguousarray(b.T).copy()
        rn = np.zeros(nvc)
        its = np.zeros(nvc, np.int32)
        hist = np.zeros(nvc * n_iter)
        f = lib().fvo_solv_cg if kind == "cg" else lib().fvo_solv_bicg
        f(C.byref(self.s), imat, reduc, n_iter, nvc, _d(phi), nvc, _d(res), _d(rn), int(not_r0),
          _i(its), _d(hist))
        return phi[1:].copy(), res.T.copy(), rn, its, hist.reshape(nvc, n_iter)


def read_dump(path):
    """Read the record stream written by oracle/ref/fvk_ref_driver.f90."""
    out = {}
    with open(path, "rb") as fh:
        data = fh.read()
    o = 0
    while o < len(data):
        name = data[o:o + 16].decode().strip()
        kind = int(np.frombuffer(data, np.int32, 1, o + 16)[0])
        n = int(np.frombuffer(data, np.int64, 1, o + 20)[0])
        o += 28
        if kind == 1:
            out[name] = np.frombuffer(data, np.int32, n, o).copy()
            o += 4 * n
        else:
            out[name] = np.frombuffer(data, np.float64, n, o).copy()
            o += 8 * n
    return out


def parse_history(text):
    """The reference's print lines: reduc:, redpp:, red_p:, REDUC17:."""
    h = {"reduc": [], "redpp": [], "redpp_sum": [], "red_p": [], "red_p_sum": [], "REDUC": []}
    for line in text.splitlines():
        t = line.split()
        if not t:
            continue
        if t[0] == "reduc:":
            h["reduc"].append(float(t[4]))
        elif t[0] == "redpp:":
            h["redpp"].append(float(t[4]))
            h["redpp_sum"].append(float(t[5]))
        elif t[0] == "red_p:":
            h["red_p"].append(float(t[4]))
            h["red_p_sum"].append(float(t[5]))
        elif t[0] == "REDUC17:":
            h["REDUC"].append(float(t[3].replace("E+", "e+").replace("E-", "e-")))
    return {k: np.array(v) for k, v in h.items()}
