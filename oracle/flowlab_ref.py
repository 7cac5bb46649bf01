"""Driver of flowlab's own Fortran modules built with flang (TEST
INFRASTRUCTURE: oracle/build_ref.sh -> oracle/_ref/flowlab/flowlab_ref_driver,
source oracle/ref/flowlab_ref_driver.f90): the stream-file input and output
of the driver, used to write the golden vectors (tests/golden/make_flowlab.py)
and as the CPU baseline of the flowlab leg of bench.py.  Never used by the
product path."""
import os
import subprocess
import time

import numpy as np

DRIVER = os.path.join(os.path.dirname(os.path.abspath(__file__)), "_ref", "flowlab", "flowlab_ref_driver")


def available():
    return os.path.exists(DRIVER)


def write_input(path, m, par, n_iter, n_dump, vel_bnd):
    i4 = lambda a: np.asarray(a, "<i4").tobytes()
    f8 = lambda a: np.asarray(a, "<f8").tobytes()
    with open(path, "wb") as f:
        f.write(i4([m.n_dim, m.n_vrt, m.n_bnd, m.n_dmn, m.n_fce, m.nnz, n_iter, n_dump]))
        f.write(f8(par))
        f.write(f8(m.x))  # (n_vrt, 3) row-major = x(3, n_vrt) column-major
        f.write(i4(m.elm_vptr)); f.write(i4(m.elm_vrt))
        f.write(i4(m.bnd_tag))
        f.write(i4(m.fce_elm))  # (n_fce, 2) row-major = col(2*n_fce)
        f.write(i4(m.fce_vptr)); f.write(i4(m.fce_vrt))
        f.write(i4(m.ng_ptr)); f.write(i4(m.ng_col))
        f.write(f8(vel_bnd))  # (n_bnd, n_dim) row-major = phi%bnd(n_dim, n_bnd)


class Reader:
    def __init__(self, path):
        self.b = open(path, "rb").read()
        self.o = 0

    def f8(self, *shape):
        k = int(np.prod(shape))
        a = np.frombuffer(self.b, "<f8", k, self.o).copy()
        self.o += 8 * k
        # Fortran (a, b, c) column-major -> numpy (c, b, a)
        return a.reshape(tuple(reversed(shape))) if len(shape) > 1 else a

    def i4(self):
        v = int(np.frombuffer(self.b, "<i4", 1, self.o)[0])
        self.o += 4
        return v


def read_output(path, m, n_iter, n_dump):
    r = Reader(path)
    nd, n, nf, nb, nnz = m.n_dim, m.n_dmn, m.n_fce, m.n_bnd, m.nnz
    geo = [("x_vc", (nd, n)), ("vol", (n,)), ("x_fc", (nd, nf)), ("norm", (nd, nf)), ("area", (nf,)),
           ("x_pc", (nd, nf)), ("x_nc", (nd, nf)), ("l_pn", (nf,)), ("w_fp", (nf,)), ("ls", (nd, nd, n))]
    dump = [("pres_grad", (nd, 1, n)), ("pres_bnd", (1, nb)), ("vel_grad", (nd, nd, n)), ("mom_val", (nnz,)),
            ("mom_rhs", (nd, n)), ("mom_diag", (n,)), ("vel_bnd", (nd, nb)), ("vel_solved", (nd, n)),
            ("mom_resid", (nd, n)), ("mom_res0", (3,)), ("pp_val", (nnz,)), ("pp_rhs", (n,)), ("spd", (nf,)),
            ("pp", (1, n)), ("pp_res0", (1,)), ("pp_grad", (nd, 1, n)), ("pp_bnd", (1, nb)), ("vel", (nd, n)),
            ("pres", (1, n))]
    fin = [("vel", (nd, n)), ("pres", (1, n)), ("pp", (1, n)), ("spd", (nf,)), ("vel_bnd", (nd, nb)),
           ("pres_bnd", (1, nb))]
    sz = lambda lst: 8 * sum(int(np.prod(s)) for _, s in lst)
    # iterations run: the file holds geometry, per iteration its dumps and 5
    # doubles, then the count and the final fields
    total, k = len(r.b), None
    for kk in range(0, n_iter + 1):
        if sz(geo) + sum(sz(dump) * (it <= n_dump) + 40 for it in range(1, kk + 1)) + 4 + sz(fin) == total:
            k = kk
            break
    assert k is not None, "unexpected driver output size"
    o = {"geo_" + name: r.f8(*shp) for name, shp in geo}
    res0, gl = [], []
    for it in range(1, k + 1):
        if it <= n_dump:
            for name, shp in dump:
                a = r.f8(*shp)
                o["it%d_%s" % (it, name)] = a.reshape(a.shape[0], -1)[:, 0] if shp[0] == 1 and len(shp) == 2 else (
                    a[:, 0, :] if len(shp) == 3 and shp[1] == 1 else a)
        res0.append(r.f8(4))
        gl.append(r.f8(1)[0])
    o["n_done"] = np.int32(r.i4())
    o["res0"] = np.array(res0).reshape(-1, 4)
    o["norm_gl"] = np.array(gl)
    for name, shp in fin:
        a = r.f8(*shp)
        o["final_" + name] = a[:, 0] if shp[0] == 1 and len(shp) == 2 else a
    assert r.o == total and int(o["n_done"]) == k, (r.o, total, int(o["n_done"]), k)
    return o


def _big_stack():
    # solv_m's solvers keep their work vectors in automatic arrays
    # (solv_m.f90:48, 116: n x 6 doubles), on the stack: 42 MB at 885k cells
    import resource
    soft, hard = resource.getrlimit(resource.RLIMIT_STACK)
    want = resource.RLIM_INFINITY if hard == resource.RLIM_INFINITY else hard
    resource.setrlimit(resource.RLIMIT_STACK, (want, hard))


def run(inp, out):
    subprocess.run([DRIVER, inp, out], check=True, preexec_fn=_big_stack)


def time_outer_iterations(m, par, vel_bnd, n_iter, tmp, n_warm=2):
    """Seconds of the reference driver for n_warm + n_iter outer iterations
    minus its run with n_warm (input, mesh set-up, geometry and the first
    iterations), on the mesh m (arrays as programs_amd.flowlab.FlMesh holds
    them); also returns the n_warm run's seconds."""
    inp, out = os.path.join(tmp, "fl_time.in"), os.path.join(tmp, "fl_time.out")
    times = []
    for k in (n_warm, n_warm + n_iter):
        write_input(inp, m, par, k, -1, vel_bnd)
        t0 = time.perf_counter()
        run(inp, out)
        times.append(time.perf_counter() - t0)
    os.remove(inp)
    os.remove(out)
    return times[1] - times[0], times[0]
