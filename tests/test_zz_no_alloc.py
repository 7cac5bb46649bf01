"""fvk_solv_res_dev, the entry point that allocates nothing (its reduction
partials live in a static device array per GPU), against fvk_solv_res (the
same kernels with the context's own scratch), through the C ABI: residual
and L1 norms bit for bit, on two calls in a row on one GPU, the second on a
smaller mesh with fewer components, both with and without not_r0.  Device
memory for the inputs and outputs comes from the HIP runtime the library
itself is linked with (hipMalloc / hipMemcpy through ctypes), as a C caller
would provide it."""
import ctypes as C

import numpy as np
import pytest

from tests.helpers import require_lib

pytestmark = pytest.mark.gpu

H2D, D2H = 1, 2


def hip():
    fvk = require_lib()
    fvk.lib()  # loads libamdhip64.so.7 with the library
    h = C.CDLL("libamdhip64.so.7")
    h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    h.hipFree.argtypes = [C.c_void_p]
    h.hipSetDevice.argtypes = [C.c_int]
    for f in (h.hipMalloc, h.hipMemcpy, h.hipFree, h.hipSetDevice):
        f.restype = C.c_int
    return h


class DevArray:
    def __init__(self, h, a):
        self.h = h
        self.a = np.ascontiguousarray(a, np.float64)
        self.p = C.c_void_p()
        assert h.hipMalloc(C.byref(self.p), self.a.nbytes) == 0
        assert h.hipMemcpy(self.p, self.a.ctypes.data, self.a.nbytes, H2D) == 0

    def get(self):
        out = np.empty_like(self.a)
        assert self.h.hipMemcpy(out.ctypes.data, self.p, out.nbytes, D2H) == 0
        return out

    def free(self):
        self.h.hipFree(self.p)


def one_call(fvk, h, n, n_vc, seed, not_r0):
    m = fvk.Mesh.box(n, n, n, jitter=0.2, seed=seed)
    ctx = fvk.grd_solv(m)  # reference row order: device rows = cells
    assert h.hipSetDevice(0) == 0
    rng = np.random.default_rng(seed)
    ng = rng.standard_normal(m.n_fce)
    pl = 6.0 + rng.random(m.n_vol)
    fvk.set_matrix(ctx, pl, ng, ng * 0.5)
    phi = rng.standard_normal((m.n_vol, n_vc))
    b = rng.standard_normal((m.n_vol, n_vc))
    # host path: fvk_solv_res
    r_host = b.copy()
    rn_host = fvk.solv_res(ctx, None if not_r0 else phi, r_host, not_r0=not_r0)
    # device path: component-major device arrays in row order
    dphi, dr = DevArray(h, phi.T), DevArray(h, b.T)
    drn = DevArray(h, np.full(n_vc, np.nan))
    fvk.solv_res_dev(ctx, n_vc, dphi.p.value, dr.p.value, drn.p.value, not_r0=not_r0)
    ctx.sync()
    r_dev = dr.get().T
    rn_dev = drn.get()
    print("n", n, "n_vc", n_vc, "not_r0", not_r0, "res0_n", rn_dev, flush=True)
    assert np.array_equal(r_dev, r_host)
    assert np.array_equal(rn_dev, rn_host)
    assert np.all(np.isfinite(rn_dev)) and np.all(rn_dev > 0)
    for a in (dphi, dr, drn):
        a.free()
    ctx.destroy()
    m.free()


@pytest.mark.parametrize("not_r0", [False, True])
def test_no_alloc_residual_matches_solv_res(not_r0):
    fvk = require_lib()
    h = hip()
    one_call(fvk, h, 40, 3, 0x5EED00A1, not_r0)
    one_call(fvk, h, 9, 1, 0x5EED00A2, not_r0)


def test_no_alloc_rejects_bad_sizes():
    fvk = require_lib()
    m = fvk.Mesh.box(4, 4, 4)
    ctx = fvk.grd_solv(m)
    with pytest.raises(fvk.FvkError) as e:
        fvk.solv_res_dev(ctx, 4, 1, 1, 1)
    assert e.value.status == -1 and "n_vc" in str(e.value)
    ctx.destroy()
