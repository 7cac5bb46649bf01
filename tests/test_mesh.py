"""Host mesh layer (grd_read / grd_conn / grd_geom) of libfvk against the
reference's golden dumps, the oracle, and closed forms."""
import numpy as np
import pytest

from oracle import fvo
from tests import golden_cases as G
from tests.helpers import require_lib


def product_mesh(rec):
    fvk = require_lib()
    src = G.mesh_inputs(rec)
    if isinstance(src, str):
        return fvk.Mesh.read_neu(src)
    vx = np.asarray(src["vx"])
    return fvk.Mesh.from_cells(src["n_dx"], vx, src["v_ty"], src["v_vptr"], src["v_vrt"], src["bnd"])


@pytest.mark.parametrize("name", G.CASES)
def test_mesh_matches_reference_dump(name):
    rec = G.load(name)
    m = product_mesh(rec)
    t, g = m.topology(), m.geometry()
    for k in ("fce_ip", "fce_in", "fce_ib", "fce_bty", "fce_lst", "ng_lst"):
        assert np.array_equal(t[k], rec["d_" + k]), k
    assert m.n_bf == int(rec["d_n_bf"][0])
    for k in ("vol_x", "vol_v", "fce_x", "fce_n", "fce_dxp", "fce_dxn", "fce_s", "fce_lpn", "fce_wfp"):
        assert np.array_equal(g[k].ravel(), rec["d_" + k]), k


@pytest.mark.parametrize("args", [(6, 5, 4, 0.0, 0), (8, 8, 8, 0.2, 0x5EED0002), (9, 7, 0, 0.15, 11),
                                  (4, 4, 4, 0.3, 99)])
def test_box_generator_matches_oracle(args):
    fvk = require_lib()
    nx, ny, nz, jit, seed = args
    m = fvk.Mesh.box(nx, ny, nz, jitter=jit, seed=seed)
    o = fvo.OMesh.box(nx, ny, nz, jitter=jit, seed=seed)
    t, to = m.topology(), o.topology()
    for k in t:
        assert np.array_equal(t[k], to[k]), k
    g, go = m.geometry(), o.geometry()
    for k in g:
        assert np.array_equal(g[k], go[k]), k


def test_uniform_hex_closed_forms():
    fvk = require_lib()
    N = 6
    m = fvk.Mesh.box(N, N, N)
    g, t = m.geometry(), m.topology()
    h = 1.0 / N
    assert m.n_vol == N ** 3 and m.n_fce == 3 * N * N * (N - 1) + 6 * N * N and m.n_bf == 6 * N * N
    assert np.allclose(g["fce_s"], h * h, rtol=1e-14)
    assert np.allclose(g["vol_v"], h ** 3, rtol=1e-12)
    assert abs(g["vol_v"].sum() - 1.0) < 1e-13
    inner = t["fce_in"] > 0
    assert np.allclose(g["fce_wfp"][inner], 0.5, rtol=1e-13)
    assert np.allclose(g["fce_lpn"][inner], h, rtol=1e-13)
    assert np.allclose(g["fce_lpn"][~inner], h / 2, rtol=1e-13)
    assert np.allclose(np.abs(g["fce_n"]).sum(1), 1.0)
    # owner is the lower cell; each interior face appears in exactly two lists
    assert np.all(t["fce_ip"][inner] < t["fce_in"][inner])
    cnt = np.bincount(t["fce_lst"], minlength=m.n_fce + 1)[1:]
    assert np.array_equal(cnt, np.where(inner, 2, 1))
    # ascending face order in every cell list
    for c in range(m.n_vol):
        lst = t["fce_lst"][t["cell_ptr"][c]:t["cell_ptr"][c + 1]]
        assert np.all(np.diff(lst) > 0)


def test_jittered_volumes_close_box():
    fvk = require_lib()
    m = fvk.Mesh.box(10, 9, 8, jitter=0.2, seed=5)
    assert abs(m.geometry()["vol_v"].sum() - 1.0) < 1e-12


def test_bad_meshes_raise():
    fvk = require_lib()
    # a cell whose exterior faces carry no boundary code does not close
    vx = np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0]], float)
    with pytest.raises(fvk.FvkError):
        fvk.Mesh.from_cells(2, vx, [2], [0, 4], [1, 2, 3, 4], [31, 31, 31, 0])
    with pytest.raises(fvk.FvkError):
        fvk.Mesh.from_cells(2, vx, [9], [0, 4], [1, 2, 3, 4], [31, 31, 31, 31])
