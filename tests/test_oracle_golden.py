"""The CPU oracle against the flang-built reference's golden vectors.

Every comparison is exact (bitwise equal doubles, equal integers): the
oracle keeps the reference's loop orders and evaluation order and both are
built without FMA contraction.
"""
import numpy as np
import pytest

from oracle import fvo
from tests import golden_cases as G


@pytest.mark.parametrize("name", G.CASES)
def test_topology_and_geometry(name):
    rec = G.load(name)
    m = G.oracle_mesh(rec)
    t = m.topology()
    g = m.geometry()
    for k in ("fce_ip", "fce_in", "fce_ib", "fce_bty", "fce_lst", "ng_lst"):
        assert np.array_equal(t[k], rec["d_" + k]), k
    assert t["n_bf"] == int(rec["d_n_bf"][0])
    for k in ("vol_x", "vol_v", "vol_vr", "fce_x", "fce_n", "fce_dxp", "fce_dxn", "fce_s",
              "fce_lpn", "fce_wfp"):
        assert np.array_equal(g[k].ravel(), rec["d_" + k]), k
    m.free()


@pytest.mark.parametrize("name", G.CASES)
def test_literal_connectivity_equals_hashed(name):
    rec = G.load(name)
    a = G.oracle_mesh(rec, conn="literal").topology()
    b = G.oracle_mesh(rec, conn="hash").topology()
    for k in ("fce_ip", "fce_in", "fce_ib", "fce_lst", "ng_lst"):
        assert np.array_equal(a[k], b[k]), k


@pytest.mark.parametrize("name", G.CASES)
def test_gradient(name):
    rec = G.load(name)
    m = G.oracle_mesh(rec)
    g = m.geometry()
    nv = m.m.n_vol
    phi_c = np.prod(g["vol_x"], axis=1)
    ib = m.topology()["fce_ib"]
    bmask = m.topology()["fce_in"] == 0
    phi_b = np.zeros(m.m.n_bf)
    phi_b[ib[bmask] - nv - 1] = np.prod(g["fce_x"][bmask], axis=1)
    for k in (1, 2, 3):
        for neu in (0, 2):
            got = m.grad(phi_c, phi_b, k, neu)
            # reference layout grad(nv, 1, n_dx), column-major
            ref = rec["d_grad%d_%d" % (neu, k)].reshape(m.m.n_dx, nv).T
            assert np.array_equal(got, ref, equal_nan=True), (k, neu)


@pytest.mark.parametrize("name", G.CASES)
def test_simple_outer_iterations(name):
    rec = G.load(name)
    spec = rec["spec"]
    m = G.oracle_mesh(rec)
    f = fvo.OFlow(m, G.cfg_for(rec))
    for code, u in spec["bc"]:
        f.set_bnd_u(code, u)
    H = np.array([f.outer(it) for it in range(1, spec["n_outer"] + 1)])
    assert np.array_equal(H[:, 0], rec["h_reduc"], equal_nan=True)
    assert np.array_equal(H[:, 1], rec["h_redpp"], equal_nan=True)
    assert np.array_equal(H[:, 4], rec["h_redpp_sum"], equal_nan=True)
    assert np.array_equal(H[:, 3], rec["d_hist_REDUC"], equal_nan=True)
    if spec["variant"] == "shipped":
        assert np.array_equal(H[:, 2], rec["h_red_p"], equal_nan=True)
    F = f.fields()
    pairs = [("u_cell", "u_cell"), ("u_bnd", "u_bnd"), ("p", "p"), ("dpdx", "dpdx"), ("spd", "spd"),
             ("frc_dns", "frc_dns"), ("cs", "cs"), ("pl", "pp_pl"), ("ng_p", "pp_ngp")]
    for a, b in pairs:
        assert np.array_equal(F[a].ravel(), rec["d_" + b]), a


def _mom_sys(rec, m):
    return fvo.OSys(m, rec["d_mom_pl"], rec["d_mom_ngp"], rec["d_mom_ngn"])


@pytest.mark.parametrize("name", G.CASES)
@pytest.mark.parametrize("imat", [0, 1, 2])
def test_standalone_solvers(name, imat):
    rec = G.load(name)
    m = G.oracle_mesh(rec)
    nv, nd = m.m.n_vol, m.m.n_dx
    i = np.arange(1, nv + 1)
    b3 = np.stack([(np.mod(i * 7919 + c * 104729, 1013)) / 1013.0 - 0.5 for c in range(1, nd + 1)], 1)
    A = _mom_sys(rec, m)
    c, cr = A.precond(imat)
    assert np.array_equal(c, rec["d_mom_c%d" % imat], equal_nan=True)
    assert np.array_equal(cr, rec["d_mom_cr%d" % imat], equal_nan=True)
    assert np.array_equal(A.direct(imat, b3[:, 0]), rec["d_mom_dir%d" % imat], equal_nan=True)
    x, r, rn, its, _ = A.solve("bicg", imat, 1e-30, 6, b3)
    assert np.array_equal(x.ravel(), rec["d_mom_bicg_x%d" % imat], equal_nan=True)
    assert np.array_equal(r.T.ravel(), rec["d_mom_bicg_r%d" % imat], equal_nan=True)
    assert np.array_equal(rn, rec["d_mom_bicg_n%d" % imat], equal_nan=True)
    x, *_ = A.solve("bicg", imat, 0.1, 5, b3)
    assert np.array_equal(x.ravel(), rec["d_mom_bicgr_x%d" % imat], equal_nan=True)
    P = fvo.OSys(m, rec["d_pp_pl"], rec["d_pp_ngp"], rec["d_pp_ngp"])
    bb = b3[:, 0] - np.cumsum(b3[:, 0])[-1] / nv  # sequential sum, as Fortran sum()
    c, _ = P.precond(imat)
    assert np.array_equal(c, rec["d_pp_c%d" % imat], equal_nan=True)
    assert np.array_equal(P.direct(imat, bb), rec["d_pp_dir%d" % imat], equal_nan=True)
    x, r, rn, its, _ = P.solve("cg", imat, 1e-30, 15, bb)
    assert np.array_equal(x[:, 0], rec["d_pp_cg_x%d" % imat], equal_nan=True)
    assert np.array_equal(rn, rec["d_pp_cg_n%d" % imat], equal_nan=True)
    x, *_ = P.solve("cg", imat, 0.1, 20, bb, not_r0=True)
    assert np.array_equal(x[:, 0], rec["d_pp_cgr_x%d" % imat], equal_nan=True)
