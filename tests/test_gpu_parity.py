"""HIP path (through the C ABI) against the flang-built reference's golden
vectors and the CPU oracle.

Bars (stated per assertion):
  * exact (bitwise) for integer work and for every row-local floating-point
    kernel: factor (solv_precond), triangular sweeps (solv_direct), residual
    (solv_res), Green-Gauss gradients (solv_grad);
  * 1e-10 relative (max-norm) for anything downstream of a dot product or L1
    norm (PCG, BiCGStab, SIMPLE fields and residual histories): the device
    reduces in a fixed tree, the reference sequentially.  Where the reference
    algorithm itself is ill-conditioned, the bar widens to 3x the change the
    oracle shows when only its own summation order is switched from
    sequential to pairwise (sens_tol), capped per case (WIDEN_CAP); that
    change and the error reached are printed.
"""
import numpy as np
import pytest

from oracle import fvo
from tests import golden_cases as G
from tests.helpers import assert_rel, require_lib

pytestmark = pytest.mark.gpu

TOL = 1e-10

# The only solves of the golden cases whose reference algorithm moves by more
# than TOL/3 when its summation order changes: BiCGStab preconditioned with
# DIC, a factor for symmetric matrices, on the non-symmetric momentum matrix
# of the 2D channel, a combination the reference does not ship (its momentum
# solves use DILU, cfd_setup.f90:528).  The oracle moves by 6.8e-4 after 6
# iterations and by 1.8e-7 after the 5 iterations with reduc 0.1.  The
# widened bar of every other solve is capped at TOL.
# Beyond the golden cases, SIMPLE on the 32^3 cavity over 200 outer iterations
# (configuration C1, relaxed) moves by 2.5e-5 with the summation order.
WIDEN_CAP = {("chan2d_shipped", "bicg6", 1): 2.5e-3, ("chan2d_shipped", "bicg5r", 1): 1e-6,
             ("box32", "simple200", 0): 1e-4}


def sens_tol(run, key):
    """max(TOL, 3 * |run(sequential) - run(pairwise)| / |run(sequential)|)
    for a callable returning a flat array from the oracle, capped at
    WIDEN_CAP.get(key, TOL)."""
    a = np.asarray(run(), np.float64).ravel()
    with fvo.sum_order(1):
        b = np.asarray(run(), np.float64).ravel()
    den = np.max(np.abs(a)) if a.size else 0.0
    sens = float(np.max(np.abs(a - b)) / den) if den > 0 else 0.0
    tol = max(TOL, 3.0 * sens)
    cap = WIDEN_CAP.get(key, TOL)
    if 3.0 * sens > TOL:
        print("%s: order sensitivity of the reference algorithm %.2e, bar %.2e (cap %.1e)" % (key, sens, tol, cap))
    assert tol <= cap, "%s: the reference algorithm's order sensitivity %.2e exceeds the cap %.1e" % (key, sens, cap)
    return tol


def check_rel(x, ref, tol, what):
    """assert_rel, printing the error reached where the bar is widened."""
    if tol > TOL:
        from tests.helpers import relerr
        print("%s: relative error %.2e against the widened bar %.2e" % (what, relerr(x, ref), tol))
    assert_rel(x, ref, tol, what)


def product_mesh(rec):
    fvk = require_lib()
    src = G.mesh_inputs(rec)
    if isinstance(src, str):
        return fvk.Mesh.read_neu(src)
    return fvk.Mesh.from_cells(src["n_dx"], np.asarray(src["vx"]), src["v_ty"], src["v_vptr"], src["v_vrt"],
                               src["bnd"])


def b3_for(nv, nd):
    i = np.arange(1, nv + 1)
    return np.stack([np.mod(i * 7919 + c * 104729, 1013) / 1013.0 - 0.5 for c in range(1, nd + 1)], 1)


@pytest.mark.parametrize("rcm", [False, True])
@pytest.mark.parametrize("name", G.CASES)
def test_solvers_match_reference(name, rcm):
    fvk = require_lib()
    rec = G.load(name)
    m = product_mesh(rec)
    ctx = fvk.grd_solv(m, rcm=rcm)
    nv, nd = m.n_vol, m.n_dx
    b3 = b3_for(nv, nd)
    fvk.set_matrix(ctx, rec["d_mom_pl"], rec["d_mom_ngp"], rec["d_mom_ngn"])
    om = G.oracle_mesh(rec)
    A = fvo.OSys(om, rec["d_mom_pl"], rec["d_mom_ngp"], rec["d_mom_ngn"])
    for imat in (0, 1, 2):
        tol6 = sens_tol(lambda: A.solve("bicg", imat, 1e-30, 6, b3)[0], (name, "bicg6", imat))
        tol5 = sens_tol(lambda: A.solve("bicg", imat, 0.1, 5, b3)[0], (name, "bicg5r", imat))
        c, cr = fvk.solv_precond(ctx, imat)
        assert np.array_equal(c, rec["d_mom_c%d" % imat]), ("precond c exact", imat)
        assert np.array_equal(cr, rec["d_mom_cr%d" % imat]), ("precond c_r exact", imat)
        assert np.array_equal(fvk.solv_direct(ctx, imat, b3[:, 0]), rec["d_mom_dir%d" % imat]), ("direct exact", imat)
        x = np.zeros((nv, nd))
        r = b3.copy()
        rn, st = fvk.solv_bicg(ctx, imat, 1e-30, 6, x, r)
        check_rel(x.ravel(), rec["d_mom_bicg_x%d" % imat], tol6, "bicg x imat %d" % imat)
        assert_rel(r.T.ravel(), rec["d_mom_bicg_r%d" % imat], 0.0, "bicg residual (solv_res, exact)")
        assert_rel(rn, rec["d_mom_bicg_n%d" % imat], TOL, "bicg res0_n")
        x = np.zeros((nv, nd))
        r = b3.copy()
        fvk.solv_bicg(ctx, imat, 0.1, 5, x, r)
        check_rel(x.ravel(), rec["d_mom_bicgr_x%d" % imat], tol5, "bicg reduc=0.1 x imat %d" % imat)
    fvk.set_matrix(ctx, rec["d_pp_pl"], rec["d_pp_ngp"], rec["d_pp_ngp"])
    P = fvo.OSys(om, rec["d_pp_pl"], rec["d_pp_ngp"], rec["d_pp_ngp"])
    bb = b3[:, 0] - np.cumsum(b3[:, 0])[-1] / nv
    for imat in (0, 1, 2):
        tolc = sens_tol(lambda: P.solve("cg", imat, 1e-30, 15, bb)[0], (name, "cg15", imat))
        tolr = sens_tol(lambda: P.solve("cg", imat, 0.1, 20, bb, not_r0=True)[0], (name, "cg20r", imat))
        c, _ = fvk.solv_precond(ctx, imat)
        assert np.array_equal(c, rec["d_pp_c%d" % imat])
        assert np.array_equal(fvk.solv_direct(ctx, imat, bb), rec["d_pp_dir%d" % imat])
        x = np.zeros((nv, 1))
        r = bb.reshape(-1, 1).copy()
        rn, st = fvk.solv_cg(ctx, imat, 1e-30, 15, x, r)
        assert st.iters[0] == 15
        check_rel(x[:, 0], rec["d_pp_cg_x%d" % imat], tolc, "cg x imat %d" % imat)
        x = np.zeros((nv, 1))
        r = bb.reshape(-1, 1).copy()
        fvk.solv_cg(ctx, imat, 0.1, 20, x, r, not_r0=True)
        check_rel(x[:, 0], rec["d_pp_cgr_x%d" % imat], tolr, "cg reduc=0.1 x imat %d" % imat)
    ctx.destroy()


@pytest.mark.parametrize("name", G.CASES)
def test_residual_exact_vs_oracle(name):
    fvk = require_lib()
    rec = G.load(name)
    m = product_mesh(rec)
    om = G.oracle_mesh(rec)
    ctx = fvk.grd_solv(m)
    fvk.set_matrix(ctx, rec["d_mom_pl"], rec["d_mom_ngp"], rec["d_mom_ngn"])
    A = fvo.OSys(om, rec["d_mom_pl"], rec["d_mom_ngp"], rec["d_mom_ngn"])
    nv, nd = m.n_vol, m.n_dx
    phi = np.sin(np.arange(nv * nd)).reshape(nv, nd)
    b = b3_for(nv, nd)
    r = b.copy()
    rn = fvk.solv_res(ctx, phi, r)
    # oracle solv_res through a zero-iteration solve (same routine)
    _, rref, rnref, _, _ = A.solve("cg", 0, 0.1, 0, b, x0=phi)
    assert np.array_equal(r, rref)
    assert_rel(rn, rnref, 1e-14, "L1 norms")
    y = A.spmv(phi[:, 0])
    r1 = np.zeros((nv, 1))
    fvk.solv_res(ctx, phi[:, :1], r1)
    assert np.array_equal(-r1[:, 0], y)


@pytest.mark.parametrize("rcm", [False, True])
@pytest.mark.parametrize("name", G.CASES)
def test_gradient_exact(name, rcm):
    fvk = require_lib()
    rec = G.load(name)
    m = product_mesh(rec)
    ctx = fvk.grd_solv(m, rcm=rcm)
    g, t = m.geometry(), m.topology()
    nv = m.n_vol
    phi = np.zeros(nv + m.n_bf)
    phi[:nv] = np.prod(g["vol_x"], axis=1)
    bm = t["fce_in"] == 0
    phi[t["fce_ib"][bm] - 1] = np.prod(g["fce_x"][bm], axis=1)
    for k in (1, 2, 3):
        for neu in (0, 2):
            got = fvk.solv_grad(ctx, phi, k, neu)[:, 0, :]
            ref = rec["d_grad%d_%d" % (neu, k)].reshape(m.n_dx, nv).T
            assert np.array_equal(got, ref), (k, neu)


def run_flow(fvk, m, spec, cfg, rcm=False):
    ctx = fvk.grd_solv(m, rcm=rcm)
    fl = fvk.Flow(ctx, cfg)
    for code, u in spec["bc"]:
        fl.set_bnd_u(code, u)
    hist = [fl.outer(it) for it in range(1, spec["n_outer"] + 1)]
    return ctx, fl, hist


def flow_cfg_for(fvk, rec):
    spec = rec["spec"]
    if spec["variant"] == "shipped":
        return fvk.flow_cfg(rho=spec["rho"], mu=spec["mu"], urf_p=spec["urf_p"])
    return fvk.flow_cfg(simpler=0, imat_p=0, rho=spec["rho"], mu=spec["mu"], urf_p=spec["urf_p"])


def oracle_flow_fields(rec):
    spec = rec["spec"]
    om = G.oracle_mesh(rec)
    of = fvo.OFlow(om, G.cfg_for(rec))
    for code, u in spec["bc"]:
        of.set_bnd_u(code, u)
    H = np.array([of.outer(it) for it in range(1, spec["n_outer"] + 1)])
    F = of.fields()
    return np.concatenate([H[:, :4].ravel()] + [F[k].ravel() for k in ("u_cell", "p", "dpdx", "spd")])


@pytest.mark.parametrize("name", G.CASES)
def test_simple_matches_reference(name):
    fvk = require_lib()
    rec = G.load(name)
    spec = rec["spec"]
    tol = sens_tol(lambda: oracle_flow_fields(rec), (name, "simple", 0))
    m = product_mesh(rec)
    ctx, fl, hist = run_flow(fvk, m, spec, flow_cfg_for(fvk, rec))
    H = {k: np.array([h[k] for h in hist]) for k in ("reduc", "redpp", "red_p", "mx_res_n", "redpp_sum")}
    assert_rel(H["reduc"], rec["h_reduc"], tol, "reduc: history")
    assert_rel(H["redpp"], rec["h_redpp"], tol, "redpp: history")
    assert_rel(H["mx_res_n"], rec["d_hist_REDUC"], tol, "REDUC: history")
    if spec["variant"] == "shipped":
        assert_rel(H["red_p"], rec["h_red_p"], tol, "red_p: history")
    s = fl.state()
    pairs = [("u_cell", "u_cell"), ("u_bnd", "u_bnd"), ("p", "p"), ("dpdx", "dpdx"), ("spd", "spd"),
             ("cs", "cs"), ("frc_dns", "frc_dns")]
    for a, b in pairs:
        assert_rel(s[a].ravel(), rec["d_" + b], tol, a)
    pl, gp, _ = fl.matrix()
    assert_rel(pl, rec["d_pp_pl"], tol, "pp pl")
    assert_rel(gp, rec["d_pp_ngp"], tol, "pp ng%p")
    ctx.destroy()


@pytest.mark.parametrize("n,jit,variant,n_outer", [(16, 0.1, "relaxed", 25), (12, 0.0, "shipped", 8),
                                                   (20, 0.0, "relaxed", 20)])
def test_simple_matches_oracle_box(n, jit, variant, n_outer):
    fvk = require_lib()
    kw = dict(simpler=0, imat_p=0, urf_p=0.3) if variant == "relaxed" else dict(urf_p=0.1)
    o = fvo.OMesh.box(n, n, n, jitter=jit, seed=0x5EED0002)
    of = fvo.OFlow(o, fvo.default_cfg(**kw))
    of.set_bnd_u(31, [1, 0, 0])
    H = np.array([of.outer(it) for it in range(1, n_outer + 1)])
    F = of.fields()
    m = fvk.Mesh.box(n, n, n, jitter=jit, seed=0x5EED0002)
    for rcm in (False, True):
        ctx, fl, hist = run_flow(fvk, m, dict(bc=[(31, (1, 0, 0))], n_outer=n_outer), fvk.flow_cfg(**kw), rcm=rcm)
        assert_rel([h["reduc"] for h in hist], H[:, 0], TOL, "reduc history")
        assert_rel([h["mx_res_n"] for h in hist], H[:, 3], TOL, "REDUC history")
        s = fl.state()
        for k in ("u_cell", "p", "dpdx", "spd"):
            assert_rel(s[k].ravel(), F[k].ravel(), TOL, k)
        ctx.destroy()


def test_context_uses_the_rcm_order():
    """A context built with rcm=True puts the cells in the order
    fvk_mesh_rcm gives (pinned on the host by tests/test_rcm.py); without
    it, in the reference's order."""
    fvk = require_lib()
    m = fvk.Mesh.box(9, 9, 9, jitter=0.1, seed=0x5EED0073)
    c1 = fvk.grd_solv(m, rcm=True)
    c2 = fvk.grd_solv(m)
    assert np.array_equal(c1.perm(), m.rcm())
    assert np.array_equal(c2.perm(), np.arange(1, m.n_vol + 1))
    c1.destroy()
    c2.destroy()
