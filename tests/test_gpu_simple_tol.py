"""SIMPLE on the GPU against the CPU oracle at the configurations' own sizes
and with pressure solves to a tolerance.

  * 32^3 uniform lid-driven cavity, Re = 100, relaxed pressure update
    (urf 0.7 / 0.3), BiCGStab momentum and Jacobi-PCG pressure correction,
    FP64, 200 outer iterations: residual histories and fields.
  * 128^3 cavity with every interior vertex jittered by +-2% of the spacing
    (the +-20% mesh gives negative cell volumes in the reference's volume
    formula), pressure solves to a relative L1 residual of 1e-8 from x = 0
    (p_solver 3), 3 outer iterations.
  * p_solver 3 on the golden meshes against the oracle's p_solver 3
    (fvo_pcg_tol: the reference's solv_cg with reduc = tol, pinned bit for bit
    against the flang build in tests/test_ref_cg.py).
  * FP32-inner (p_solver 1) and learned (p_solver 2) pressure solves at a
    tight tolerance against the oracle's p_solver 3.

Bars: 1e-10 relative (max norm), or where the reference algorithm itself is
sensitive to the order of its sums, 30x the change the oracle shows when
only that order changes (sens_tol, printed).  Solves to a tolerance by
other methods (p_solver 1, 2) agree to what the tolerance fixes (stated at
the assertion).
"""
import time

import numpy as np
import pytest

from oracle import fvo
from tests import golden_cases as G
from tests.helpers import assert_rel, relerr, require_lib
from tests.test_gpu_parity import TOL, product_mesh, sens_tol

pytestmark = pytest.mark.gpu

RELAXED = dict(simpler=0, imat_p=0, urf_p=0.3)
FIELDS = ("u_cell", "p", "dpdx", "spd")


def oracle_run(om, kw, n_outer, bc=((31, (1.0, 0.0, 0.0)),)):
    of = fvo.OFlow(om, fvo.default_cfg(**kw))
    for code, u in bc:
        of.set_bnd_u(code, u)
    H = np.array([of.outer(it) for it in range(1, n_outer + 1)])
    F = of.fields()
    of.free()
    return H, F


def flat(H, F):
    return np.concatenate([H[:, :4].ravel()] + [F[k].ravel() for k in FIELDS])


def gpu_run(fvk, m, kw, n_outer, bc=((31, (1.0, 0.0, 0.0)),), rcm=False, graph=False):
    ctx = fvk.grd_solv(m, rcm=rcm)
    fl = fvk.Flow(ctx, fvk.flow_cfg(**kw))
    for code, u in bc:
        fl.set_bnd_u(code, u)
    hist = fl.run_graph(1, n_outer) if graph else fl.run(1, n_outer)
    s = fl.state()
    fl.destroy()
    ctx.destroy()
    return hist, s


def compare(hist, s, H, F, tol, what):
    errs = {k: relerr(s[k].ravel(), F[k].ravel()) for k in FIELDS}
    errs["reduc"] = relerr([h["reduc"] for h in hist], H[:, 0])
    errs["REDUC"] = relerr([h["mx_res_n"] for h in hist], H[:, 3])
    print(what, "max rel. error vs oracle:", " ".join("%s %.2e" % kv for kv in errs.items()), "bar %.1e" % tol,
          flush=True)
    assert_rel([h["reduc"] for h in hist], H[:, 0], tol, what + " reduc: history")
    assert_rel([h["mx_res_n"] for h in hist], H[:, 3], tol, what + " REDUC: history")
    for k in FIELDS:
        assert_rel(s[k].ravel(), F[k].ravel(), tol, what + " " + k)
    return errs


@pytest.mark.timeout(300)
def test_cavity_32cube_200_outer_iterations():
    fvk = require_lib()
    n, n_outer = 32, 200
    om = fvo.OMesh.box(n, n, n)
    try:
        t0 = time.time()
        tol = sens_tol(lambda: flat(*oracle_run(om, RELAXED, n_outer)), ("box32", "simple200", 0))
        H, F = oracle_run(om, RELAXED, n_outer)
        print("oracle: %.1f s for 3 runs of %d outer iterations" % (time.time() - t0, n_outer), flush=True)
    finally:
        om.free()
    m = fvk.Mesh.box(n, n, n)
    hist, s = gpu_run(fvk, m, RELAXED, n_outer, graph=True)
    compare(hist, s, H, F, tol, "32^3 x 200")
    # the flow has converged a long way: the last normalised residual is small
    assert hist[-1]["mx_res_n"] < 1e-2


@pytest.mark.timeout(300)
def test_cavity_32cube_200_outer_iterations_pressure_to_tolerance():
    """The same 200 iterations with every pressure correction solved to a
    relative L1 residual of 1e-10 (p_solver 3): without the reference's
    20-iteration, reduc-0.1 exits (whose discrete stops carry rounding
    differences into the trajectory), GPU and oracle agree to 1e-10."""
    fvk = require_lib()
    n, n_outer = 32, 200
    kw = dict(RELAXED, p_solver=3, p_tol=1e-10)
    om = fvo.OMesh.box(n, n, n)
    try:
        t0 = time.time()
        H, F = oracle_run(om, kw, n_outer)
        print("oracle: %.1f s" % (time.time() - t0), flush=True)
    finally:
        om.free()
    m = fvk.Mesh.box(n, n, n)
    hist, s = gpu_run(fvk, m, {k: v for k, v in kw.items()}, n_outer, graph=True)
    compare(hist, s, H, F, TOL, "32^3 x 200 p_tol 1e-10")


@pytest.mark.timeout(600)
def test_cavity_128cube_pressure_to_tolerance_3_outer():
    fvk = require_lib()
    n, n_outer, jit = 128, 3, 0.02
    kw = dict(RELAXED, p_solver=3, p_tol=1e-8)
    t0 = time.time()
    om = fvo.OMesh.box(n, n, n, jitter=jit, seed=0x5EED0002)
    try:
        H, F = oracle_run(om, kw, n_outer)
    finally:
        om.free()
    print("oracle: %.1f s" % (time.time() - t0), flush=True)
    m = fvk.Mesh.box(n, n, n, jitter=jit, seed=0x5EED0002)
    hist, s = gpu_run(fvk, m, kw, n_outer)
    assert [h["it_pp"] for h in hist][1:] and min(h["it_pp"] for h in hist[1:]) > 100
    compare(hist, s, H, F, TOL, "128^3 p_tol 1e-8")


@pytest.mark.parametrize("name", G.CASES)
def test_pressure_to_tolerance_matches_oracle(name):
    fvk = require_lib()
    rec = G.load(name)
    spec = rec["spec"]
    cfg = G.cfg_for(rec)
    kw = {k: getattr(cfg, k) for k, _ in fvo.Cfg._fields_}
    kw.update(p_solver=3, p_tol=1e-10)
    bc = [(c, tuple(u)) for c, u in spec["bc"]]
    om = G.oracle_mesh(rec)
    try:
        run = lambda: flat(*oracle_run(om, kw, spec["n_outer"], bc))  # noqa: E731
        tol = sens_tol(run, (name, "p_tol", 0))
        H, F = oracle_run(om, kw, spec["n_outer"], bc)
    finally:
        om.free()
    gkw = {k: v for k, v in kw.items() if k != "p_cap"}
    for graph in (False, True):
        hist, s = gpu_run(fvk, product_mesh(rec), gkw, spec["n_outer"], bc, graph=graph)
        compare(hist, s, H, F, tol, "%s p_solver 3%s" % (name, " graph" if graph else ""))


@pytest.mark.parametrize("p_solver", [1, 2])
def test_mixed_and_learned_pressure_solves_match_oracle(p_solver):
    """The FP32-inner defect correction and the learned factor solve the
    same systems to the same FP64 tolerance as the oracle's plain PCG: at
    p_tol 1e-12 the fields after 10 outer iterations agree to 1e-9 (each
    solve leaves an error of at most about cond * 1e-12 in p')."""
    fvk = require_lib()
    n, n_outer, jit = 16, 10, 0.02
    kw = dict(RELAXED, p_solver=3, p_tol=1e-12)
    om = fvo.OMesh.box(n, n, n, jitter=jit, seed=0x5EED0002)
    try:
        H, F = oracle_run(om, kw, n_outer)
    finally:
        om.free()
    m = fvk.Mesh.box(n, n, n, jitter=jit, seed=0x5EED0002)
    ctx = fvk.grd_solv(m)
    fl = fvk.Flow(ctx, fvk.flow_cfg(**dict(kw, p_solver=p_solver)))
    model = None
    if p_solver == 2:
        model = fvk.Model(seed=11)
        fl.set_model(model, train_steps=2, lr=3e-3)
    fl.set_bnd_u(31, [1.0, 0.0, 0.0])
    hist = fl.run(1, n_outer)
    s = fl.state()
    compare(hist, s, H, F, 1e-9, "16^3 p_solver %d p_tol 1e-12" % p_solver)
    fl.destroy()
    if model is not None:
        model.destroy()
    ctx.destroy()
