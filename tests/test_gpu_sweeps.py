"""DIC / DILU factor and triangular sweeps in one launch each against the
level-by-level launches and the oracle.  One launch: blocks owning ranges of
rows consecutive in the reference numbering, walked level by level with the
values inside a range passing through LDS (the default; ranges of 4096 rows,
and of 256 rows so that each block takes many ranges), or rows taken in level
order in chunks; rows wait for the values they need from elsewhere.

Bar: bitwise.  Both forms perform the same operations in the same order for
every row (solv_precond.f90:10-29, solv_direct.f90:11-42); only the
scheduling differs.  Cases: jittered hex boxes up to 96^3 (884,736 rows,
286 levels per sweep) in natural and reverse Cuthill-McKee order, one and
three right-hand sides, the non-symmetric DILU matrix, PCG / BiCGStab
solves built on them, and SIMPLE outer iterations in the reference's
shipped configuration (SIMPLER, DILU momentum, DIC pressure), eager and
from HIP graphs.  The wait-bound failure path is forced from the host and
must end every launch and be reported.
"""
import sys

import numpy as np
import pytest

from oracle import fvo
from tests.helpers import require_lib

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def random_matrix(m, seed):
    """A diagonally dominant non-symmetric face matrix on the mesh."""
    rng = np.random.default_rng(seed)
    nf = m.n_fce
    ngp = -rng.uniform(0.5, 1.5, nf)
    ngn = -rng.uniform(0.5, 1.5, nf)
    topo = m.topology()
    ip, inn = topo["fce_ip"], topo["fce_in"]
    pl = np.full(m.n_vol, 1.0)
    it = inn > 0
    np.add.at(pl, ip[it] - 1, np.abs(ngp[it]))
    np.add.at(pl, inn[it] - 1, np.abs(ngn[it]))
    return pl, ngp, ngn


MODES = [
    ("one launch per level", {"sweep_one_launch": 0}),
    ("chunks in level order", {"sweep_one_launch": 1, "sweep_parts": 0}),
    ("ranges of 4096 rows", {"sweep_one_launch": 1, "sweep_parts": 1, "sweep_part_rows": 4096}),
    ("ranges of 256 rows", {"sweep_one_launch": 1, "sweep_parts": 1, "sweep_part_rows": 256}),
]


def set_mode(fvk, ctx, params):
    for k, v in params.items():
        fvk.set_param(ctx, k, v)


def all_modes(fvk, ctx, fn):
    out = []
    for label, params in MODES:
        set_mode(fvk, ctx, params)
        out.append((label, fn()))
    set_mode(fvk, ctx, MODES[2][1])
    return out


@pytest.mark.parametrize("n,rcm", [(32, False), (32, True), (96, False)])
def test_factor_and_sweeps_one_launch_bitwise(n, rcm):
    fvk = require_lib()
    say("sweeps: box %d^3 rcm %s" % (n, rcm))
    m = fvk.Mesh.box(n, n, n, jitter=0.05, seed=0x5EED0100 + n)
    ctx = fvk.grd_solv(m, rcm=rcm)
    pl, ngp, ngn = random_matrix(m, 7 + n)
    fvk.set_matrix(ctx, pl, ngp, ngn)
    b = np.random.default_rng(n).standard_normal(m.n_vol)
    for imat in (1, 2):
        f = all_modes(fvk, ctx, lambda: fvk.solv_precond(ctx, imat))
        c0, cr0 = f[0][1]
        for label, (c1, cr1) in f[1:]:
            assert np.array_equal(c0, c1) and np.array_equal(cr0, cr1), ("factor", imat, label)
        x = all_modes(fvk, ctx, lambda: fvk.solv_direct(ctx, imat, b))
        for label, x1 in x[1:]:
            assert np.array_equal(x[0][1], x1), ("sweeps", imat, label)
        assert np.all(np.isfinite(x[0][1]))
        say("  imat %d: factor and sweeps bitwise, levels %s" % (imat, ctx.levels()))
    if n <= 32:
        A = fvo.OSys(fvo.OMesh.box(n, n, n, jitter=0.05, seed=0x5EED0100 + n), pl, ngp, ngn)
        for imat in (1, 2):
            assert np.array_equal(A.direct(imat, b), fvk.solv_direct(ctx, imat, b)), ("oracle", imat)
    ctx.destroy()
    m.free()


@pytest.mark.parametrize("rcm", [False, True])
def test_sweeps_on_tetrahedra_bitwise(rcm, tmp_path):
    """flowlab's duct_3d.msh read as a spray-simulation mesh (24,258 cells,
    up to 5 lower neighbours a row: the kernels that hold 8 slots) in every
    form against one launch per level."""
    import gzip
    import os
    import shutil
    fvk = require_lib()
    golden = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
    p = os.path.join(str(tmp_path), "duct_3d.msh")
    with gzip.open(os.path.join(golden, "duct_3d.msh.gz")) as a, open(p, "wb") as b:
        shutil.copyfileobj(a, b)
    say("sweeps: duct_3d rcm %s" % rcm)
    m = fvk.Mesh.read_gmsh(p, 3)
    ctx = fvk.grd_solv(m, rcm=rcm)
    pl, ngp, ngn = random_matrix(m, 17)
    fvk.set_matrix(ctx, pl, ngp, ngn)
    b = np.random.default_rng(5).standard_normal(m.n_vol)
    for imat in (1, 2):
        f = all_modes(fvk, ctx, lambda: fvk.solv_precond(ctx, imat))
        for label, (c1, cr1) in f[1:]:
            assert np.array_equal(f[0][1][0], c1) and np.array_equal(f[0][1][1], cr1), ("factor", imat, label)
        x = all_modes(fvk, ctx, lambda: fvk.solv_direct(ctx, imat, b))
        for label, x1 in x[1:]:
            assert np.array_equal(x[0][1], x1), ("sweeps", imat, label)
        say("  imat %d: factor and sweeps bitwise, levels %s" % (imat, ctx.levels()))
    ctx.destroy()
    m.free()


def test_krylov_on_one_launch_sweeps_bitwise():
    fvk = require_lib()
    n = 48
    say("krylov: box %d^3" % n)
    m = fvk.Mesh.box(n, n, n, jitter=0.05, seed=0x5EED0201)
    ctx = fvk.grd_solv(m, rcm=True)
    pl, ngp, ngn = random_matrix(m, 11)
    rng = np.random.default_rng(3)
    b3 = rng.standard_normal((m.n_vol, 3))
    fvk.set_matrix(ctx, pl, ngp, ngn)

    def bicg():
        x = np.zeros((m.n_vol, 3))
        r = b3.copy()
        rn, st = fvk.solv_bicg(ctx, 2, 1e-30, 6, x, r)
        return np.concatenate([x.ravel(), r.ravel(), rn])

    r = all_modes(fvk, ctx, bicg)
    for label, b in r[1:]:
        assert np.array_equal(r[0][1], b), ("BiCGStab (3 right-hand sides, DILU)", label)
    fvk.set_matrix(ctx, pl, ngp, ngp)
    bb = b3[:, 0] - b3[:, 0].mean()

    def cg():
        x = np.zeros((m.n_vol, 1))
        r = bb.reshape(-1, 1).copy()
        rn, st = fvk.solv_cg(ctx, 1, 1e-30, 20, x, r)
        return np.concatenate([x.ravel(), r.ravel(), rn])

    r = all_modes(fvk, ctx, cg)
    for label, b in r[1:]:
        assert np.array_equal(r[0][1], b), ("PCG (DIC)", label)
    ctx.destroy()
    m.free()


@pytest.mark.parametrize("graph", [False, True])
def test_shipped_simple_one_launch_bitwise(graph):
    fvk = require_lib()
    n = 24
    say("shipped SIMPLE: box %d^3 graph %s" % (n, graph))
    m = fvk.Mesh.box(n, n, n, jitter=0.02, seed=0x5EED0301)
    out = []
    for label, params in MODES:
        ctx = fvk.grd_solv(m, rcm=True)
        set_mode(fvk, ctx, params)
        fl = fvk.Flow(ctx, fvk.flow_cfg())
        fl.set_bnd_u(31, [1.0, 0.0, 0.0])
        h = (fl.run_graph if graph else fl.run)(1, 5)
        st = fl.state()
        out.append((st, [x["mx_res_n"] for x in h]))
        fl.destroy()
        ctx.destroy()
    s0, h0 = out[0]
    for (label, _), (s1, h1) in zip(MODES[1:], out[1:]):
        for k in ("u_cell", "p", "dpdx", "spd"):
            assert np.array_equal(s0[k], s1[k]), (label, k)
        assert h0 == h1, label
    say("  residual history %s" % h0)
    m.free()


def test_sweep_wait_failure_is_reported():
    """The failure path of the one-launch factor and sweeps, forced from the
    host: with the wait-bound flag raised ("sweep_fail"), rows of the ranges
    form that wait give up within 256 passes, blocks stop taking ranges, the
    launches end, and the factorisation and sweeps report FVK_ERR_HIP; a new
    context on the same mesh gives the bits of the first."""
    fvk = require_lib()
    n = 32
    say("sweep failure path: box %d^3" % n)
    m = fvk.Mesh.box(n, n, n, jitter=0.05, seed=0x5EED0401)
    pl, ngp, ngn = random_matrix(m, 19)
    b = np.random.default_rng(9).standard_normal(m.n_vol)
    ctx = fvk.grd_solv(m)
    fvk.set_matrix(ctx, pl, ngp, ngn)
    x0 = fvk.solv_direct(ctx, 2, b)
    fvk.set_param(ctx, "sweep_fail", 1)
    for call in (lambda: fvk.solv_precond(ctx, 1), lambda: fvk.solv_direct(ctx, 2, b)):
        with pytest.raises(fvk.FvkError) as e:
            call()
        assert e.value.status == fvk.FVK_ERR_HIP and "exceeded its bound" in str(e.value)
    ctx.destroy()
    ctx = fvk.grd_solv(m)
    fvk.set_matrix(ctx, pl, ngp, ngn)
    assert np.array_equal(fvk.solv_direct(ctx, 2, b), x0)
    ctx.destroy()
    m.free()
    say("  reported, and a new context sweeps as before")
