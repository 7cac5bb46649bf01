"""FP32-inner defect correction and the factorised preconditioner on the GPU.

Every solve is decided by the FP64 true residual |b - A x|_1 / |b|_1,
recomputed here on the CPU by the oracle's own SpMV (oracle/fvo, the
reference's face-scatter product) from the solution the library returns.
The factor's action z = G G^T r is compared with oracle/learn.apply on the
same FP16-rounded L and Delta^2 (the library keeps both in FP16, Delta^2 =
exp(2 s) rounded once from FP64 with 2 s held in [-9.7, 11]; float32
arithmetic inside: tolerance 2e-6 relative to max |z|).
"""
import numpy as np
import pytest

from oracle import fvo
from oracle import learn as L
from tests.helpers import require_lib

pytestmark = pytest.mark.gpu


def setup(n, jitter, seed, rcm=False, sigma=1.0):
    fvk = require_lib()
    m = fvk.Mesh.box(n, n, n, jitter=jitter, seed=seed)
    ctx = fvk.grd_solv(m, rcm=rcm)
    b = fvk.sys_random(ctx, seed + 1, sigma)
    pl, a, _ = fvk.sys_get(ctx)
    om = fvo.OMesh.box(n, n, n, jitter=jitter, seed=seed)
    t, g = om.topology(), om.geometry()
    S = L.System.from_mesh(t, g, pl, -a)
    A = fvo.OSys(om, pl, -a, -a)
    return fvk, m, ctx, om, S, A, b


def true_res(A, x, b):
    return np.abs(b - A.spmv(x)).sum() / np.abs(b).sum()


@pytest.mark.parametrize("rcm", [False, True])
def test_dc32_jacobi_converges_to_fp64_residual(rcm):
    fvk, m, ctx, om, S, A, b = setup(16, 0.2, 0x51, rcm)
    try:
        plain = fvk.sys_solve(ctx, fvk.SOLVE_PCG64, 0, 1e-8)
        xp = fvk.sys_x(ctx)
        mixed = fvk.sys_solve(ctx, fvk.SOLVE_DC32, 0, 1e-8)
        xm = fvk.sys_x(ctx)
        print("plain", plain["iters"], "mixed", mixed["iters"], mixed["res_true_l1"])
        assert plain["status"] == fvk.FVK_OK and mixed["status"] == fvk.FVK_OK
        assert true_res(A, xp, b) < 1e-8
        assert true_res(A, xm, b) < 1e-8
        assert abs(true_res(A, xm, b) - mixed["res_true_l1"]) < 1e-3 * mixed["res_true_l1"]
        # same solution up to the null space (constants) of the Neumann system
        dx = (xm - xm.mean()) - (xp - xp.mean())
        assert np.max(np.abs(dx)) < 1e-4 * np.max(np.abs(xp - xp.mean()))
        # FP32 inner solves restart the recurrence, but not by much
        assert mixed["iters"] < 1.5 * plain["iters"]
    finally:
        om.free()


@pytest.mark.parametrize("factor", [False, True])
def test_dc32_carried_direction_saves_iterations(factor):
    """Correction steps that continue the previous step's search direction
    reach the same FP64 residual as restarted ones, in fewer iterations.  The
    inner target is held at 1e-3 so that several correction steps run."""
    fvk, m, ctx, om, S, A, b = setup(20, 0.2, 0x55, rcm=True)
    try:
        method = fvk.SOLVE_DC32
        if factor:
            lf, sf = L.fsai_closed_form(S)
            inner = om.topology()["fce_in"] > 0
            l = np.zeros(ctx.n_fce, np.float32)
            sig = np.zeros(ctx.n_fce, np.float32)
            l[inner] = lf
            sig[inner] = sf
            fvk.factor_set(ctx, l, sig)
            method = fvk.SOLVE_DC32_FACTOR
        fvk.set_param(ctx, "dc_inner_reduc", 1e-3)
        fvk.set_param(ctx, "dc_replace_every", 0)
        out = {}
        for carry in (0, 1):
            fvk.set_param(ctx, "dc_carry", carry)
            info = fvk.sys_solve(ctx, method, 0, 1e-9)
            x = fvk.sys_x(ctx)
            assert info["status"] == fvk.FVK_OK and info["fallback"] == 0
            assert true_res(A, x, b) < 1e-9
            out[carry] = info
            print("carry", carry, info["iters"], info["inner_iters"])
        assert out[1]["outer"] >= 3
        assert out[1]["iters"] < out[0]["iters"]
    finally:
        om.free()


@pytest.mark.parametrize("factor", [False, True])
def test_dc32_residual_replacement_saves_iterations(factor):
    """Replacing the recursive FP32 residual by the FP64 residual of the
    solution so far every 32 inner iterations (the direction carries on)
    reaches the FP64 residual in one inner solve and in fewer iterations than
    restarted correction steps (dc_replace_every 0: two inner solves, the
    first stopped at 1e-5); replacing every 8 iterations converges too."""
    fvk, m, ctx, om, S, A, b = setup(40, 0.2, 0x5B)
    try:
        method = fvk.SOLVE_DC32
        if factor:
            lf, sf = L.fsai_closed_form(S)
            inner = om.topology()["fce_in"] > 0
            l = np.zeros(ctx.n_fce, np.float32)
            sig = np.zeros(ctx.n_fce, np.float32)
            l[inner] = lf
            sig[inner] = sf
            fvk.factor_set(ctx, l, sig)
            method = fvk.SOLVE_DC32_FACTOR
        out = {}
        for R in (0, 32, 8):
            fvk.set_param(ctx, "dc_replace_every", R)
            info = fvk.sys_solve(ctx, method, 0, 1e-8)
            x = fvk.sys_x(ctx)
            print("replace every", R, info["iters"], info["inner_iters"], true_res(A, x, b))
            assert info["status"] == fvk.FVK_OK and info["fallback"] == 0, (R, info)
            assert true_res(A, x, b) < 1e-8
            out[R] = info
        assert out[0]["outer"] == 2 and out[32]["outer"] == 1
        assert out[32]["iters"] < out[0]["iters"]
        assert out[8]["iters"] < out[0]["iters"] + 10
    finally:
        om.free()


def test_dc32_factor_spike_rhs_converges_without_fallback():
    """A right-hand side concentrated in one cell (minus its mean: the
    Neumann system is singular) on 48^3 rows: after the inner solve's
    normalisation to mean |r| = 1 the spike is ~1e5, beyond FP16's range.
    The FP16 intermediate of the factor path is scaled by the RMS of
    Delta u, not by the mean, so the solve must converge without breaking
    down (no fallback), checked by the FP64 true residual."""
    fvk = require_lib()
    n = 48
    m = fvk.Mesh.box(n, n, n, jitter=0.2, seed=0x56)
    ctx = fvk.grd_solv(m, rcm=True)
    fvk.sys_random(ctx, 0x57, 1.0)
    pl, a, _ = fvk.sys_get(ctx)
    b = np.full(ctx.n_vol, -1.0 / ctx.n_vol)
    b[ctx.n_vol // 3] += 1.0
    fvk.sys_set(ctx, pl, a, b)
    fvk.factor_set(ctx, np.zeros(ctx.n_fce, np.float32), np.zeros(ctx.n_fce, np.float32))
    out = fvk.sys_solve(ctx, fvk.SOLVE_DC32_FACTOR, 0, 1e-8)
    print("spike: iterations", out["iters"], "inner", out["inner_iters"], "res", out["res_true_l1"])
    assert out["status"] == fvk.FVK_OK and out["fallback"] == 0
    assert out["res_true_l1"] < 1e-8


def test_dc32_shuffled_triangles_32bit_columns():
    """A Delaunay channel of ~40,000 triangles in shuffled cell order, solved
    in that order (no RCM): neighbours lie up to n rows apart, beyond 16-bit
    column offsets, so the inner system takes 32-bit columns (seen in the
    SpMV's byte count).  The Jacobi and the factor paths (zero factor and a
    briefly trained network) reach the FP64 residual computed by the oracle's
    SpMV, without fallback."""
    from programs_amd import meshgen
    fvk = require_lib()
    vx, ty, vp, vr, bnd = meshgen.channel_delaunay(40000, seed=0x58)
    nt = len(ty)
    perm = np.random.default_rng(0x59).permutation(nt)
    vr = vr.reshape(nt, 3)[perm].ravel()
    bnd = bnd.reshape(nt, 3)[perm].ravel()
    m = fvk.Mesh.from_cells(2, vx, ty, vp, vr, bnd)
    ctx = fvk.grd_solv(m)
    b = fvk.sys_random(ctx, 0x5A, 1.0)
    pl, a, _ = fvk.sys_get(ctx)
    vx3 = np.zeros((len(vx), 3))
    vx3[:, :2] = vx
    om = fvo.OMesh.raw(2, vx3, ty, vp, vr, bnd)
    try:
        A = fvo.OSys(om, pl, -a, -a)
        nnz_off = 2 * int(np.sum(om.topology()["fce_in"] > 0))
        out = fvk.sys_solve(ctx, fvk.SOLVE_DC32, 0, 1e-8)
        assert out["status"] == fvk.FVK_OK and true_res(A, fvk.sys_x(ctx), b) < 1e-8
        fvk.factor_set(ctx, np.zeros(ctx.n_fce, np.float32), np.zeros(ctx.n_fce, np.float32))
        out = fvk.sys_solve(ctx, fvk.SOLVE_DC32_FACTOR, 0, 1e-8, time_kernel=True)
        assert out["status"] == fvk.FVK_OK and out["fallback"] == 0
        assert true_res(A, fvk.sys_x(ctx), b) < 1e-8
        # 32-bit columns; on triangles (three neighbours a row) the SpMV reads
        # the full rows, a column and a value per off-diagonal
        assert out["kern_bytes"] == 8.0 * ctx.n_vol + 8.0 * nnz_off
        model = fvk.Model(seed=11)
        model.train(ctx, 20, lr=3e-3)
        model.build_factor(ctx)
        lrn = fvk.sys_solve(ctx, fvk.SOLVE_DC32_FACTOR, 0, 1e-8)
        model.destroy()
        print("shuffled triangles: jacobi-factor", out["iters"], "learned", lrn["iters"])
        assert lrn["status"] == fvk.FVK_OK and lrn["fallback"] == 0
        assert true_res(A, fvk.sys_x(ctx), b) < 1e-8
    finally:
        om.free()


def test_factor_apply_matches_oracle():
    fvk, m, ctx, om, S, A, b = setup(12, 0.2, 0x52, rcm=True)
    try:
        rng = np.random.default_rng(3)
        inner = np.zeros(ctx.n_fce, bool)
        t = om.topology()
        inner[t["fce_in"] > 0] = True
        l = np.zeros(ctx.n_fce, np.float32)
        sig = np.zeros(ctx.n_fce, np.float32)
        l[inner] = (0.3 * rng.standard_normal(inner.sum())).astype(np.float32)
        sig[inner] = (0.1 * rng.standard_normal(inner.sum())).astype(np.float32)
        fvk.factor_set(ctx, l, sig)
        r = rng.standard_normal(ctx.n_vol)
        z = fvk.factor_apply(ctx, r)
        # the library keeps L and Delta^2 in FP16
        l16 = l[inner].astype(np.float16).astype(np.float64)
        w, e, l64, s = L.factor(S, l16, sig[inner].astype(np.float64))
        e16 = np.exp(np.clip(2.0 * s, -9.7, 11.0)).astype(np.float16).astype(np.float64)
        zr = L.apply(S, w, e16, l64, r)
        assert np.max(np.abs(z - zr)) <= 2e-6 * np.max(np.abs(zr))
        # zero factor: Jacobi
        fvk.factor_set(ctx, np.zeros_like(l), np.zeros_like(sig))
        z0 = fvk.factor_apply(ctx, r)
        np.testing.assert_allclose(z0, r / S.d, rtol=2e-6)
    finally:
        om.free()


def test_dc32_factor_fsai_cuts_iterations():
    fvk, m, ctx, om, S, A, b = setup(20, 0.2, 0x53)
    try:
        jac = fvk.sys_solve(ctx, fvk.SOLVE_DC32, 0, 1e-8)
        lf, sf = L.fsai_closed_form(S)
        t = om.topology()
        inner = t["fce_in"] > 0
        l = np.zeros(ctx.n_fce, np.float32)
        sig = np.zeros(ctx.n_fce, np.float32)
        l[inner] = lf
        sig[inner] = sf
        fvk.factor_set(ctx, l, sig)
        fs = fvk.sys_solve(ctx, fvk.SOLVE_DC32_FACTOR, 0, 1e-8)
        x = fvk.sys_x(ctx)
        print("jacobi", jac["iters"], "fsai", fs["iters"])
        assert fs["status"] == fvk.FVK_OK and fs["fallback"] == 0
        assert true_res(A, x, b) < 1e-8
        assert fs["iters"] < 0.75 * jac["iters"]
    finally:
        om.free()


def test_dc32_factor_falls_back_to_plain_pcg():
    fvk, m, ctx, om, S, A, b = setup(12, 0.2, 0x54)
    try:
        plain = fvk.sys_solve(ctx, fvk.SOLVE_PCG64, 0, 1e-8)
        # a valid (SPD) but useless factor: needs far more iterations than
        # the plain solver, so the budget of plain iterations runs out
        rng = np.random.default_rng(0)
        l = (3.0 * rng.standard_normal(ctx.n_fce)).astype(np.float32)
        sig = (1.5 * rng.standard_normal(ctx.n_fce)).astype(np.float32)
        fvk.factor_set(ctx, l, sig)
        out = fvk.sys_solve(ctx, fvk.SOLVE_DC32_FACTOR, 0, 1e-8, max_iter=plain["iters"])
        x = fvk.sys_x(ctx)
        assert out["fallback"] == 1 and out["status"] == fvk.FVK_FALLBACK
        assert true_res(A, x, b) < 1e-8
    finally:
        om.free()


def test_zero_rhs_and_short_budget():
    """b = 0 returns x = 0 at once for every method (solv_cg.f90:22 returns
    when max res0_n < small); a budget below what the tolerance needs ends
    with FVK_NOT_CONVERGED and a finite, partly converged x."""
    fvk, m, ctx, om, S, A, b = setup(10, 0.2, 0x56)
    try:
        lf, sf = L.fsai_closed_form(S)
        inner = om.topology()["fce_in"] > 0
        l = np.zeros(ctx.n_fce, np.float32)
        sig = np.zeros(ctx.n_fce, np.float32)
        l[inner] = lf
        sig[inner] = sf
        fvk.factor_set(ctx, l, sig)
        pl, a, _ = fvk.sys_get(ctx)
        fvk.sys_set(ctx, pl, a, np.zeros_like(b))
        for method in (fvk.SOLVE_PCG64, fvk.SOLVE_DC32, fvk.SOLVE_DC32_FACTOR):
            out = fvk.sys_solve(ctx, method, 0, 1e-8)
            assert out["status"] == fvk.FVK_OK and out["iters"] == 0 and out["fallback"] == 0, (method, out)
            assert not np.any(fvk.sys_x(ctx)), method
        fvk.sys_set(ctx, pl, a, b)
        for method in (fvk.SOLVE_PCG64, fvk.SOLVE_DC32):
            out = fvk.sys_solve(ctx, method, 0, 1e-8, max_iter=10)
            x = fvk.sys_x(ctx)
            print(method, out["iters"], out["res_true_l1"])
            assert out["status"] == fvk.FVK_NOT_CONVERGED, (method, out)
            assert out["iters"] <= 10 and np.all(np.isfinite(x))
            assert 1e-8 < true_res(A, x, b) < 1.0
    finally:
        om.free()


def test_dc32_factor_breakdown_falls_back_to_plain_pcg():
    """An indefinite system (the Laplacian shifted below its first nonzero
    eigenvalue, so the near-constant mode is negative) with a right-hand side
    rich in that mode: p.q < 0 in the first inner iteration, a breakdown.
    The defect correction restarts twice from its finite x, breaks down
    again, and the plain FP64 Jacobi PCG takes over (the reference
    recurrence allows p.q < 0 and converges here in about 100 iterations):
    FVK_FALLBACK with the FP64 residual met."""
    fvk, m, ctx, om, S, A, b = setup(8, 0.2, 0x57)
    try:
        Ad = S.dense()
        ev = np.linalg.eigvalsh(Ad)
        shift = 0.3 * ev[1]
        pl, a, _ = fvk.sys_get(ctx)
        b2 = np.random.default_rng(3).standard_normal(len(pl)) + 1.0
        fvk.sys_set(ctx, pl - shift, a, b2)
        A2 = fvo.OSys(om, pl - shift, -a, -a)
        assert np.linalg.eigvalsh(Ad - shift * np.eye(len(pl)))[0] < 0
        fvk.factor_set(ctx, np.zeros(ctx.n_fce, np.float32), np.zeros(ctx.n_fce, np.float32))
        out = fvk.sys_solve(ctx, fvk.SOLVE_DC32_FACTOR, 0, 1e-8, max_iter=2000)
        x = fvk.sys_x(ctx)
        print("indefinite:", out["status"], out["fallback"], out["iters"], out["outer"], true_res(A2, x, b2))
        assert out["fallback"] == 1 and out["status"] == fvk.FVK_FALLBACK
        assert out["outer"] >= 2  # the inner solve was restarted after a breakdown
        assert true_res(A2, x, b2) < 2e-8
    finally:
        om.free()


def test_non_finite_rhs_is_an_argument_error():
    """A NaN or infinite right-hand side is rejected (FVK_ERR_ARG) by every
    method, before any iteration."""
    fvk, m, ctx, om, S, A, b = setup(6, 0.2, 0x58)
    try:
        fvk.factor_set(ctx, np.zeros(ctx.n_fce, np.float32), np.zeros(ctx.n_fce, np.float32))
        pl, a, _ = fvk.sys_get(ctx)
        for bad in (np.nan, np.inf):
            b2 = b.copy()
            b2[5] = bad
            fvk.sys_set(ctx, pl, a, b2)
            for method in (fvk.SOLVE_PCG64, fvk.SOLVE_DC32, fvk.SOLVE_DC32_FACTOR):
                with pytest.raises(fvk.FvkError) as e:
                    fvk.sys_solve(ctx, method, 0, 1e-8)
                assert e.value.status == -1 and "non-finite" in str(e.value)
        # the context stays usable
        fvk.sys_set(ctx, pl, a, b)
        out = fvk.sys_solve(ctx, fvk.SOLVE_DC32, 0, 1e-8)
        assert out["status"] == fvk.FVK_OK and true_res(A, fvk.sys_x(ctx), b) < 1e-8
    finally:
        om.free()


def test_gpu_fsai_closed_form_matches_oracle():
    """fvk_factor_fsai builds on the GPU the factor the oracle's
    fsai_closed_form gives (l, sig to float32 rounding), and the solve with
    it takes the iterations of the oracle's factor loaded by fvk_factor_set."""
    fvk, m, ctx, om, S, A, b = setup(14, 0.2, 0x59)
    try:
        fvk.factor_fsai(ctx)
        l, sig = fvk.factor_get(ctx)
        lf, sf = L.fsai_closed_form(S)
        inner = om.topology()["fce_in"] > 0
        assert np.allclose(l[inner], lf, rtol=2e-7, atol=0)
        assert np.allclose(sig[inner], sf, rtol=2e-6, atol=1e-12)
        assert not np.any(l[~inner]) and not np.any(sig[~inner])
        a = fvk.sys_solve(ctx, fvk.SOLVE_DC32_FACTOR, 0, 1e-8)
        ll = np.zeros(ctx.n_fce, np.float32)
        ss = np.zeros(ctx.n_fce, np.float32)
        ll[inner] = lf
        ss[inner] = sf
        fvk.factor_set(ctx, ll, ss)
        b2 = fvk.sys_solve(ctx, fvk.SOLVE_DC32_FACTOR, 0, 1e-8)
        print("fsai on the GPU", a["iters"], "from the oracle", b2["iters"])
        assert a["status"] == fvk.FVK_OK and a["fallback"] == 0 and a["res_true_l1"] < 1e-8
        assert abs(a["iters"] - b2["iters"]) <= 2
    finally:
        om.free()
