"""The oracle's pressure solve to a tolerance (fvo_pcg_tol, the flow's
p_solver 3) and its singular-system handling, against what linear algebra
fixes: the dense inverse of a nonsingular system, the minimum-norm
(pseudo-inverse) solution of a pure-Neumann one, zero mean, the residual
bound, and the threshold of the singularity test.  Inside a SIMPLE
iteration the p' solution it returns is the pseudo-inverse solution of the
p' system the oracle assembled."""
import numpy as np
import pytest

from oracle import fvo


def dense(om, pl, ngp, ngn):
    t = om.topology()
    n = t["n_vol"]
    A = np.diag(np.asarray(pl, np.float64))
    m = t["fce_in"] > 0
    ip, inn = t["fce_ip"][m] - 1, t["fce_in"][m] - 1
    np.add.at(A, (ip, inn), np.asarray(ngp)[m])
    np.add.at(A, (inn, ip), np.asarray(ngn)[m])
    return A, n


def laplacian(n, seed, jitter=0.2):
    om = fvo.OMesh.box(n, n, n, jitter=jitter, seed=seed)
    t, g = om.topology(), om.geometry()
    m = t["fce_in"] > 0
    rng = np.random.default_rng(seed)
    a = np.zeros(len(m))
    a[m] = g["fce_s"][m] / g["fce_lpn"][m] * np.exp(rng.standard_normal(int(m.sum())))
    pl = np.zeros(t["n_vol"])
    np.add.at(pl, t["fce_ip"][m] - 1, a[m])
    np.add.at(pl, t["fce_in"][m] - 1, a[m])
    return om, pl, a, rng.standard_normal(t["n_vol"])


def test_nonsingular_solution_is_the_dense_solve():
    om, pl, a, b = laplacian(5, 0x5EED0041)
    try:
        pl = pl * (1.0 + 1e-2)
        S = fvo.OSys(om, pl, -a, -a)
        assert not S.singular()
        x, its, r0, sing = S.solve_tol(b, 1e-13)
        A, n = dense(om, pl, -a, -a)
        xd = np.linalg.solve(A, b)
        assert not sing and its > 1
        assert r0 == pytest.approx(np.abs(b).sum(), rel=1e-14)
        assert np.max(np.abs(x - xd)) <= 1e-9 * np.max(np.abs(xd))
        # the recurrence stops on its own residual; the true one is as small
        assert np.abs(b - A @ x).sum() / np.abs(b).sum() < 1e-12
        # a solve to a looser tolerance stops earlier and is looser
        x2, its2, _, _ = S.solve_tol(b, 1e-6)
        assert its2 < its
        assert 1e-9 < np.abs(b - A @ x2).sum() / np.abs(b).sum() < 2e-6
    finally:
        om.free()


def test_singular_solution_is_the_pseudo_inverse_solution():
    om, pl, a, b = laplacian(5, 0x5EED0042)
    try:
        S = fvo.OSys(om, pl, -a, -a)
        assert S.singular()
        A, n = dense(om, pl, -a, -a)
        x, its, r0, sing = S.solve_tol(b, 1e-13)
        assert sing
        bp = b - b.mean()
        assert r0 == pytest.approx(np.abs(bp).sum(), rel=1e-14)
        xp = np.linalg.pinv(A) @ b  # min-norm: orthogonal to the constants
        assert abs(x.sum()) <= 1e-12 * np.abs(x).sum()
        assert np.max(np.abs(x - xp)) <= 1e-9 * np.max(np.abs(xp))
        assert np.abs(bp - A @ x).sum() / np.abs(bp).sum() < 1e-12
        # scaling the right-hand side scales the solution (the normalisation
        # to a mean |b| of 1 is undone); a constant added to b changes nothing
        x3, _, _, _ = S.solve_tol(1e-7 * (b + 3.0), 1e-13)
        assert np.max(np.abs(x3 - 1e-7 * xp)) <= 1e-9 * 1e-7 * np.max(np.abs(xp))
    finally:
        om.free()


def test_singularity_threshold():
    om, pl, a, b = laplacian(4, 0x5EED0043, jitter=0.0)
    try:
        n = len(pl)
        # relative row defect 1e-13 per row: below the 1e-12 mean threshold
        assert fvo.OSys(om, pl * (1.0 + 1e-13), -a, -a).singular()
        assert not fvo.OSys(om, pl * (1.0 + 1e-11), -a, -a).singular()
        # one Dirichlet row is enough: defect 1/n of a unit row defect
        pl2 = pl.copy()
        pl2[0] *= 2.0
        assert not fvo.OSys(om, pl2, -a, -a).singular()
        assert n > 1
    finally:
        om.free()


@pytest.mark.parametrize("simpler", [0, 1])
def test_simple_p_solver3_solves_the_assembled_system(simpler):
    """In a SIMPLE iteration with p_solver 3 the p' solution is the
    pseudo-inverse solution of the p' system the oracle assembled (the
    closed cavity's is singular), and the relaxed pressure update uses it."""
    n = 5
    om = fvo.OMesh.box(n, n, n, jitter=0.05, seed=0x5EED0044)
    try:
        cfg = fvo.default_cfg(simpler=simpler, imat_p=0, urf_p=0.3, p_solver=3, p_tol=1e-13)
        F = fvo.OFlow(om, cfg)
        F.set_bnd_u(31, [1.0, 0.0, 0.0])
        for it in (1, 2, 3):
            p_old = F.fields()["p"]
            F.outer(it)
            f = F.fields()
            rhs, pc = F.last_pp_system()
            A, _ = dense(om, f["pl"], f["ng_p"], f["ng_n"])
            assert fvo.OSys(om, f["pl"], f["ng_p"], f["ng_n"]).singular()
            xp = np.linalg.pinv(A) @ rhs
            assert np.max(np.abs(pc - xp)) <= 1e-8 * max(np.max(np.abs(xp)), 1e-300)
            # the first iteration's p' right-hand side is zero (no mass
            # imbalance yet): nothing to solve
            assert F.f.it_pp > 1 if it > 1 else (F.f.it_pp == 0 and not np.any(pc))
            if not simpler:
                assert np.allclose(f["p"] - p_old, 0.3 * (pc - pc[0]), rtol=0, atol=1e-14 * np.max(np.abs(pc)))
        F.free()
    finally:
        om.free()


def test_all_cores_baseline_pcg_is_the_reference_recurrence():
    """The all-cores CPU baseline's PCG (fvo_pcg_omp: row-gather product,
    OpenMP) on one thread is the oracle's solv_cg bit for bit (the gather
    adds each row's terms in the face scatter's order); on several threads
    only its reductions' rounding differs."""
    om, pl, a, b = laplacian(9, 0x5EED0045)
    try:
        b = b - b.mean()
        S = fvo.OSys(om, pl, -a, -a)
        x1 = S.solve("cg", 0, 0.0, 40, b)[0][:, 0]
        _, x2 = S.pcg_omp(b, 40, 1)
        assert np.array_equal(x1, x2)
        _, x4 = S.pcg_omp(b, 40, 4)
        assert np.max(np.abs(x4 - x1)) <= 1e-10 * np.max(np.abs(x1))
    finally:
        om.free()
