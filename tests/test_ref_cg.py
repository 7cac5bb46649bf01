"""The timing driver of the reference's PCG (oracle/_ref/cg/fvk_ref_cg, the
flang build of solv_cg.f90 behind bench.py's CPU baseline) on a small
lognormal pressure system: after a fixed number of iterations its solution
equals, bit for bit, that of the oracle's restatement of solv_cg on the same
face-based system, so the driver hands the reference the right matrix and
right-hand side.  Skipped where the reference build is absent."""
import numpy as np
import pytest

from oracle import fvo, refcg

pytestmark = pytest.mark.skipif(not refcg.available(), reason="reference PCG driver not built")


def lognormal_system(n, seed):
    om = fvo.OMesh.box(n, n, n, jitter=0.2, seed=seed)
    t, g = om.topology(), om.geometry()
    ip, inn = t["fce_ip"], t["fce_in"]
    m = inn > 0
    rng = np.random.default_rng(seed)
    a = np.zeros(len(ip))
    a[m] = g["fce_s"][m] / g["fce_lpn"][m] * np.exp(rng.standard_normal(int(m.sum())))
    pl = np.zeros(t["n_vol"])
    np.add.at(pl, ip[m] - 1, a[m])
    np.add.at(pl, inn[m] - 1, a[m])
    pl *= 1.0 + 1e-3  # a little diagonal dominance: nonsingular
    b = rng.standard_normal(t["n_vol"])
    return om, ip, inn, a, pl, b


@pytest.mark.parametrize("n,iters", [(6, 25), (9, 60)])
def test_reference_pcg_driver_matches_oracle(n, iters):
    om, ip, inn, a, pl, b = lognormal_system(n, 0x5EED0020 + n)
    try:
        _, x_ref = refcg.run(ip, inn, -a, -a, pl, b, iters, want_x=True)
        x_orc, _, _, its, _ = fvo.OSys(om, pl, -a, -a).solve("cg", 0, 0.0, iters, b)
        assert its[0] == iters
        assert np.array_equal(x_ref, x_orc[:, 0])
        assert np.max(np.abs(x_ref)) > 0
    finally:
        om.free()


def seqsum(a):
    """Sequential left-to-right sum (the order of the reference's sum())."""
    return float(np.cumsum(np.asarray(a, np.float64))[-1])


@pytest.mark.parametrize("singular", [False, True])
def test_reference_pcg_to_tolerance_matches_oracle_mode(singular):
    """The oracle's solve to a tolerance (fvo_pcg_tol) is the reference's
    solv_cg with reduc = tol on the right-hand side scaled to a mean |b| of
    1 (after taking out its mean when the rows sum to zero), the solution
    scaled back (and shifted to mean zero): bit for bit."""
    n, tol = 8, 1e-9
    om, ip, inn, a, pl, b = lognormal_system(n, 0x5EED0031)
    try:
        if singular:
            pl = pl / (1.0 + 1e-3)
            np.add.at(pl, [0], 0.0)
        A = fvo.OSys(om, pl, -a, -a)
        assert A.singular() == singular
        x_orc, its, r0, sing = A.solve_tol(b, tol)
        assert sing == singular and 1 < its < 20000
        nv = len(b)
        bp = b - seqsum(b) / nv if singular else b.copy()
        l1b = seqsum(np.abs(bp))
        assert r0 == l1b
        sigma = l1b / nv
        _, x_ref = refcg.run(ip, inn, -a, -a, pl, (1.0 / sigma) * bp, 20000, reduc=tol)
        x_ref = sigma * x_ref
        if singular:
            x_ref = x_ref - seqsum(x_ref) / nv
        assert np.array_equal(x_ref, x_orc)
    finally:
        om.free()
