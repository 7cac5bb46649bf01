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
