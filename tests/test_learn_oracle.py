"""Pins the CPU restatement of the learned factor (oracle/learn.py) against
what the mathematics fixes, so that the GPU kernels can be checked against it:

* G G^T apply equals the dense product and is symmetric positive definite;
* the loss equals Kaporin's measure computed densely, log(tr/N) - log det/N,
  up to the constant log det(Ah)/N;
* the loss gradient matches central finite differences, on a hexahedral
  mesh and on a triangulated one (whose cell graph has 3-cycles);
* the loss is minimised, with zero gradient, by the FSAI factor, which on a
  hexahedral mesh has a closed form, itself checked against FSAI's
  definition (local solves);
* the MLP backward pass matches finite differences of the forward pass
  (float64, no rounding), and the bfloat16 emulation stays close to it.
"""
import numpy as np
import pytest

from oracle import fvo
from oracle import learn as L


def box_system(nx, ny, nz, jitter, seed, sigma=1.0):
    m = fvo.OMesh.box(nx, ny, nz, jitter=jitter, seed=seed)
    try:
        t, g = m.topology(), m.geometry()
    finally:
        m.free()
    rng = np.random.default_rng(seed)
    ip, inn = t["fce_ip"] - 1, t["fce_in"] - 1
    inner = inn >= 0
    a = g["fce_s"] / g["fce_lpn"] * np.exp(sigma * rng.standard_normal(len(ip)))
    ngp = np.where(inner, -a, 0.0)
    pl = np.zeros(t["n_vol"])
    np.add.at(pl, ip[inner], a[inner])
    np.add.at(pl, inn[inner], a[inner])
    pl += 1e-3 * pl.mean()  # a little diagonal weight: nonsingular, for dense checks
    return L.System.from_mesh(t, g, pl, ngp)


def tri_system(seed=3):
    """Triangles from a jittered structured split: cell graph with 3-cycles is
    not guaranteed, so we add a fan around an interior vertex of degree 3."""
    rng = np.random.default_rng(seed)
    # a small planar triangulation with a degree-3 vertex: the three triangles
    # around it are pairwise adjacent
    pts = np.array([[0, 0], [2, 0], [1, 1.7], [1, 0.6], [3, 1.5], [-1, 1.5], [1, -1.2]], float)
    tris = [(0, 1, 3), (1, 2, 3), (2, 0, 3), (1, 4, 2), (0, 2, 5), (0, 6, 1)]
    n = len(tris)
    edges = {}
    for c, tr in enumerate(tris):
        for k in range(3):
            e = tuple(sorted((tr[k], tr[(k + 1) % 3])))
            edges.setdefault(e, []).append(c)
    p, q = [], []
    for e, cs in sorted(edges.items()):
        if len(cs) == 2:
            p.append(min(cs)); q.append(max(cs))
    p, q = np.array(p), np.array(q)
    a = np.exp(rng.standard_normal(len(p)))
    d = np.zeros(n)
    np.add.at(d, p, a)
    np.add.at(d, q, a)
    d += 0.1
    k = len(p)
    geo = dict(s=np.ones(k), nrm=np.tile([1.0, 0, 0], (k, 1)), lpn=np.ones(k), wfp=np.full(k, 0.5),
               dxp=np.zeros((k, 3)), dxn=np.zeros((k, 3)), vol=np.ones(n))
    return L.System(p, q, a, d, geo)


def has_3cycle(S):
    for c in range(S.N):
        js = [S.n[k] for k in S.upper[c]]
        for x in js:
            for y in js:
                if x != y and y in S.nbr[x]:
                    return True
    return False


def random_factor(S, seed, scale=0.3):
    rng = np.random.default_rng(seed)
    return scale * rng.standard_normal(len(S.a)), 0.2 * rng.standard_normal(len(S.a))


def test_apply_is_dense_ggt_and_spd():
    S = box_system(4, 3, 3, 0.2, 11)
    l, sig = random_factor(S, 1)
    w, e, l, s = L.factor(S, l, sig)
    G = L.dense_factor(S, w, e, l)
    M = G @ G.T
    rng = np.random.default_rng(2)
    r = rng.standard_normal(S.N)
    np.testing.assert_allclose(L.apply(S, w, e, l, r), M @ r, rtol=1e-12, atol=1e-12)
    assert np.all(np.linalg.eigvalsh(M) > 0)
    # untrained (zero output layer): Jacobi
    w0, e0, l0, _ = L.factor(S, np.zeros(len(S.a)), np.zeros(len(S.a)))
    np.testing.assert_allclose(L.apply(S, w0, e0, l0, r), r / S.d, rtol=1e-14)


@pytest.mark.parametrize("which", ["hex", "tri"])
def test_loss_is_kaporin_measure(which):
    S = box_system(3, 3, 2, 0.2, 5) if which == "hex" else tri_system()
    if which == "tri":
        assert has_3cycle(S)
    l, sig = random_factor(S, 4)
    J, T = L.kaporin(S, l, sig, grad=False)
    A = S.dense()
    Ah = A / np.sqrt(np.outer(S.d, S.d))
    w, e, l64, s = L.factor(S, l, sig)
    Gh = np.diag(np.sqrt(S.d)) @ L.dense_factor(S, w, e, l64)
    K = Gh.T @ Ah @ Gh
    np.testing.assert_allclose(T, np.trace(K), rtol=1e-12)
    sign, logdet = np.linalg.slogdet(K)
    _, logdet_a = np.linalg.slogdet(Ah)
    assert sign > 0
    np.testing.assert_allclose(J, np.log(np.trace(K) / S.N) - (logdet - logdet_a) / S.N, rtol=1e-10, atol=1e-12)


@pytest.mark.parametrize("which", ["hex", "tri"])
def test_loss_gradient_finite_differences(which):
    S = box_system(3, 3, 3, 0.2, 8) if which == "hex" else tri_system()
    l, sig = random_factor(S, 9)
    J, T, gl, gs = L.kaporin(S, l, sig)
    h = 1e-6
    rng = np.random.default_rng(0)
    for k in rng.choice(len(l), size=min(12, len(l)), replace=False):
        e = np.zeros(len(l)); e[k] = h
        fd = (L.kaporin(S, l + e, sig, grad=False)[0] - L.kaporin(S, l - e, sig, grad=False)[0]) / (2 * h)
        assert abs(fd - gl[k]) <= 1e-6 * max(1.0, abs(fd)), (k, fd, gl[k])
        fd = (L.kaporin(S, l, sig + e, grad=False)[0] - L.kaporin(S, l, sig - e, grad=False)[0]) / (2 * h)
        # dJ/dsig_f is the same for all faces owned by one cell; check face k
        assert abs(fd - gs[k]) <= 1e-6 * max(1.0, abs(fd)), (k, fd, gs[k])


def test_fsai_is_the_minimiser():
    S = box_system(4, 4, 3, 0.2, 21)
    l, sig = L.fsai_closed_form(S)
    # the closed form is FSAI by its definition (local solves)
    Gd = L.fsai_dense(S)
    w, e, l64, _ = L.factor(S, l, sig)
    Gc = np.diag(np.sqrt(S.d)) @ L.dense_factor(S, w, e, l64)
    np.testing.assert_allclose(Gc, Gd, rtol=1e-10, atol=1e-12)
    # zero gradient there, and any perturbation raises the loss
    J, T, gl, gs = L.kaporin(S, l, sig)
    assert np.max(np.abs(gl)) < 1e-12 and np.max(np.abs(gs)) < 1e-12
    rng = np.random.default_rng(1)
    for _ in range(5):
        dl, ds = 1e-2 * rng.standard_normal(len(l)), 1e-2 * rng.standard_normal(len(l))
        assert L.kaporin(S, l + dl, sig + ds, grad=False)[0] > J
    # and it cuts PCG iterations against Jacobi (what a trained net aims at)
    A = S.dense()
    b = np.random.default_rng(3).standard_normal(S.N)
    it_j, _ = L.pcg_iterations(A, b, lambda r: r / S.d)
    it_f, _ = L.pcg_iterations(A, b, lambda r: L.apply(S, w, e, l64, r))
    assert it_f < it_j


def test_mlp_backward_finite_differences():
    S = box_system(3, 3, 2, 0.2, 2)
    X = L.features(S)
    rng = np.random.default_rng(5)
    theta = L.init_params(7).astype(np.float64)
    theta[L.OFF_W3:L.NPARAM] = 0.3 * rng.standard_normal(L.NPARAM - L.OFF_W3)
    dO = rng.standard_normal((len(X), 2))

    def f(th):
        o, _ = L.mlp_forward(th, X, emulate=False)
        return float(np.sum(o * dO))

    o, cache = L.mlp_forward(theta, X, emulate=False)
    g = L.mlp_backward(cache, dO, emulate=False)
    h = 1e-6
    for k in list(rng.choice(L.NPARAM, size=40, replace=False)) + [L.OFF_B1, L.OFF_B2 + 3, L.OFF_B3 + 1]:
        e = np.zeros(L.NPARAM); e[k] = h
        fd = (f(theta + e) - f(theta - e)) / (2 * h)
        assert abs(fd - g[k]) <= 1e-6 * max(1.0, abs(fd)), (k, fd, g[k])
    # bfloat16 emulation: same function to bfloat16 accuracy
    oe, ce = L.mlp_forward(theta.astype(np.float32), X, emulate=True)
    assert np.max(np.abs(oe - o)) < 0.05 * np.max(np.abs(o))
    ge = L.mlp_backward(ce, dO, emulate=True)
    assert np.linalg.norm(ge - g) < 0.05 * np.linalg.norm(g)


def test_bf16_rounding():
    x = np.array([1.0, 1.00390625, 1.005859375, -3.0e-3, 65504.0, 0.0], np.float32)
    y = L.bf16(x)
    # 1 + 2^-8 is a tie: to even (1.0); 1 + 1.5*2^-8 rounds up to 1 + 2^-7
    np.testing.assert_array_equal(y[:3], np.array([1.0, 1.0, 1.0078125], np.float32))
    assert y[5] == 0.0
    assert np.all(np.abs(y[3:5] - x[3:5]) <= np.abs(x[3:5]) * 2.0 ** -8)


def hex_system(n, L=1.0, scale=1.0):
    """Uniform n^3 hex box of side L; a_f = scale * s/l_pn (pure Neumann)."""
    om = fvo.OMesh.box(n, n, n, L=(L, L, L))
    t, g = om.topology(), om.geometry()
    om.free()
    m = t["fce_in"] > 0
    a = np.zeros(len(m))
    a[m] = scale * g["fce_s"][m] / g["fce_lpn"][m]
    pl = np.zeros(t["n_vol"])
    np.add.at(pl, t["fce_ip"][m] - 1, a[m])
    np.add.at(pl, t["fce_in"][m] - 1, a[m])
    return L_.System.from_mesh(t, g, pl, -a), t, g


L_ = L


@pytest.mark.parametrize("n,side", [(5, 1.0), (6, 3.7)])
def test_features_closed_form_on_uniform_hex(n, side):
    """On a uniform hex box with a_f = s/l_pn every feature has a closed form:
    a/sqrt(a_pp a_nn) = 1/sqrt(k_p k_n), a/a_pp = 1/k_p, a/a_nn = 1/k_n with
    k_c the number of interior neighbours of c; s/h^2 = 1; the normal is the
    unit vector from owner to neighbour centre; l_pn/h = 1; w_fp = 1/2;
    |dx_p|/h = |dx_n|/h = 0 (the auxiliary poles are the cell centres);
    log(|V_p|/|V_n|) = 0.  Interior-interior faces are
    (1/6, 1/6, 1/6, 1, e_i, 1, 1/2, 0, 0, 0).  Independent of the box's size
    (dimensionless: s/h instead of s/h^2 fails here) and of a scaling of
    the matrix."""
    S, t, g = hex_system(n, side)
    X = L_.features(S)
    k = np.zeros(S.N)
    np.add.at(k, S.p, 1)
    np.add.at(k, S.n, 1)
    kp, kn = k[S.p], k[S.n]
    assert np.allclose(X[:, 0], 1.0 / np.sqrt(kp * kn), rtol=1e-12, atol=0)
    assert np.allclose(X[:, 1], 1.0 / kp, rtol=1e-12, atol=0)
    assert np.allclose(X[:, 2], 1.0 / kn, rtol=1e-12, atol=0)
    assert np.allclose(X[:, 3], 1.0, rtol=1e-12, atol=0)
    c = g["vol_x"]
    e = (c[S.n] - c[S.p]) / (side / n)
    assert np.allclose(X[:, 4:7], e, rtol=0, atol=1e-12)
    assert np.allclose(np.abs(X[:, 4:7]).sum(1), 1.0)
    assert np.allclose(X[:, 7], 1.0, rtol=1e-12, atol=0)
    assert np.allclose(X[:, 8], 0.5, rtol=1e-12, atol=0)
    assert np.allclose(X[:, 9:11], 0.0, rtol=0, atol=1e-12)
    assert np.allclose(X[:, 11], 0.0, rtol=0, atol=1e-12)
    inner = (kp == 6) & (kn == 6)
    assert inner.sum() == 3 * (n - 2) ** 2 * (n - 3)
    assert np.allclose(X[inner, :4], [1 / 6, 1 / 6, 1 / 6, 1.0], rtol=1e-12)
    # boundary-adjacent: an owner on a box face (5 interior neighbours) next
    # to an interior cell
    bnd = (kp == 5) & (kn == 6)
    assert bnd.any()
    assert np.allclose(X[bnd, :3], [1 / np.sqrt(30), 1 / 5, 1 / 6], rtol=1e-12)
    # scaling the matrix changes nothing
    S2, _, _ = hex_system(n, side, scale=7.5)
    assert np.allclose(L_.features(S2), X, rtol=1e-12, atol=1e-15)


def test_features_scale_invariant_on_jittered_mesh():
    """The features are dimensionless: the same jittered mesh scaled by 10
    (with a matrix scaled likewise) gives the same features."""
    def case(side):
        om = fvo.OMesh.box(5, 5, 5, jitter=0.2, seed=0x5EED0051, L=(side, side, side))
        t, g = om.topology(), om.geometry()
        om.free()
        m = t["fce_in"] > 0
        rng = np.random.default_rng(1)
        a = np.zeros(len(m))
        a[m] = g["fce_s"][m] / g["fce_lpn"][m] * np.exp(rng.standard_normal(int(m.sum())))
        pl = np.zeros(t["n_vol"])
        np.add.at(pl, t["fce_ip"][m] - 1, a[m])
        np.add.at(pl, t["fce_in"][m] - 1, a[m])
        return L_.features(L_.System.from_mesh(t, g, pl, -a))
    X1, X2 = case(1.0), case(10.0)
    assert np.allclose(X1, X2, rtol=1e-9, atol=1e-12)
    assert np.std(X1[:, 9]) > 0 and np.std(X1[:, 11]) > 0
