"""CPU restatement of the learned factorised preconditioner (numpy).

TEST INFRASTRUCTURE: imported only by tests/ and bench.py's cpu_baseline leg,
as the checker.  The product (programs_amd/) never imports it.

This part has no counterpart in the reference (spray-simulation solves its
pressure systems with solv_cg.f90 and the solv_precond.f90 factors only).  It
restates, in plain numpy, what the GPU library computes:

* per interior face f = (p, n), p < n in the reference numbering, twelve
  dimensionless features of the Jacobi-scaled matrix and the face geometry;
* a 12-64-64-2 tanh MLP (weights rounded to bfloat16, activations rounded to
  bfloat16 before each product, float32 accumulation) giving l_f and sig_f;
* the factor G = D^-1/2 (I + L) Delta: L strictly lower (row n, column p)
  with L[n, p] = l_f, Delta = diag(exp(s_c)), s_c = sum of sig_f over the
  faces owned by c (the entries of column c); the preconditioner applies
  z = G G^T r;
* Kaporin's measure of the Jacobi-scaled preconditioned matrix as training
  loss, J = log(tr(Gh^T Ah Gh) / N) - (2/N) sum_i s_i, with Ah = D^-1/2 A D^-1/2
  and Gh = D^1/2 G, and its gradient.  For a fixed pattern, the minimiser of
  J is the factorised sparse approximate inverse (FSAI, Kolotilina and
  Yeremin 1993), which the tests use as a known answer.

Parameter vector layout (float32, 5122 entries): W1[12][64], b1[64],
W2[64][64], b2[64], W3[64][2], b3[2] (input-major matrices).
"""
import numpy as np

NIN, NH, NOUT = 12, 64, 2
OFF_W1 = 0
OFF_B1 = OFF_W1 + NIN * NH
OFF_W2 = OFF_B1 + NH
OFF_B2 = OFF_W2 + NH * NH
OFF_W3 = OFF_B2 + NH
OFF_B3 = OFF_W3 + NH * NOUT
NPARAM = OFF_B3 + NOUT


def unpack(theta):
    t = np.asarray(theta)
    return (t[OFF_W1:OFF_B1].reshape(NIN, NH), t[OFF_B1:OFF_W2], t[OFF_W2:OFF_B2].reshape(NH, NH),
            t[OFF_B2:OFF_W3], t[OFF_W3:OFF_B3].reshape(NH, NOUT), t[OFF_B3:NPARAM])


def pack(W1, b1, W2, b2, W3, b3):
    return np.concatenate([np.ravel(W1), b1, np.ravel(W2), b2, np.ravel(W3), b3])


def bf16(x):
    """Round float32 values to bfloat16 (nearest, ties to even), as float32."""
    a = np.ascontiguousarray(x, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).reshape(a.shape)


def init_params(seed):
    """Xavier-uniform hidden layers, zero output layer: the untrained factor is
    G = D^-1/2 (L = 0, Delta = I), i.e. Jacobi."""
    rng = np.random.default_rng(seed)
    W1 = rng.uniform(-1, 1, (NIN, NH)) * np.sqrt(6.0 / (NIN + NH))
    W2 = rng.uniform(-1, 1, (NH, NH)) * np.sqrt(6.0 / (NH + NH))
    return pack(W1, np.zeros(NH), W2, np.zeros(NH), np.zeros((NH, NOUT)), np.zeros(NOUT)).astype(np.float32)


class System:
    """An SPD (or singular Neumann) face-based system A on a mesh.

    p, n : interior faces' cells (0-based, p < n in reference numbering)
    a    : off-diagonal magnitudes, A[p, n] = A[n, p] = -a (a > 0)
    d    : diagonal of A
    geo  : dict with per-interior-face s, nrm (k x 3), lpn, wfp, dxp (k x 3),
           dxn (k x 3) and per-cell vol
    """

    def __init__(self, p, n, a, d, geo):
        self.p = np.asarray(p, np.int64)
        self.n = np.asarray(n, np.int64)
        self.a = np.asarray(a, np.float64)
        self.d = np.asarray(d, np.float64)
        self.geo = geo
        self.N = len(self.d)
        assert np.all(self.p < self.n)
        self.ah = -self.a / np.sqrt(self.d[self.p] * self.d[self.n])  # Ah[p, n] (< 0)
        # neighbour lists for the 3-cycle terms of the loss
        self.nbr = [dict() for _ in range(self.N)]
        for k, (i, j) in enumerate(zip(self.p, self.n)):
            self.nbr[i][j] = k
            self.nbr[j][i] = k
        self.upper = [[] for _ in range(self.N)]  # faces whose owner is c
        for k, i in enumerate(self.p):
            self.upper[i].append(k)

    @classmethod
    def from_mesh(cls, topo, geom, pl, ngp):
        """From oracle/fvo mesh dictionaries (1-based ids), the diagonal pl and
        the face coefficients ng%p (negative on interior faces)."""
        ip = topo["fce_ip"] - 1
        inn = topo["fce_in"] - 1
        m = inn >= 0
        nd = topo["n_dx"]

        def v3(a):
            out = np.zeros((int(m.sum()), 3))
            out[:, :nd] = a[m]
            return out
        geo = dict(s=geom["fce_s"][m], nrm=v3(geom["fce_n"]), lpn=geom["fce_lpn"][m], wfp=geom["fce_wfp"][m],
                   dxp=v3(geom["fce_dxp"]), dxn=v3(geom["fce_dxn"]), vol=geom["vol_v"])
        return cls(ip[m], inn[m], -np.asarray(ngp)[m], pl, geo)

    def dense(self):
        A = np.diag(self.d.copy())
        A[self.p, self.n] = -self.a
        A[self.n, self.p] = -self.a
        return A


def features(S):
    """Twelve features per interior face (float64)."""
    g = S.geo
    dp, dn = S.d[S.p], S.d[S.n]
    # |V|: the reference's volumes (x_f s_f n_x summed over faces from the
    # origin, grd_geom.f90:151-153,172-173) turn negative on strongly
    # warped cells
    vp = np.maximum(np.abs(g["vol"][S.p]), 1e-300)
    vn = np.maximum(np.abs(g["vol"][S.n]), 1e-300)
    h = np.cbrt(0.5 * (vp + vn))
    X = np.empty((len(S.a), NIN))
    X[:, 0] = S.a / np.sqrt(dp * dn)
    X[:, 1] = S.a / dp
    X[:, 2] = S.a / dn
    X[:, 3] = g["s"] / (h * h)
    X[:, 4:7] = g["nrm"]
    X[:, 7] = g["lpn"] / h
    X[:, 8] = g["wfp"]
    X[:, 9] = np.linalg.norm(g["dxp"], axis=1) / h
    X[:, 10] = np.linalg.norm(g["dxn"], axis=1) / h
    X[:, 11] = np.log(vp / vn)
    return X


def mlp_forward(theta, X, emulate=True):
    """Outputs (k x 2) and the activations kept for the backward pass.
    emulate: bfloat16 rounding of weights and activations, float32
    arithmetic (the GPU kernels); otherwise plain float64."""
    W1, b1, W2, b2, W3, b3 = unpack(theta)
    if emulate:
        f = np.float32
        X_, W1_, W2_, W3_ = bf16(X), bf16(W1), bf16(W2), bf16(W3)
        h1 = np.tanh((X_ @ W1_ + b1.astype(f)).astype(f))
        h1_ = bf16(h1)
        h2 = np.tanh((h1_ @ W2_ + b2.astype(f)).astype(f))
        h2_ = bf16(h2)
        o = (h2_ @ W3_ + b3.astype(f)).astype(f)
    else:
        X_, W1_, W2_, W3_ = X, W1, W2, W3
        h1 = np.tanh(X_ @ W1_ + b1)
        h1_ = h1
        h2 = np.tanh(h1_ @ W2_ + b2)
        h2_ = h2
        o = h2_ @ W3_ + b3
    return o, (X_, h1, h1_, h2, h2_, W1_, W2_, W3_)


def mlp_backward(cache, dO, emulate=True):
    """Parameter gradient (packed) for output gradient dO (k x 2)."""
    X_, h1, h1_, h2, h2_, W1_, W2_, W3_ = cache
    if emulate:
        dO = np.asarray(dO, np.float32)
        dO_ = bf16(dO)
        dz2 = ((dO_ @ W3_.T) * (1.0 - h2 * h2)).astype(np.float32)
        dz2_ = bf16(dz2)
        dz1 = ((dz2_ @ W2_.T) * (1.0 - h1 * h1)).astype(np.float32)
        dz1_ = bf16(dz1)
    else:
        dO_ = dO
        dz2 = (dO_ @ W3_.T) * (1.0 - h2 * h2)
        dz2_ = dz2
        dz1 = (dz2_ @ W2_.T) * (1.0 - h1 * h1)
        dz1_ = dz1
    return pack(X_.T @ dz1_, dz1.sum(0), h1_.T @ dz2_, dz2.sum(0), h2_.T @ dO_, np.asarray(dO).sum(0))


def factor(S, l, sig):
    """(w, e, l, s): w_i = 1/sqrt(a_ii), e_c = Delta_c^2 = exp(2 s_c) with s_c
    the sum of sig_f over the faces owned by c (column c of L), and L."""
    s = np.zeros(S.N)
    np.add.at(s, S.p, np.asarray(sig, np.float64))
    return 1.0 / np.sqrt(S.d), np.exp(2.0 * s), np.asarray(l, np.float64), s


def apply(S, w, e, l, r):
    """z = G G^T r with G = D^-1/2 (I + L) Delta."""
    u = w * r
    t = u.copy()
    np.add.at(t, S.p, l * u[S.n])      # (I + L^T) u
    t *= e                             # Delta^2
    y = t.copy()
    np.add.at(y, S.n, l * t[S.p])      # (I + L) t
    return w * y


def dense_factor(S, w, e, l):
    G = np.eye(S.N)
    G[S.n, S.p] = l
    return np.diag(w) @ G @ np.diag(np.sqrt(e))


def kaporin(S, l, sig, grad=True):
    """J and (dJ/dl, dJ/dsig) for the factor Gh = (I + L) Delta of the
    Jacobi-scaled matrix Ah, Delta = exp(s).

    Column c of Gh is Delta_c (1, l_f for the faces f owned by c), so
    tr(Gh^T Ah Gh) = sum_c Delta_c^2 tau_c with tau_c the quadratic form of
    the unscaled column: tau_c = Pc_c + sum_f l_f Pf_f, where
    Pc_c = 1 + sum_f Ah_f l_f and Pf_f = l_f + Ah_f + (3-cycle terms) are the
    entries of Ah times that column."""
    l = np.asarray(l, np.float64)
    sig = np.asarray(sig, np.float64)
    N = S.N
    s = np.zeros(N)
    np.add.at(s, S.p, sig)
    e = np.exp(2.0 * s)
    Pc = np.ones(N)
    np.add.at(Pc, S.p, S.ah * l)
    Pf = l + S.ah
    # 3-cycles: j in U(p), j != n, j adjacent to n
    for k in range(len(l)):
        c, i = S.p[k], S.n[k]
        for k2 in S.upper[c]:
            j = S.n[k2]
            if j == i:
                continue
            kk = S.nbr[i].get(j)
            if kk is not None:
                Pf[k] += S.ah[kk] * l[k2]
    tau = Pc.copy()
    np.add.at(tau, S.p, l * Pf)
    T = float(np.sum(e * tau))
    J = np.log(T / N) - 2.0 / N * np.sum(s)
    if not grad:
        return J, T
    g_l = 2.0 / T * e[S.p] * Pf
    g_s = 2.0 / T * e * tau - 2.0 / N
    return J, T, g_l, g_s[S.p]


def fsai_closed_form(S):
    """FSAI of Ah for the factor (I + L) Delta when no two higher neighbours of
    a cell are adjacent (hexahedral and quadrilateral meshes): l_f = -Ah_f,
    Delta_c = (1 - sum_f Ah_f^2)^-1/2 over the faces owned by c.  Returns
    per-face (l, sig) with sig spread evenly over the faces owned by c."""
    q = np.zeros(S.N)
    np.add.at(q, S.p, S.ah ** 2)
    cnt = np.bincount(S.p, minlength=S.N)
    return -S.ah, (-0.5 * np.log1p(-q))[S.p] / cnt[S.p]


def fsai_dense(S):
    """FSAI by its definition, dense, for small systems: column c of the lower
    factor minimises the Kaporin measure over its pattern {c} and the higher
    neighbours of c, i.e. Gh[Q, c] = y / sqrt(y_c) with Ah[Q, Q] y = e_c."""
    Ah = S.dense() / np.sqrt(np.outer(S.d, S.d))
    G = np.zeros((S.N, S.N))
    for c in range(S.N):
        Q = [c] + sorted(S.n[k] for k in S.upper[c])
        e = np.zeros(len(Q))
        e[0] = 1.0
        y = np.linalg.solve(Ah[np.ix_(Q, Q)], e)
        G[Q, c] = y / np.sqrt(y[0])
    return G


def pcg_iterations(A, b, apply_minv, tol=1e-8, maxit=100000):
    """Iterations of PCG (relative L1 residual) from x = 0; test helper."""
    x = np.zeros_like(b)
    r = b.copy()
    z = apply_minv(r)
    p = z.copy()
    rz = r @ z
    nb = np.abs(b).sum()
    for k in range(maxit):
        q = A @ p
        al = rz / (p @ q)
        x += al * p
        r -= al * q
        if np.abs(r).sum() / nb < tol:
            return k + 1, x
        z = apply_minv(r)
        rz2 = r @ z
        p = z + (rz2 / rz) * p
        rz = rz2
    return maxit, x
