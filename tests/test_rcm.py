"""Reverse Cuthill-McKee cell order (fvk_mesh_rcm, the order a context
built with FVK_CTX_RCM uses): a valid permutation, deterministic, a
bandwidth no worse than the textbook algorithm's (scipy's
reverse_cuthill_mckee) and far below a shuffled numbering's, and within the
16-bit column offsets the FP32 inner solve needs on a shuffled channel.
Host only (no device)."""
import numpy as np
import pytest

from programs_amd import meshgen
from tests.helpers import require_lib


def edges(m):
    t = m.topology()
    inner = t["fce_in"] > 0
    return t["fce_ip"][inner] - 1, t["fce_in"][inner] - 1


def bandwidth(order, ip, inn):
    """max |row(i) - row(j)| over the graph's edges, order[k] = cell (0-based)
    at row k."""
    row = np.empty(len(order), np.int64)
    row[order] = np.arange(len(order))
    return int(np.max(np.abs(row[ip] - row[inn])))


def scipy_rcm(n, ip, inn):
    from scipy.sparse import coo_matrix
    from scipy.sparse.csgraph import reverse_cuthill_mckee
    A = coo_matrix((np.ones(2 * len(ip)), (np.r_[ip, inn], np.r_[inn, ip])), shape=(n, n)).tocsr()
    return np.asarray(reverse_cuthill_mckee(A, symmetric_mode=True))


@pytest.mark.parametrize("n,jit", [(10, 0.0), (12, 0.2)])
def test_rcm_on_box(n, jit):
    fvk = require_lib()
    m = fvk.Mesh.box(n, n, n, jitter=jit, seed=0x5EED0071)
    p = m.rcm()
    assert np.array_equal(np.sort(p), np.arange(1, m.n_vol + 1))
    assert np.array_equal(m.rcm(), p)
    assert np.array_equal(fvk.Mesh.box(n, n, n, jitter=jit, seed=0x5EED0071).rcm(), p)
    ip, inn = edges(m)
    bw = bandwidth(p - 1, ip, inn)
    assert bw <= n * n  # natural order: exactly n^2
    assert bandwidth(np.arange(m.n_vol), ip, inn) == n * n
    assert bw <= bandwidth(scipy_rcm(m.n_vol, ip, inn), ip, inn)


def test_rcm_restores_locality_of_a_shuffled_channel():
    """A 40,000-triangle channel with its cells numbered at random: RCM
    brings every neighbour within 32767 rows (the FP32 inner solve's 16-bit
    column offsets) and matches the textbook algorithm's bandwidth to 10%."""
    fvk = require_lib()
    vx, ty, vp, vr, bnd = meshgen.channel_delaunay(40000, seed=0x5EED0072)
    nt = len(ty)
    rng = np.random.default_rng(5)
    sh = rng.permutation(nt)
    vr3 = vr.reshape(nt, 3)[sh].ravel()
    bnd3 = bnd.reshape(nt, 3)[sh].ravel()
    m = fvk.Mesh.from_cells(2, vx, ty, vp, vr3, bnd3)
    ip, inn = edges(m)
    assert bandwidth(np.arange(nt), ip, inn) > 32767
    p = m.rcm()
    assert np.array_equal(np.sort(p), np.arange(1, nt + 1))
    bw = bandwidth(p - 1, ip, inn)
    assert bw <= 32767
    assert bw <= 1.1 * bandwidth(scipy_rcm(nt, ip, inn), ip, inn)
    assert bw < 1000  # a 2D strip about 45 cells high
