"""Synthetic unstructured meshes for the workloads: a 2D channel of seeded
Delaunay triangles (Gambit type-3 cells, one cell deep in the reference's 2D
mode) with inlet, outlet and wall boundary codes.

The boundary codes follow the reference's naming rule (grd_read.f90:115-128):
inlet_1 -> 11, outlet_1 -> 21, wall_2 -> 32.
"""
import numpy as np


def channel_delaunay(n_tri, length=10.0, height=1.0, jitter=0.3, seed=0x5EED0005):
    """About n_tri triangles on [0, length] x [0, height]: a jittered grid of
    points (interior points moved by U(-jitter, jitter) of the spacing in x
    and y, boundary points along their side only) triangulated by Delaunay.
    Returns (vx (n, 2), v_ty, v_vptr, v_vrt (1-based), bnd) for
    fvk.Mesh.from_cells(2, ...)."""
    from scipy.spatial import Delaunay
    ny = max(2, int(round(np.sqrt(n_tri * height / (2.0 * length)))))
    nx = max(2, int(round(ny * length / height)))
    hx, hy = length / nx, height / ny
    rng = np.random.default_rng(seed)
    i, j = np.meshgrid(np.arange(nx + 1), np.arange(ny + 1), indexing="xy")
    i, j = i.ravel(), j.ravel()
    x = i * hx
    y = j * hy
    ux = rng.uniform(-jitter, jitter, x.shape) * hx
    uy = rng.uniform(-jitter, jitter, y.shape) * hy
    inner_x = (i > 0) & (i < nx)
    inner_y = (j > 0) & (j < ny)
    x = np.where(inner_x, x + ux, x)
    y = np.where(inner_y, y + uy, y)
    pts = np.stack([x, y], 1)
    tri = Delaunay(pts)
    simp = tri.simplices.astype(np.int64)
    nb = tri.neighbors
    a = pts[simp]
    area = (a[:, 1, 0] - a[:, 0, 0]) * (a[:, 2, 1] - a[:, 0, 1]) - (a[:, 2, 0] - a[:, 0, 0]) * (a[:, 1, 1] - a[:, 0, 1])
    keep = np.abs(area) > 1e-12 * hx * hy
    simp, nb = simp[keep], nb[keep]
    flip = area[keep] < 0
    simp[flip] = simp[flip][:, [0, 2, 1]]
    nb = nb.copy()
    nb[flip] = nb[flip][:, [0, 2, 1]]
    nt = len(simp)
    # Gambit triangle faces: (1,2) opposite vertex 3, (2,3) opposite 1, (3,1)
    # opposite 2
    bnd = np.zeros((nt, 3), np.int32)
    for k, (va, vb, opp) in enumerate(((0, 1, 2), (1, 2, 0), (2, 0, 1))):
        edge_b = nb[:, opp] < 0
        mid = 0.5 * (pts[simp[:, va]] + pts[simp[:, vb]])
        code = np.full(nt, 32, np.int32)
        code[mid[:, 0] < 0.25 * hx] = 11
        code[mid[:, 0] > length - 0.25 * hx] = 21
        bnd[:, k] = np.where(edge_b, code, 0)
    v_ty = np.full(nt, 3, np.int32)
    v_vptr = (np.arange(nt + 1) * 3).astype(np.int32)
    v_vrt = (simp.ravel() + 1).astype(np.int32)
    return pts, v_ty, v_vptr, v_vrt, bnd.ravel()
