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


def box_records(n, jitter=0.0, seed=0x5EED0006, lid=31, wall=32):
    """Lid-driven cavity as gmsh element records for the flowlab path: the
    unit cube in n^3 hexahedra (gmsh vertex order), interior vertices moved by
    U(-jitter h, jitter h) per coordinate (numpy's PCG64 seeded with seed),
    boundary quadrilaterals tagged lid (z = 1) or wall.  Returns
    (x3, e_ty, e_tag, e_vptr, e_vrt) for flowlab.FlMesh.from_records:
    boundary elements first, then the cells in x-fastest order."""
    m = n + 1
    g = np.arange(m, dtype=np.float64) / n
    z, y, x = np.meshgrid(g, g, g, indexing="ij")
    x3 = np.stack([x.ravel(), y.ravel(), z.ravel()], 1)
    if jitter > 0:
        idx = np.arange(m)
        kz, ky, kx = np.meshgrid(idx, idx, idx, indexing="ij")
        inner = ((kx > 0) & (kx < n) & (ky > 0) & (ky < n) & (kz > 0) & (kz < n)).ravel()
        rng = np.random.default_rng(seed)
        d = rng.uniform(-jitter / n, jitter / n, size=x3.shape)
        x3[inner] += d[inner]
    node = lambda i, j, k: (k * m + j) * m + i + 1
    i, j, k = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    i, j, k = (a.transpose(2, 1, 0).ravel() for a in (i, j, k))  # x fastest
    hexes = np.stack([node(i, j, k), node(i + 1, j, k), node(i + 1, j + 1, k), node(i, j + 1, k),
                      node(i, j, k + 1), node(i + 1, j, k + 1), node(i + 1, j + 1, k + 1), node(i, j + 1, k + 1)], 1)
    a, b = np.meshgrid(np.arange(n), np.arange(n), indexing="ij")
    a, b = a.ravel(), b.ravel()
    quads, tags = [], []
    for side in range(6):
        ax, hi = side // 2, side % 2
        c = np.full_like(a, n if hi else 0)
        if ax == 0:
            q = [node(c, a, b), node(c, a + 1, b), node(c, a + 1, b + 1), node(c, a, b + 1)]
        elif ax == 1:
            q = [node(a, c, b), node(a + 1, c, b), node(a + 1, c, b + 1), node(a, c, b + 1)]
        else:
            q = [node(a, b, c), node(a + 1, b, c), node(a + 1, b + 1, c), node(a, b + 1, c)]
        quads.append(np.stack(q, 1))
        tags.append(np.full(len(a), lid if (ax == 2 and hi) else wall))
    quads = np.concatenate(quads)
    tags = np.concatenate(tags)
    nq, nh = len(quads), len(hexes)
    e_ty = np.concatenate([np.full(nq, 3), np.full(nh, 5)]).astype(np.int32)
    e_tag = np.concatenate([tags, np.zeros(nh)]).astype(np.int32)
    e_vptr = np.concatenate([[0], np.cumsum(np.concatenate([np.full(nq, 4), np.full(nh, 8)]))]).astype(np.int32)
    e_vrt = np.concatenate([quads.ravel(), hexes.ravel()]).astype(np.int32)
    return x3, e_ty, e_tag, e_vptr, e_vrt


def box_bnd(nx, ny, nz, lid=31, wall=32):
    """Boundary codes per (cell, local Gambit face) of fvk.Mesh.box's cells
    (cell i + nx (j + ny k), hex faces as gambit.vrt numbers them: 1 y-, 2
    x+, 3 y+, 4 x-, 5 z-, 6 z+; the lid is z+).  Flat, 6 per cell."""
    k, j, i = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    b = np.zeros((nz, ny, nx, 6), np.int32)
    b[..., 0][j == 0] = wall
    b[..., 1][i == nx - 1] = wall
    b[..., 2][j == ny - 1] = wall
    b[..., 3][i == 0] = wall
    b[..., 4][k == 0] = wall
    b[..., 5][k == nz - 1] = lid
    return b.reshape(-1)


def write_neu(path, n_dx, vx, v_ty, v_vptr, v_vrt, bnd, faces_per_cell):
    """Gambit neutral file in the fixed formats the reference's grd_read
    expects (grd_read.f90:28,42,59,83,85): nodal coordinates (E20.12),
    elements (7 vertices a line), one boundary-condition set per code (its
    name from grd_read.f90:115-128's rule: 11 inlet_1, 21 outlet_1, 31
    wall_1, ...).  v_vptr 0-based offsets, v_vrt 1-based vertex ids; bnd the
    code of each (cell, local face), faces_per_cell[i] of them for cell i."""
    nv = len(v_ty)
    fptr = np.concatenate([[0], np.cumsum(faces_per_cell)]).astype(np.int64)
    codes = sorted(set(int(c) for c in bnd if c > 0))
    kinds = {1: "inlet", 2: "outlet", 3: "wall", 4: "symmetry", 5: "pressure"}
    L = ["        CONTROL INFO 2.3.16", "** GAMBIT NEUTRAL FILE", "generated", "PROGRAM:  fvk", "",
         "     NUMNP     NELEM     NGRPS    NBSETS     NDFCD     NDFVL",
         "".join(" %9d" % v for v in (len(vx), nv, 1, len(codes), n_dx, n_dx)), "ENDOFSECTION",
         "   NODAL COORDINATES 2.3.16"]
    for k, x in enumerate(vx):
        L.append("%10d " % (k + 1) + "".join("%20.12E" % x[d] for d in range(n_dx)))
    L += ["ENDOFSECTION", "      ELEMENTS/CELLS 2.3.16"]
    for i in range(nv):
        vs = v_vrt[v_vptr[i]:v_vptr[i + 1]]
        head = "%8d %2d %2d " % (i + 1, v_ty[i], len(vs))
        first, rest = vs[:7], vs[7:]
        L.append(head + "".join("%8d" % v for v in first))
        while len(rest):
            L.append(" " * 15 + "".join("%8d" % v for v in rest[:7]))
            rest = rest[7:]
    L.append("ENDOFSECTION")
    for c in codes:
        ent = [(i + 1, v_ty[i], k + 1) for i in range(nv) for k, cc in enumerate(bnd[fptr[i]:fptr[i + 1]]) if cc == c]
        L.append(" BOUNDARY CONDITIONS 2.3.16")
        L.append("%32s%10d%10d%10d%10d" % (kinds[c // 10] + "_%d" % (c % 10), 1, len(ent), 0, 6))
        L += ["%10d%5d%5d" % e for e in ent]
        L.append("ENDOFSECTION")
    with open(path, "w") as fh:
        fh.write("\n".join(L) + "\n")
