"""gmsh 2.2 ASCII meshes written by this project for the reader's tests:
a jittered hex box and its split into prisms (3D; boundary quadrilaterals,
and triangles for the prisms' caps) and the seeded-Delaunay channel (2D
triangles, boundary lines).  Node labels are not 1..n and boundary elements
are interleaved with the cells, as gmsh files have them."""
import numpy as np

HEX_GMSH = (0, 1, 3, 2, 4, 5, 7, 6)  # Gambit brick order <-> gmsh hexahedron order (an involution)
# faces of a Gambit brick (1-based local vertices)
BRICK_FACES = ((1, 2, 6, 5), (2, 4, 8, 6), (4, 3, 7, 8), (3, 1, 5, 7), (2, 1, 3, 4), (5, 6, 8, 7))


def label(i):
    return 3 * i + 7


def write(path, vx, elems):
    """vx (n, 3); elems: list of (gmsh type, physical tag, 0-based vertex list)."""
    with open(path, "w") as f:
        f.write("$MeshFormat\n2.2 0 8\n$EndMeshFormat\n$Nodes\n%d\n" % len(vx))
        for i, x in enumerate(vx):
            f.write("%d %r %r %r\n" % (label(i), float(x[0]), float(x[1]), float(x[2])))
        f.write("$EndNodes\n$Elements\n%d\n" % len(elems))
        for k, (ty, tag, vl) in enumerate(elems):
            f.write("%d %d 2 %d %d %s\n" % (100 + 2 * k, ty, tag, tag + 1000, " ".join(str(label(v)) for v in vl)))
        f.write("$EndElements\n")


def box_elements(fvk, n, jitter, seed, prisms=False):
    m = fvk.Mesh.box(n, n, n, jitter=jitter, seed=seed)
    vx, ty, vp, vr = m.cells()
    zmax = vx[:, 2].max()
    cells = [vr[vp[c]:vp[c + 1]] - 1 for c in range(m.n_vol)]
    faces = {}
    for c, cv in enumerate(cells):
        for k, fc in enumerate(BRICK_FACES):
            key = tuple(sorted(cv[a - 1] for a in fc))
            faces.setdefault(key, []).append((k, cv))
    bnd = []
    for fl in faces.values():
        if len(fl) != 1:
            continue
        k, a = fl[0]
        if prisms and k in (4, 5):
            # the prisms' caps: the two triangles of the split
            o = 0 if k == 4 else 4
            bnd.append([a[o], a[o + 1], a[o + 3]])
            bnd.append([a[o], a[o + 3], a[o + 2]])
        else:
            bnd.append([a[i - 1] for i in BRICK_FACES[k]])
    elems = []
    rng = np.random.default_rng(seed)
    bq = list(rng.permutation(len(bnd)))
    for c, cv in enumerate(cells):
        if prisms:
            a = cv  # Gambit: 1 (000) 2 (100) 3 (010) 4 (110) 5..8 above
            elems.append((6, 0, [a[0], a[1], a[3], a[4], a[5], a[7]]))
            elems.append((6, 0, [a[0], a[3], a[2], a[4], a[7], a[6]]))
        else:
            elems.append((5, 0, [cv[HEX_GMSH[k]] for k in range(8)]))
        # a boundary face now and then, between the cells
        for _ in range(2):
            if bq:
                fv = bnd[bq.pop()]
                tag = 31 if np.all(vx[fv, 2] == zmax) else 32
                elems.append((3 if len(fv) == 4 else 2, tag, fv))
    for j in bq:
        fv = bnd[j]
        elems.append((3 if len(fv) == 4 else 2, 31 if np.all(vx[fv, 2] == zmax) else 32, fv))
    return m, vx, elems


def channel_elements(n_tri, seed):
    from programs_amd import meshgen
    vx2, ty, vp, vr, bnd = meshgen.channel_delaunay(n_tri, seed=seed)
    vx = np.zeros((len(vx2), 3))
    vx[:, :2] = vx2
    elems = []
    nt = len(ty)
    tri = (vr - 1).reshape(nt, 3)
    b = bnd.reshape(nt, 3)
    # Gambit triangle faces: (1,2), (2,3), (3,1)
    for t in range(nt):
        elems.append((2, 0, list(tri[t])))
        for k, (a, c) in enumerate(((0, 1), (1, 2), (2, 0))):
            if b[t, k]:
                elems.append((1, int(b[t, k]), [tri[t, a], tri[t, c]]))
    return vx, elems
