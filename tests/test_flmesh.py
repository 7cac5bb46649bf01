"""The flowlab-layout mesh builder (fvk_fl_mesh_*, host only): counts,
ordering and orientation rules, and gmsh file vs in-memory records."""
import gzip
import os
import shutil

import numpy as np
import pytest

from tests.helpers import require_lib

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def fl():
    require_lib()
    from programs_amd import flowlab
    return flowlab


def face_normal(x, fv):
    """geometry_m.f90:174-204: fan of cross products from the first vertex."""
    v0 = x[fv[0]]
    s = np.zeros(3)
    e1 = None
    for k in range(1, len(fv)):
        e = x[fv[k]] - v0
        if e1 is not None:
            s = s + 0.5 * np.cross(e1, e)
        e1 = e
    return s


def check(m):
    nf, nb, n = m.n_fce, m.n_bnd, m.n_dmn
    fe = m.fce_elm
    bnd = fe[:, 1] < 0
    # boundary faces first, each boundary element once
    k = int(bnd.sum())
    assert k == nb and bnd[:k].all() and not bnd[k:].any()
    assert sorted((-fe[:k, 1]).tolist()) == list(range(1, nb + 1))
    # owners and neighbours are domain elements; interior pairs unique
    assert ((fe[:, 0] > nb) & (fe[:, 0] <= nb + n)).all()
    pairs = {tuple(sorted(p)) for p in fe[k:].tolist()}
    assert len(pairs) == nf - k
    # the matrix pattern: diagonal, interior neighbours, columns ascending
    for c in range(n):
        cols = m.ng_col[m.ng_ptr[c] - 1: m.ng_ptr[c + 1] - 1]
        assert (np.diff(cols) > 0).all() and c + 1 in cols
    assert m.nnz == n + 2 * (nf - k)
    # every face's normal points out of its owner
    x = m.x
    xc = np.array([x[m.elm_vrt[m.elm_vptr[c] - 1: m.elm_vptr[c + 1] - 1] - 1].mean(0) for c in range(n)])
    for f in range(nf):
        fv = m.fce_vrt[m.fce_vptr[f] - 1: m.fce_vptr[f + 1] - 1] - 1
        nv = face_normal(x, fv)
        xf = x[fv].mean(0)
        assert np.dot(nv, xf - xc[fe[f, 0] - nb - 1]) > 0, f


def test_box_records_counts_and_rules():
    F = fl()
    from programs_amd import meshgen
    N = 5
    m = F.FlMesh.from_records(3, *meshgen.box_records(N, jitter=0.15))
    assert (m.n_dmn, m.n_bnd, m.n_fce) == (N ** 3, 6 * N * N, 3 * N * N * (N + 1))
    check(m)
    # interior faces in reverse order of discovery: the last face belongs to
    # the first cell
    assert m.fce_elm[-1, 0] == m.n_bnd + 1
    # lid tags
    assert set(m.bnd_tag.tolist()) == {31, 32} and (m.bnd_tag == 31).sum() == N * N


@pytest.mark.parametrize("fname,nd", [("msh_box4j.msh", 3), ("msh_prism3.msh", 3), ("msh_chan400.msh", 2)])
def test_gmsh_meshes_follow_the_rules(fname, nd):
    m = fl().FlMesh.read_gmsh(os.path.join(GOLDEN, fname), nd)
    if nd == 3:
        check(m)
    else:
        m3 = m
        assert (m3.fce_elm[: m3.n_bnd, 1] < 0).all()


def test_duct_3d_counts(tmp_path):
    p = os.path.join(str(tmp_path), "duct_3d.msh")
    with gzip.open(os.path.join(GOLDEN, "duct_3d.msh.gz")) as a, open(p, "wb") as b:
        shutil.copyfileobj(a, b)
    m = fl().FlMesh.read_gmsh(p, 3)
    assert (m.n_vrt, m.n_bnd, m.n_dmn, m.n_fce, m.nnz) == (27650, 6550, 24258, 76049, 163256)
    assert {int(t): int(c) for t, c in zip(*np.unique(m.bnd_tag, return_counts=True))} == {11: 311, 21: 311,
                                                                                           31: 2280, 32: 3648}


def test_file_and_records_agree(tmp_path):
    F = fl()
    from programs_amd import meshgen
    x3, ty, tag, vp, vr = meshgen.box_records(3, jitter=0.1)
    p = os.path.join(str(tmp_path), "box3.msh")
    with open(p, "w") as f:
        f.write("$MeshFormat\n2.2 0 8\n$EndMeshFormat\n$Nodes\n%d\n" % len(x3))
        for i, x in enumerate(x3):
            f.write("%d %r %r %r\n" % (i + 1, float(x[0]), float(x[1]), float(x[2])))
        f.write("$EndNodes\n$Elements\n%d\n" % len(ty))
        for e in range(len(ty)):
            f.write("%d %d 2 %d 0 %s\n" % (e + 1, ty[e], tag[e], " ".join(str(v) for v in vr[vp[e]:vp[e + 1]])))
        f.write("$EndElements\n")
    a = F.FlMesh.read_gmsh(p, 3)
    b = F.FlMesh.from_records(3, x3, ty, tag, vp, vr)
    for k in ("x", "elm_vptr", "elm_vrt", "bnd_tag", "fce_elm", "fce_vptr", "fce_vrt", "ng_ptr", "ng_col"):
        assert np.array_equal(getattr(a, k), getattr(b, k)), k


def test_bad_meshes_are_errors(tmp_path):
    F = fl()
    from programs_amd.fvk import FvkError
    x3 = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [1, 1, 0]], float)
    # two triangles, no boundary lines: faces without neighbour or boundary
    with pytest.raises(FvkError, match="neither a neighbour nor a boundary element"):
        F.FlMesh.from_records(2, x3, [2, 2], [0, 0], [0, 3, 6], [1, 2, 3, 2, 4, 3])
    # a boundary line on an interior edge
    with pytest.raises(FvkError, match="interior face"):
        F.FlMesh.from_records(2, x3, [1, 2, 2], [31, 0, 0], [0, 2, 5, 8], [2, 3, 1, 2, 3, 2, 4, 3])
