"""gmsh input and output (host side, no device).

* The reader (fvk_mesh_read_gmsh) against flowlab's own read_gmsh_mesh
  (g++ build of flowlab/src/mesh.cpp, dumps in tests/golden/gmsh_read_*.npz):
  node coordinates bit for bit, the cells' vertex lists (gmsh order
  converted to Gambit's), element numbers and boundary codes, on flowlab's
  duct_3d.msh and on meshes written by this project (hexahedra, prisms,
  triangles).
* A hex box written as gmsh and read back is the box the Gambit path builds:
  connectivity and geometry bit for bit.
* Closed forms on the prism and duct meshes: volumes sum to the domain's,
  every interior face has two cells, the channel's boundary codes.
* The field writers against flowlab's write_gmsh_field_ /
  write_gmsh_vector_field_ output bytes (tests/golden/gmsh_fields.npz).
"""
import gzip
import os
import shutil

import numpy as np
import pytest

from tests.golden import msh_gen
from tests.helpers import ROOT, require_lib

GOLDEN = os.path.join(ROOT, "tests", "golden")


def msh_path(name, tmp_path):
    if name == "duct3d":
        p = tmp_path / "duct_3d.msh"
        with gzip.open(os.path.join(GOLDEN, "duct_3d.msh.gz"), "rb") as f, open(p, "wb") as g:
            shutil.copyfileobj(f, g)
        return str(p)
    return os.path.join(GOLDEN, "msh_%s.msh" % name)


DIM = {1: 1, 2: 2, 3: 2, 4: 3, 5: 3, 6: 3, 7: 3}
HEX = msh_gen.HEX_GMSH


@pytest.mark.parametrize("name", ["duct3d", "box4j", "prism3", "chan400"])
def test_reader_matches_flowlab(name, tmp_path):
    fvk = require_lib()
    ref = np.load(os.path.join(GOLDEN, "gmsh_read_%s.npz" % name))
    m = fvk.Mesh.read_gmsh(msh_path(name, tmp_path))
    vx, ty, vp, vr = m.cells()
    nd = m.n_dx
    assert np.array_equal(vx[:, :nd], ref["x"][:, :nd])
    pos = {int(l): i for i, l in enumerate(ref["vrt_idx"])}
    tp, tl, ep, el = ref["tag_ptr"], ref["tag_lst"], ref["vrt_ptr"], ref["vrt_lst"]
    gtype = np.array([tl[tp[e]] for e in range(len(tp) - 1)])
    dims = np.array([DIM[t] for t in gtype])
    cells = np.nonzero(dims == nd)[0]
    assert m.n_vol == len(cells)
    assert np.array_equal(m.elm_idx(), ref["elm_idx"][cells])
    for c, e in enumerate(cells):
        g = [pos[int(v)] + 1 for v in el[ep[e]:ep[e + 1]]]
        want = [g[HEX[k]] for k in range(8)] if gtype[e] == 5 else g
        assert list(vr[vp[c]:vp[c + 1]]) == want, (c, e)
    # boundary faces carry the physical tag of the boundary element with
    # the same vertices; every boundary element is a face of the mesh
    t = m.topology()
    bf = np.nonzero(t["fce_in"] == 0)[0]
    assert len(bf) == m.n_bf
    codes = {}
    for e in np.nonzero(dims == nd - 1)[0]:
        codes[tuple(sorted(pos[int(v)] for v in el[ep[e]:ep[e + 1]]))] = int(tl[tp[e] + 1])
    assert len(codes) == m.n_bf
    assert sorted(np.unique(t["fce_bty"][bf])) == sorted(set(codes.values()))
    assert np.bincount(t["fce_bty"][bf]).sum() == len(codes)


def test_gmsh_box_is_the_gambit_box(tmp_path):
    fvk = require_lib()
    n, jit, seed = 5, 0.2, 0x5EED00C5
    m0, vx, el = msh_gen.box_elements(fvk, n, jit, seed)
    p = str(tmp_path / "box.msh")
    msh_gen.write(p, vx, el)
    m = fvk.Mesh.read_gmsh(p)
    t0, t1 = m0.topology(), m.topology()
    for k in t0:
        assert np.array_equal(t0[k], t1[k]), k
    g0, g1 = m0.geometry(), m.geometry()
    for k in g0:
        assert np.array_equal(g0[k], g1[k]), k


def test_prism_and_duct_meshes_are_closed(tmp_path):
    fvk = require_lib()
    # unjittered prisms: planar faces, so each cell's face vectors close
    _, vx, el = msh_gen.box_elements(fvk, 3, 0.0, 0, prisms=True)
    p = str(tmp_path / "prisms.msh")
    msh_gen.write(p, vx, el)
    m = fvk.Mesh.read_gmsh(p)
    assert m.n_vol == 2 * 27
    g, t = m.geometry(), m.topology()
    assert abs(g["vol_v"].sum() - 1.0) < 1e-12
    inner = t["fce_in"] > 0
    assert inner.sum() + m.n_bf == m.n_fce
    # each cell's outward face vectors close (planar faces)
    s = np.zeros((m.n_vol, 3))
    sn = g["fce_s"][:, None] * g["fce_n"]
    np.add.at(s, t["fce_ip"] - 1, sn)
    np.add.at(s, t["fce_in"][inner] - 1, -sn[inner])
    assert np.max(np.abs(s)) < 1e-12
    d = fvk.Mesh.read_gmsh(msh_path("duct3d", tmp_path))
    gd = d.geometry()
    assert d.n_vol > 20000 and d.n_dx == 3
    assert np.all(gd["fce_s"] > 0)
    assert len(set(np.unique(d.topology()["fce_bty"][d.topology()["fce_in"] == 0]))) >= 2


def test_code_map(tmp_path):
    fvk = require_lib()
    m = fvk.Mesh.read_gmsh(msh_path("chan400", tmp_path), code_map={32: 33, 11: 12})
    t = m.topology()
    codes = set(np.unique(t["fce_bty"][t["fce_in"] == 0]))
    assert codes == {12, 21, 33}


def test_field_writers_match_flowlab(tmp_path):
    fvk = require_lib()
    z = np.load(os.path.join(GOLDEN, "gmsh_fields.npz"))
    for name in ("scalar", "vec2", "vec3"):
        ref = bytes(z[name + "_bytes"]).decode()
        fname = str(tmp_path / (name + ".msh"))
        phi = z[name + "_phi"]
        nd = int(z[name + "_n_dim"])
        fvk.write_gmsh_field(fname, z[name + "_idx"], phi if nd == 1 else phi.reshape(-1, nd))
        got = open(fname).read()
        # the sixth line is the view name: the file name as written
        rl, gl = ref.split("\n"), got.split("\n")
        assert rl[5] == str(z[name + "_fname"]) and gl[5] == fname
        rl[5] = gl[5] = ""
        assert rl == gl, name
