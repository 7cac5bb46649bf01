"""Connectivity and geometry built on the GPU (fvk_mesh_set_build_device)
against the host build (which is bitwise the reference's grd_conn /
grd_geom, tests/test_mesh.py): every topology and geometry array equal bit
for bit, on the golden fixtures (hexahedra, 2D quadrilaterals, triangles),
jittered boxes, the shuffled Delaunay channel and a prism mesh read from
gmsh."""
import numpy as np
import pytest

from programs_amd import meshgen
from tests import golden_cases as G
from tests.golden import msh_gen
from tests.helpers import require_lib
from tests.test_gpu_parity import product_mesh

pytestmark = pytest.mark.gpu


def both(fvk, make):
    fvk.mesh_build_device(-1)
    a = make()
    fvk.mesh_build_device(0)
    try:
        b = make()
    finally:
        fvk.mesh_build_device(-1)
    return a, b


def same(a, b):
    assert (a.n_vol, a.n_fce, a.n_bf, a.n_slot) == (b.n_vol, b.n_fce, b.n_bf, b.n_slot)
    ta, tb = a.topology(), b.topology()
    for k in ta:
        assert np.array_equal(ta[k], tb[k]), k
    ga, gb = a.geometry(), b.geometry()
    for k in ga:
        assert np.array_equal(ga[k], gb[k]), k


@pytest.mark.parametrize("name", G.CASES)
def test_golden_meshes(name):
    fvk = require_lib()
    rec = G.load(name)
    same(*both(fvk, lambda: product_mesh(rec)))


@pytest.mark.parametrize("n,jit", [(7, 0.0), (9, 0.2), (24, 0.1)])
def test_boxes(n, jit):
    fvk = require_lib()
    same(*both(fvk, lambda: fvk.Mesh.box(n, n, n, jitter=jit, seed=0x5EED00D1)))
    same(*both(fvk, lambda: fvk.Mesh.box(n, n + 3, 0, jitter=jit, seed=0x5EED00D2)))


def test_channel_and_prisms(tmp_path):
    fvk = require_lib()
    args = meshgen.channel_delaunay(20000, seed=0x5EED00D3)
    same(*both(fvk, lambda: fvk.Mesh.from_cells(2, *args)))
    _, vx, el = msh_gen.box_elements(fvk, 6, 0.15, 0x5EED00D4, prisms=True)
    p = str(tmp_path / "prisms.msh")
    msh_gen.write(p, vx, el)
    same(*both(fvk, lambda: fvk.Mesh.read_gmsh(p)))
