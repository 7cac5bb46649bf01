"""Tecplot output of the flow's fields on the GPU path against the files
the reference's own writers wrote (spray-simulation/cfd_outp.f90 eqn_outp
and prp_outp(p=.true.), flang build, tests/golden/tecplot_*.dat made by
tests/golden/make_tecplot.py): the flow's state is set to the reference's
fields (read back exactly from the files: the writers print the shortest
digits that read back to the same double), written by fvk_eqn_outp /
fvk_prp_outp_p, and compared byte for byte.  A gmsh ElementData file of the
same fields reads back to the same values to %g precision."""
import os

import numpy as np
import pytest

from tests import golden_cases as G
from tests.helpers import ROOT, require_lib
from tests.test_gpu_parity import product_mesh

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(ROOT, "tests", "golden")


def parse(path, n_vrt, n_vol, n_dx, ncomp):
    lines = open(path).read().split("\n")
    k = 4 + n_dx * n_vrt
    vals = np.array([float(x) for x in lines[k:k + ncomp * n_vol]])
    return vals.reshape(ncomp, n_vol).T


@pytest.mark.parametrize("name", ["cav5_shipped", "chan2d_shipped", "box8j_shipped"])
def test_tecplot_files_match_reference_bytes(name, tmp_path):
    fvk = require_lib()
    rec = G.load(name)
    m = product_mesh(rec)
    ctx = fvk.grd_solv(m)
    fl = fvk.Flow(ctx, fvk.flow_cfg())
    nd = m.n_dx
    eqn = os.path.join(GOLDEN, "tecplot_%s_eqn.dat" % name)
    prp = os.path.join(GOLDEN, "tecplot_%s_prp.dat" % name)
    u = parse(eqn, m.n_vrt, m.n_vol, nd, nd)
    p = parse(prp, m.n_vrt, m.n_vol, nd, 1)[:, 0]
    fl.set_state(u_cell=u, p=p)
    path = fvk.eqn_outp(m, fl, str(tmp_path))
    assert open(path, "rb").read() == open(eqn, "rb").read()
    path = fvk.prp_outp_p(m, fl, str(tmp_path))
    assert open(path, "rb").read() == open(prp, "rb").read()
    # the same fields as gmsh ElementData
    gm = str(tmp_path / "vel.msh")
    fvk.write_gmsh_field(gm, m.elm_idx(), u)
    rows = open(gm).read().split("\n")[13:13 + m.n_vol]
    back = np.array([[float(x) for x in r.split()[1:1 + nd]] for r in rows])
    assert np.allclose(back, u, rtol=1e-5, atol=1e-300)
    fl.destroy()
    ctx.destroy()
