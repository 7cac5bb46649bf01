"""Fortran drop-in: the reference's own host program (grd_read / grd_conn /
grd_geom and the golden-vector driver) with the hot-path modules solv_*_m,
grd_solv_m and cfd_setup_m replaced by the ISO_C_BINDING modules of
programs_amd/fortran, which call libfvk.so on the GPU.  Its dump must match
the golden vectors the unmodified reference wrote (bars as in
test_gpu_parity: exact where row-local, else 1e-10 or the reference's own
summation-order sensitivity).

The executable is built in the development container by
programs_amd/fortran/build_dropin.sh (it links reference objects, so it is
not part of the repository) and travels with the tree.  The test fails when
it is absent or was built from other sources than the tree's (the build
writes a hash of its Fortran sources beside it).
"""
import hashlib
import os
import subprocess

import numpy as np
import pytest

from oracle import fvo
from tests import golden_cases as G
from tests.golden.make_golden import write_neu
from tests.helpers import ROOT, assert_rel
from tests.test_gpu_parity import b3_for, sens_tol, oracle_flow_fields

pytestmark = pytest.mark.gpu

DRIVER = os.path.join(ROOT, "oracle", "_ref", "dropin", "fvk_ref_driver")
STAMP = os.path.join(ROOT, "oracle", "_ref", "dropin", "sources.sha256")
SOURCES = ["programs_amd/fortran/fvk_c.f90", "programs_amd/fortran/fvk_solv.f90",
           "programs_amd/fortran/fvk_grd_solv.f90", "programs_amd/fortran/fvk_cfd_setup.f90",
           "oracle/ref/fvk_ref_driver.f90"]


def sources_hash():
    """The hash build_dropin.sh writes: sha256 of `sha256sum` over SOURCES."""
    lines = "".join("%s  %s\n" % (hashlib.sha256(open(os.path.join(ROOT, s), "rb").read()).hexdigest(), s)
                    for s in SOURCES)
    return hashlib.sha256(lines.encode()).hexdigest()


def test_dropin_executable_is_current():
    assert os.path.exists(DRIVER), "drop-in executable missing: run programs_amd/fortran/build_dropin.sh"
    assert os.path.exists(STAMP), "drop-in executable has no source stamp: rebuild it"
    assert open(STAMP).read().strip() == sources_hash(), \
        "drop-in executable was built from other Fortran sources: rebuild it"


@pytest.mark.parametrize("name", G.CASES)
def test_fortran_dropin_matches_reference(name, tmp_path):
    test_dropin_executable_is_current()
    rec = G.load(name)
    spec = rec["spec"]
    src = G.mesh_inputs(rec)
    if isinstance(src, str):
        neu = src
    else:
        raw = dict(n_dx=src["n_dx"], vx=src["vx"], v_ty=src["v_ty"], v_vptr=src["v_vptr"],
                   v_vrt=src["v_vrt"], bnd=src["bnd"], bnd_ptr=rec["mesh_bnd_ptr"])
        neu = str(tmp_path / (name + ".neu"))
        write_neu(neu, raw)
    out = str(tmp_path / "dump.bin")
    args = [DRIVER, os.path.join(G.GOLDEN, "gambit.vrt"), neu, out, str(spec["n_outer"]), repr(spec["rho"]),
            repr(spec["mu"]), repr(spec["urf_p"])]
    for code, u in spec["bc"]:
        args += [str(code)] + [repr(float(v)) for v in u]
    env = dict(os.environ)
    if spec["variant"] == "relaxed":
        env.update(FVK_SIMPLER="0", FVK_IMAT_P="0")
    print("drop-in case", name, flush=True)
    p = subprocess.run(args, capture_output=True, text=True, env=env, cwd=str(tmp_path), timeout=240)
    assert p.returncode == 0, p.stdout[-2000:] + p.stderr[-2000:]
    d = fvo.read_dump(out)
    h = fvo.parse_history(p.stdout)
    # mesh and gradients: exact
    for k in ("fce_ip", "fce_in", "fce_ib", "fce_lst", "ng_lst"):
        assert np.array_equal(d[k], rec["d_" + k]), k
    for k in ("grad0_1", "grad0_2", "grad0_3", "grad2_1", "grad2_2", "grad2_3"):
        assert np.array_equal(d[k], rec["d_" + k]), k
    # flow fields and histories
    tol = sens_tol(lambda: oracle_flow_fields(rec), (name, "simple", 0))
    assert_rel(h["reduc"], rec["h_reduc"], tol, "reduc: history")
    assert_rel(h["redpp"], rec["h_redpp"], tol, "redpp: history")
    assert_rel(d["hist_REDUC"], rec["d_hist_REDUC"], tol, "REDUC history")
    for k in ("u_cell", "u_bnd", "p", "dpdx", "spd", "frc_dns", "cs", "pp_pl", "pp_ngp"):
        assert_rel(d[k], rec["d_" + k], tol, k)
    # stand-alone solver calls made through the drop-in modules, on the
    # matrices the drop-in assembled itself after n_outer iterations: those
    # match the reference's to the fields' bar, and solv_precond /
    # solv_direct on them (row-local) equal the oracle's on the same matrix
    # bit for bit; the short solves match the oracle's to the bar
    for k in ("mom_pl", "mom_ngp", "mom_ngn"):
        assert_rel(d[k], rec["d_" + k], tol, k)
    om = G.oracle_mesh(rec)
    try:
        nv, nd = om.m.n_vol, om.m.n_dx
        b3 = b3_for(nv, nd)
        A = fvo.OSys(om, d["mom_pl"], d["mom_ngp"], d["mom_ngn"])
        P = fvo.OSys(om, d["pp_pl"], d["pp_ngp"], d["pp_ngp"])
        bb = b3[:, 0] - np.cumsum(b3[:, 0])[-1] / nv
        for imat in (0, 1, 2):
            c, _ = A.precond(imat)
            assert np.array_equal(d["mom_c%d" % imat], c), ("mom precond", imat)
            assert np.array_equal(d["mom_dir%d" % imat], A.direct(imat, b3[:, 0])), ("mom direct", imat)
            c, _ = P.precond(imat)
            assert np.array_equal(d["pp_c%d" % imat], c), ("pp precond", imat)
            assert np.array_equal(d["pp_dir%d" % imat], P.direct(imat, bb)), ("pp direct", imat)
        tb = sens_tol(lambda: A.solve("bicg", 0, 0.1, 5, b3)[0], (name, "bicg5r", 0))
        xb = A.solve("bicg", 0, 0.1, 5, b3)[0]
        assert_rel(d["mom_bicgr_x0"], xb.ravel(), tb, "bicg reduc 0.1 jacobi")
        tc = sens_tol(lambda: P.solve("cg", 0, 0.1, 20, bb, not_r0=True)[0], (name, "cg20r", 0))
        xc = P.solve("cg", 0, 0.1, 20, bb, not_r0=True)[0]
        assert_rel(d["pp_cgr_x0"], xc.ravel(), tc, "cg reduc 0.1 jacobi")
    finally:
        om.free()
