"""flowlab's serial CSR SIMPLE on the GPU (programs_amd.flowlab, through the C
ABI) against golden vectors from flowlab's own Fortran modules built with
flang (tests/golden/make_flowlab.py, fl_*.npz).

Bars:
  * bitwise for the geometry and least-squares matrices (geometry_m.f90:97-660),
    the least-squares gradients (:665-770) and the assembled momentum and
    pressure-correction systems (transport_m.f90:58-348, 430-629) on the
    reference's own inputs: the GPU gathers each cell's face terms in face
    order, the reference's own order of additions;
  * 1e-9 relative (max-norm) downstream of the Krylov solvers' dot products
    and L1 norms (solv_m.f90), which the GPU sums in a fixed tree and the
    reference sequentially: the solver outputs, and the residual norms of
    every iteration and the fields after the reference's full run (up to 30
    outer iterations; the iteration count must be the reference's).
"""
import ast
import gzip
import os
import shutil
import sys

import numpy as np
import pytest

from tests.helpers import relerr, require_lib

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
CASES = sorted(f[:-4] for f in os.listdir(GOLDEN) if f.startswith("fl_") and f.endswith(".npz"))
SMALL_CASES = [c for c in CASES if c != "fl_duct3d"]


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def load(name, tmp_path):
    z = np.load(os.path.join(GOLDEN, name + ".npz"), allow_pickle=False)
    rec = {k: z[k] for k in z.files}
    spec = ast.literal_eval(str(rec["case"]))
    p = os.path.join(GOLDEN, spec["mesh"])
    if p.endswith(".gz"):
        q = os.path.join(str(tmp_path), spec["mesh"][:-3])
        with gzip.open(p) as a, open(q, "wb") as b:
            shutil.copyfileobj(a, b)
        p = q
    return rec, spec, p


def setup(name, tmp_path):
    require_lib()
    from programs_amd import flowlab as fl
    rec, spec, path = load(name, tmp_path)
    m = fl.FlMesh.read_gmsh(path, spec["n_dim"])
    ctx = fl.FlCtx(m)
    return fl, rec, spec, m, ctx


def bitwise(a, b, what):
    a = np.asarray(a)
    b = np.asarray(b)
    assert a.shape == b.shape, (what, a.shape, b.shape)
    if not np.array_equal(a, b):
        bad = np.argwhere(a != b)
        k = tuple(bad[0])
        raise AssertionError("%s: %d of %d differ, first at %s: %r vs %r (rel %.2e)"
                             % (what, len(bad), a.size, k, a[k], b[k], relerr(a, b)))


def close(a, b, tol, what):
    e = relerr(a, b)
    say("  %-28s rel %.2e" % (what, e))
    assert e <= tol, "%s: relative error %.3e > %.1e" % (what, e, tol)


@pytest.mark.parametrize("name", CASES)
def test_geometry_bitwise(name, tmp_path):
    fl, rec, spec, m, ctx = setup(name, tmp_path)
    say("geometry %s: %d cells, %d faces" % (name, m.n_dmn, m.n_fce))
    g = ctx.geometry()
    bitwise(g["vol"], rec["geo_vol"], "vol")
    if "geo_x_vc" in rec:
        for k in ("x_vc", "x_fc", "norm", "area", "x_pc", "x_nc", "l_pn", "w_fp"):
            bitwise(g[k], rec["geo_" + k], k)
        bitwise(np.transpose(g["ls"], (0, 2, 1)), rec["geo_ls"], "ls")
    ctx.destroy()


@pytest.mark.parametrize("name", SMALL_CASES)
def test_ls_gradient_bitwise(name, tmp_path):
    fl, rec, spec, m, ctx = setup(name, tmp_path)
    say("ls_gradient %s" % name)
    # iteration 2's pressure gradient from iteration 1's pressure, boundary
    # values extrapolated (neu_bnd 1)
    g, pb = ctx.ls_gradient(rec["it1_pres"], rec["it1_pres_bnd"], 1)
    bitwise(g[:, 0, :], rec["it2_pres_grad"], "pressure gradient")
    bitwise(pb[:, 0], rec["it2_pres_bnd"], "extrapolated pressure boundary values")
    # iteration 2's velocity gradient (neu_bnd 0: the boundary values as set
    # by iteration 1's construct_pde)
    g, _ = ctx.ls_gradient(rec["it1_vel"], rec["it1_vel_bnd"], 0)
    bitwise(g, rec["it2_vel_grad"], "velocity gradient")
    # pressure-correction gradient of iteration 1
    g, pb = ctx.ls_gradient(rec["it1_pp"], np.zeros(m.n_bnd), 1)
    bitwise(g[:, 0, :], rec["it1_pp_grad"], "pressure-correction gradient")
    bitwise(pb[:, 0], rec["it1_pp_bnd"], "pressure-correction boundary values")
    ctx.destroy()


@pytest.mark.parametrize("name", SMALL_CASES)
def test_lin_eqns_solver(name, tmp_path):
    fl, rec, spec, m, ctx = setup(name, tmp_path)
    say("lin_eqns_solver %s" % name)
    nd = m.n_dim
    # iteration 2's momentum solves (BiCGStab, from iteration 1's velocity)
    for c in range(nd):
        phi, r, nrm = ctx.lin_eqns_solver(rec["it2_mom_val"], rec["it1_vel"][:, c], rec["it2_mom_rhs"][:, c], 25,
                                          0.01, 2)
        close(phi, rec["it2_vel_solved"][:, c], 1e-9, "bicg phi, component %d" % c)
        bitwise(r, rec["it2_mom_resid"][:, c], "bicg initial residual (csr_matvec_prod), component %d" % c)
        close(nrm, rec["it2_mom_res0"][c], 1e-9, "bicg norm, component %d" % c)
    # iteration 1's pressure-correction solve (CG from zero)
    phi, r, nrm = ctx.lin_eqns_solver(rec["it1_pp_val"], np.zeros(m.n_dmn), rec["it1_pp_rhs"], 25, 0.01, 1)
    close(phi, rec["it1_pp"], 1e-9, "cg phi")
    close(nrm, rec["it1_pp_res0"][0], 1e-9, "cg norm")
    ctx.destroy()


@pytest.mark.parametrize("name", SMALL_CASES)
def test_assembly_bitwise_on_reference_inputs(name, tmp_path):
    fl, rec, spec, m, ctx = setup(name, tmp_path)
    say("assembly %s" % name)
    fw = fl.FlFlow(ctx, spec["rho"], spec["mu"], spec["urf_u"], spec["urf_p"], spec["bl_c"])
    # iteration 1 from the initial state
    fw.set(vel=np.zeros((m.n_dmn, m.n_dim)), vel_bnd=rec["vel_bnd0"], pres=np.zeros(m.n_dmn),
           pres_bnd=np.zeros(m.n_bnd))
    for s in ("zero", "grad_pres", "grad_vel", "ns_source", "construct_pde"):
        fw.step(s)
    sy = fw.system()
    bitwise(sy["val"], rec["it1_mom_val"], "momentum matrix, iteration 1")
    bitwise(sy["rhs"], rec["it1_mom_rhs"], "momentum rhs, iteration 1")
    bitwise(sy["diag"], rec["it1_mom_diag"], "1/a_P, iteration 1")
    bitwise(fw.get()["vel_bnd"], rec["it1_vel_bnd"], "boundary velocities after construct_pde")
    # construct_pp on the reference's solved velocity
    fw.set(vel=rec["it1_vel_solved"])
    fw.step("zero")
    fw.step("construct_pp")
    sy = fw.system()
    bitwise(sy["val"], rec["it1_pp_val"], "pressure-correction matrix, iteration 1")
    bitwise(sy["rhs"][:, 0], rec["it1_pp_rhs"], "pressure-correction rhs, iteration 1")
    bitwise(fw.get()["spd"], rec["it1_spd"], "face fluxes, iteration 1")
    # iteration 2's momentum system from the reference's state after
    # iteration 1 (its face fluxes are the ones construct_pp just made)
    fw.set(vel=rec["it1_vel"], vel_bnd=rec["it1_vel_bnd"], pres=rec["it1_pres"], pres_bnd=rec["it1_pres_bnd"])
    for s in ("zero", "grad_pres", "grad_vel", "ns_source", "construct_pde"):
        fw.step(s)
    sy = fw.system()
    bitwise(sy["pres_grad"], rec["it2_pres_grad"], "pressure gradient, iteration 2")
    bitwise(sy["vel_grad"], rec["it2_vel_grad"], "velocity gradient, iteration 2")
    bitwise(sy["val"], rec["it2_mom_val"], "momentum matrix, iteration 2")
    bitwise(sy["rhs"], rec["it2_mom_rhs"], "momentum rhs, iteration 2")
    bitwise(sy["diag"], rec["it2_mom_diag"], "1/a_P, iteration 2")
    fw.destroy()
    ctx.destroy()


@pytest.mark.parametrize("name", CASES)
def test_outer_iterations_match_reference(name, tmp_path):
    fl, rec, spec, m, ctx = setup(name, tmp_path)
    say("outer iterations %s: %d cells" % (name, m.n_dmn))
    fw = fl.FlFlow(ctx, spec["rho"], spec["mu"], spec["urf_u"], spec["urf_p"], spec["bl_c"])
    fw.set(vel_bnd=rec["vel_bnd0"])
    res0, gl = fw.run(spec["n_iter"], spec["res_drop"])
    assert len(gl) == int(rec["n_done"]), (len(gl), int(rec["n_done"]))
    close(res0, rec["res0"], 1e-9, "residual norms, all iterations")
    close(gl, rec["norm_gl"], 1e-9, "norm_gl, all iterations")
    st = fw.get()
    close(st["vel"], rec["final_vel"], 1e-9, "velocity")
    close(st["pres"], rec["final_pres"], 1e-9, "pressure")
    close(st["pp"], rec["final_pp"], 1e-9, "pressure correction")
    close(st["spd"], rec["final_spd"], 1e-9, "face fluxes")
    close(st["vel_bnd"], rec["final_vel_bnd"], 1e-9, "boundary velocities")
    fw.destroy()
    ctx.destroy()


SWEEP_MODES = [
    ("one launch per level", {"sweep_one_launch": 0}),
    ("rows in level order, chunks", {"sweep_parts": 0}),
    ("ranges of 4096 rows", {}),
    ("ranges of 256 rows, 7 blocks", {"sweep_part_rows": 256, "sweep_blocks": 7}),
]


@pytest.mark.parametrize("mesh", ["box40", "fl_prism3"])
def test_one_launch_sweeps_bitwise(mesh, tmp_path):
    """The factor and direct_solve's sweeps in one launch each (rows taken in
    level order in chunks; or blocks owning ranges of rows, walked level by
    level with values inside a range through LDS, including blocks that take
    many ranges) against one launch per level: the same bits after three
    outer iterations.  A 40^3 jittered cavity (64,000 cells, 118 levels per
    sweep, three lower entries a row) and the prism mesh (five)."""
    require_lib()
    from programs_amd import flowlab as fl
    from programs_amd import meshgen
    if mesh == "box40":
        m = fl.FlMesh.from_records(3, *meshgen.box_records(40, jitter=0.1))
        bnd = ("Wall", 1, (1, 0, 0, 1))
    else:
        rec, spec, path = load(mesh, tmp_path)
        m = fl.FlMesh.read_gmsh(path, spec["n_dim"])
        bnd = None
    out = []
    for label, params in SWEEP_MODES:
        ctx = fl.FlCtx(m)
        for k, v in params.items():
            ctx.set_param(k, v)
        say("%s: %s" % (mesh, label))
        fw = fl.FlFlow(ctx, 1.0, 0.01, 0.7, 0.3, 0.5)
        if bnd:
            fw.set(vel_bnd=fl.boundary_values(m, ctx.geometry()["norm"], *bnd))
        else:
            fw.set(vel_bnd=rec["vel_bnd0"])
        res0, gl = fw.run(3, 0.0)
        out.append((label, res0, fw.get()))
        fw.destroy()
        ctx.destroy()
    _, r0, s0 = out[0]
    for label, r1, s1 in out[1:]:
        bitwise(r1, r0, label + ": residual norms")
        for k in ("vel", "pres", "pp", "spd"):
            bitwise(s1[k], s0[k], label + ": " + k)



def test_sweep_wait_failure_is_reported():
    """The one-launch sweeps' failure path, forced from the host ("sweep_fail"
    raises the wait-bound flag): the outer iteration's solves end and report
    FVK_ERR_HIP; a new context on the same mesh gives the first run's bits."""
    fvk = require_lib()
    from programs_amd import flowlab as fl
    from programs_amd import meshgen
    say("flowlab sweep failure path: box 24^3")
    m = fl.FlMesh.from_records(3, *meshgen.box_records(24, jitter=0.1))

    def run(fail):
        ctx = fl.FlCtx(m)
        fw = fl.FlFlow(ctx, 1.0, 0.01, 0.7, 0.3, 0.5)
        fw.set(vel_bnd=fl.boundary_values(m, ctx.geometry()["norm"], "Wall", 1, (1, 0, 0, 1)))
        try:
            if fail:
                ctx.set_param("sweep_fail", 1)
                with pytest.raises(fvk.FvkError) as e:
                    fw.run(1, 0.0)
                assert e.value.status == fvk.FVK_ERR_HIP and "exceeded its bound" in str(e.value)
                return None
            fw.run(1, 0.0)
            return fw.get()
        finally:
            fw.destroy()
            ctx.destroy()

    s0 = run(False)
    run(True)
    s1 = run(False)
    for k in ("vel", "pres", "pp", "spd"):
        bitwise(s1[k], s0[k], "after the failed context: " + k)
    say("  reported, and a new context runs as before")


# ---------------------------------------------------------------- Fortran drop-in
DROPIN = os.path.join(os.path.dirname(GOLDEN), "..", "oracle", "_ref", "dropin_flowlab", "flowlab_ref_driver")
DROPIN_STAMP = os.path.join(os.path.dirname(DROPIN), "sources.sha256")
DROPIN_SOURCES = ["programs_amd/fortran/fvk_fl_solv.f90", "oracle/ref/flowlab_ref_driver.f90"]


def dropin_sources_hash():
    import hashlib
    root = os.path.dirname(os.path.dirname(GOLDEN))
    lines = "".join("%s  %s\n" % (hashlib.sha256(open(os.path.join(root, s), "rb").read()).hexdigest(), s)
                    for s in DROPIN_SOURCES)
    return hashlib.sha256(lines.encode()).hexdigest()


@pytest.mark.parametrize("name", CASES)
def test_fortran_dropin_solv_m(name, tmp_path):
    """flowlab's own Fortran (geometry_m, transport_m and the golden driver's
    outer loop) linked with the drop-in solv_m (programs_amd/fortran/
    fvk_fl_solv.f90: lin_eqns_solver on the GPU through the C ABI): its run
    must give the reference's residual norms and fields (1e-9).  The
    executable is built by programs_amd/fortran/build_dropin.sh in the
    development container and travels with the tree; a missing or stale one
    fails."""
    from oracle import flowlab_ref
    assert os.path.exists(DROPIN), "drop-in executable missing: run programs_amd/fortran/build_dropin.sh"
    assert open(DROPIN_STAMP).read().strip() == dropin_sources_hash(), "drop-in executable is stale: rebuild it"
    fl, rec, spec, m, ctx = setup(name, tmp_path)
    ctx.destroy()
    say("drop-in %s: %d cells" % (name, m.n_dmn))
    inp, out = os.path.join(str(tmp_path), "a.in"), os.path.join(str(tmp_path), "a.out")
    par = [spec["rho"], spec["mu"], spec["urf_u"], spec["urf_p"], spec["bl_c"], spec["res_drop"]]
    flowlab_ref.write_input(inp, m, par, spec["n_iter"], 0, rec["vel_bnd0"])
    import subprocess
    subprocess.run([DROPIN, inp, out], check=True, preexec_fn=flowlab_ref._big_stack)
    o = flowlab_ref.read_output(out, m, spec["n_iter"], 0)
    assert int(o["n_done"]) == int(rec["n_done"])
    bitwise(o["geo_vol"], rec["geo_vol"], "vol (reference code)")
    close(o["res0"], rec["res0"], 1e-9, "residual norms, all iterations")
    for k in ("vel", "pres", "pp", "spd", "vel_bnd"):
        close(o["final_" + k], rec["final_" + k], 1e-9, k)
