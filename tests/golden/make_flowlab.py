"""Golden vectors of flowlab's serial CSR SIMPLE, from flowlab's own Fortran
modules built with flang in this container (oracle/build_ref.sh ->
oracle/_ref/flowlab/flowlab_ref_driver, driver oracle/ref/flowlab_ref_driver.f90).

For each case the mesh is read from a gmsh file into the arrays flowlab's
Fortran receives (programs_amd.flowlab.FlMesh: boundary faces first, interior
faces in reverse order of discovery), the velocity boundary values are set as
navierstokes' set_boundary_conditions sets them, and the driver runs the
outer loop of navierstokes.f90:515-641.  Stored: the geometry and the least-
squares matrices, the intermediate arrays of the first outer iterations
(gradients, assembled matrices and right-hand sides, solver outputs,
corrected fields), the residual norms of every iteration and the final
fields.  Run from the repository root:  python -m tests.golden.make_flowlab
"""
import gzip
import os
import shutil
import sys
import tempfile

import numpy as np

from oracle.flowlab_ref import available, read_output, run, write_input

HERE = os.path.dirname(os.path.abspath(__file__))

# name: mesh file, n_dim, density, viscosity, urf_u, urf_p, blending, boundary (type, index, value),
#       outer iterations, residual drop, iterations dumped
CASES = {
    # flowlab/examples/duct_3d/duct_3d.json on flowlab's own duct_3d.msh
    "fl_duct3d": ("duct_3d.msh.gz", 3, 1.2, 0.3, 0.55, 0.15, 0.5, ("Inlet", 1, (0, 0, 0, 1)), 10, 0.1, 1),
    # lid-driven cavity on a jittered 4x4x4 hex box (lid: physical tag 31)
    "fl_box4j": ("msh_box4j.msh", 3, 1.2, 0.03, 0.55, 0.15, 0.5, ("Wall", 1, (1, 0, 0, 1)), 30, 1e-4, 3),
    # prism mesh, the same lid
    "fl_prism3": ("msh_prism3.msh", 3, 1.0, 0.01, 0.7, 0.3, 0.5, ("Wall", 1, (1, 0, 0, 1)), 20, 1e-4, 2),
    # 2D Delaunay channel (inlet 11, outlet 21, walls 32), inflow along -normal
    "fl_chan400": ("msh_chan400.msh", 2, 1.0, 0.02, 0.55, 0.15, 0.5, ("Inlet", 1, (0, 0, 0, 1)), 30, 1e-4, 2),
}


SLIM = 5000  # cells above which a fixture keeps a subset of the arrays


def mesh_path(fname, tmp):
    p = os.path.join(HERE, fname)
    if fname.endswith(".gz"):
        q = os.path.join(tmp, fname[:-3])
        with gzip.open(p) as a, open(q, "wb") as b:
            shutil.copyfileobj(a, b)
        return q
    return p


def run_case(name, spec, tmp):
    from programs_amd import flowlab as fl
    fname, nd, rho, mu, urf_u, urf_p, blc, bnd, n_iter, drop, n_dump = spec
    m = fl.FlMesh.read_gmsh(mesh_path(fname, tmp), nd)
    par = [rho, mu, urf_u, urf_p, blc, drop]
    inp, out = os.path.join(tmp, name + ".in"), os.path.join(tmp, name + ".out")
    # geometry first (no iterations): the boundary values need the normals
    write_input(inp, m, par, 0, 0, np.zeros((m.n_bnd, nd)))
    run(inp, out)
    g = read_output(out, m, 0, 0)
    ub = fl.boundary_values(m, g["geo_norm"], *bnd)
    write_input(inp, m, par, n_iter, n_dump, ub)
    run(inp, out)
    o = read_output(out, m, n_iter, n_dump)
    o["vel_bnd0"] = ub
    if m.n_dmn > SLIM:
        # large meshes keep the geometry's volumes, the first iteration's
        # solver outputs and corrected fields, and the final fields
        keep = ("geo_vol", "it1_vel_solved", "it1_mom_res0", "it1_pp", "it1_pp_res0", "it1_vel", "it1_pres",
                "it1_spd")
        o = {k: v for k, v in o.items() if not (k.startswith("geo_") or k.startswith("it")) or k in keep}
    o["case"] = np.array(repr(dict(mesh=fname, n_dim=nd, rho=rho, mu=mu, urf_u=urf_u, urf_p=urf_p, bl_c=blc,
                                   bnd=bnd, n_iter=n_iter, res_drop=drop, n_dump=n_dump)))
    np.savez_compressed(os.path.join(HERE, name + ".npz"), **o)
    print(name, "cells", m.n_dmn, "iterations", int(o["n_done"]), "norm_gl", o["norm_gl"][-1])


def main():
    if not available():
        sys.exit("build the reference first: bash oracle/build_ref.sh")
    names = sys.argv[1:] or list(CASES)
    with tempfile.TemporaryDirectory() as tmp:
        for name in names:
            run_case(name, CASES[name], tmp)


if __name__ == "__main__":
    main()
