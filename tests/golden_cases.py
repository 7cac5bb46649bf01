"""Loading of the golden fixtures (reference outputs) and matching oracle runs."""
import ast
import os

import numpy as np

from oracle import fvo

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
# the reference's SIMPLE runs: <mesh>_<shipped|relaxed>.npz (the other
# fixtures there are the output-format and gmsh vectors)
CASES = sorted(f[:-4] for f in os.listdir(GOLDEN)
               if f.endswith("_shipped.npz") or f.endswith("_relaxed.npz"))


def load(name):
    z = np.load(os.path.join(GOLDEN, name + ".npz"), allow_pickle=False)
    rec = {k: z[k] for k in z.files}
    rec["spec"] = ast.literal_eval(str(rec["case_json"]))
    return rec


def mesh_inputs(rec):
    """Raw mesh description (n_dx, vx, v_ty, v_vptr, v_vrt, bnd) or a .neu path."""
    if "mesh_neu" in rec:
        return os.path.join(GOLDEN, str(rec["mesh_neu"]))
    return dict(n_dx=int(rec["mesh_n_dx"]), vx=rec["mesh_vx"], v_ty=rec["mesh_v_ty"],
                v_vptr=rec["mesh_v_vptr"], v_vrt=rec["mesh_v_vrt"], bnd=rec["mesh_bnd"])


def oracle_mesh(rec, conn="hash"):
    src = mesh_inputs(rec)
    if isinstance(src, str):
        return fvo.OMesh.read_neu(src, conn=conn)
    vx = np.zeros((len(src["vx"]), 3))
    vx[:, : src["n_dx"]] = src["vx"][:, : src["n_dx"]]
    return fvo.OMesh.raw(src["n_dx"], vx, src["v_ty"], src["v_vptr"], src["v_vrt"], src["bnd"], conn=conn)


def cfg_for(rec):
    spec = rec["spec"]
    if spec["variant"] == "shipped":
        return fvo.default_cfg(rho=spec["rho"], mu=spec["mu"], urf_p=spec["urf_p"])
    return fvo.default_cfg(simpler=0, imat_p=0, rho=spec["rho"], mu=spec["mu"], urf_p=spec["urf_p"])
