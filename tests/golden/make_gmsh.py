"""Golden outputs of flowlab's gmsh reader and field writers
(flowlab/src/mesh.cpp: read_gmsh_mesh, write_gmsh_field_,
write_gmsh_vector_field_), through the g++ build oracle/_ref/gmsh/gmsh_ref.

    python tests/golden/make_gmsh.py   # development container only

Writes tests/golden/gmsh_fields.npz (inputs and the written files' bytes)
and tests/golden/gmsh_read_<name>.npz (the arrays read_gmsh_mesh fills) for
the reference's own duct_3d.msh and for the .msh files tests/golden/msh_*.msh
(written here by this project from its box and channel generators).  Data
only."""
import os
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

DRIVER = os.path.join(ROOT, "oracle", "_ref", "gmsh", "gmsh_ref")
DUCT = "/root/reference/flowlab/examples/duct_3d/duct_3d.msh"


def run(*args):
    subprocess.run([DRIVER] + list(args), check=True, capture_output=True)


def fields():
    rng = np.random.default_rng(0x5EED00C1)
    out = {}
    with tempfile.TemporaryDirectory() as d:
        for name, n, nd in (("scalar", 40, 1), ("vec2", 30, 2), ("vec3", 25, 3)):
            idx = rng.permutation(1000)[:n].astype(np.int32) + 1
            phi = rng.standard_normal(n * nd) * 10.0 ** rng.integers(-12, 12, n * nd)
            phi[:3] = [0.0, -0.0, 1.0]
            fin = os.path.join(d, "in.bin")
            with open(fin, "wb") as f:
                np.array([n], np.int32).tofile(f)
                if nd > 1:
                    np.array([nd], np.int32).tofile(f)
                idx.tofile(f)
                phi.tofile(f)
            fout = os.path.join(d, name + ".msh")
            run("field" if nd == 1 else "vector", fin, fout)
            out[name + "_idx"] = idx
            out[name + "_phi"] = phi
            out[name + "_n_dim"] = np.array(nd)
            # the file name is written into the file: keep it relative
            out[name + "_bytes"] = np.frombuffer(open(fout, "rb").read(), np.uint8)
            out[name + "_fname"] = np.array(fout)
    np.savez_compressed(os.path.join(HERE, "gmsh_fields.npz"), **out)


def read_case(name, path):
    with tempfile.TemporaryDirectory() as d:
        fout = os.path.join(d, "dump.bin")
        run("read", path, fout)
        b = open(fout, "rb").read()
    o = 0

    def ints(n):
        nonlocal o
        a = np.frombuffer(b, np.int32, n, o).copy()
        o += 4 * n
        return a
    n_vrt = int(ints(1)[0])
    vrt_idx = ints(n_vrt)
    n_elm = int(ints(1)[0])
    elm_idx = ints(n_elm)
    tag_ptr = ints(n_elm + 1)
    tag_lst = ints(int(tag_ptr[-1]))
    vrt_ptr = ints(n_elm + 1)
    vrt_lst = ints(int(vrt_ptr[-1]))
    x = np.frombuffer(b, np.float64, 3 * n_vrt, o).copy()
    np.savez_compressed(os.path.join(HERE, "gmsh_read_%s.npz" % name), vrt_idx=vrt_idx, elm_idx=elm_idx,
                        tag_ptr=tag_ptr, tag_lst=tag_lst, vrt_ptr=vrt_ptr, vrt_lst=vrt_lst, x=x.reshape(-1, 3))


def main():
    write_meshes()
    fields()
    read_case("duct3d", DUCT)
    for f in sorted(os.listdir(HERE)):
        if f.startswith("msh_") and f.endswith(".msh"):
            read_case(f[4:-4], os.path.join(HERE, f))
    print("wrote gmsh fixtures")



def write_meshes():
    """The msh_*.msh inputs (this project's own meshes)."""
    sys.path.insert(0, HERE)
    import msh_gen
    from programs_amd import fvk
    _, vx, el = msh_gen.box_elements(fvk, 4, 0.2, 0x5EED00C2)
    msh_gen.write(os.path.join(HERE, "msh_box4j.msh"), vx, el)
    _, vx, el = msh_gen.box_elements(fvk, 3, 0.1, 0x5EED00C3, prisms=True)
    msh_gen.write(os.path.join(HERE, "msh_prism3.msh"), vx, el)
    vx, el = msh_gen.channel_elements(400, 0x5EED00C4)
    msh_gen.write(os.path.join(HERE, "msh_chan400.msh"), vx, el)


if __name__ == "__main__":
    main()
