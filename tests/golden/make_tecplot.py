"""Golden Tecplot files written by the reference's own writers
(spray-simulation/cfd_outp.f90: eqn_outp for the velocity, prp_outp with
p=.true. for the pressure), through the flang build
oracle/_ref/outp/fvk_ref_outp after a few SIMPLE iterations of golden cases.

    python tests/golden/make_tecplot.py   # development container only

Writes tests/golden/tecplot_<case>_eqn.dat and tecplot_<case>_prp.dat: the
reference's output bytes, data only.  The element table tecplot.vrt beside
them is the reference's own fixture (spray-simulation/tecplot.vrt)."""
import os
import shutil
import subprocess
import sys
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
from tests.golden.make_golden import CASES, box_raw, write_neu  # noqa: E402

DRIVER = os.path.join(ROOT, "oracle", "_ref", "outp", "fvk_ref_outp")
# case: n_outer of the output run
RUNS = {"cav5_shipped": 4, "chan2d_shipped": 3, "box8j_shipped": 3}


def main():
    for name, n_outer in RUNS.items():
        spec = CASES[name]
        with tempfile.TemporaryDirectory() as d:
            kind, arg = spec["mesh"]
            if kind == "neu":
                neu = os.path.join(HERE, arg)
            else:
                neu = os.path.join(d, "mesh.neu")
                write_neu(neu, box_raw(*arg))
            out = d + "/"
            args = [DRIVER, os.path.join(HERE, "gambit.vrt"), os.path.join(HERE, "tecplot.vrt"), neu, out,
                    str(n_outer), repr(spec["rho"]), repr(spec["mu"]), repr(spec["urf_p"])]
            for code, u in spec["bc"]:
                args += [str(code)] + [repr(float(v)) for v in u]
            subprocess.run(["bash", "-c", 'ulimit -s hard 2>/dev/null; "$0" "$@"'] + args, check=True,
                           capture_output=True, cwd=d)
            for kind_, fn in (("eqn", "grd1%phs1%eqn1.dat"), ("prp", "grd1%phs1%prp1.dat")):
                shutil.copy(os.path.join(d, fn), os.path.join(HERE, "tecplot_%s_%s.dat" % (name, kind_)))
        print("wrote", name)


if __name__ == "__main__":
    main()
