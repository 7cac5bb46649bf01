"""The reference's own PCG (solv_cg.f90, compiled with flang by
oracle/build_ref.sh into oracle/_ref/cg/fvk_ref_cg) on a stand-alone system.

TEST INFRASTRUCTURE: used only by tests/ (which pin the driver against the
oracle's restatement) and by bench.py's cpu_baseline leg (which times it).
The system goes to the driver as a stream file: int32 n_vol, n_fce; int32
ip, in (1-based, in = 0 on boundary faces); float64 p = A(ip, in),
n = A(in, ip), pl (diagonal), b.
"""
import os
import subprocess
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
DRIVER = os.path.join(HERE, "_ref", "cg", "fvk_ref_cg")


def available():
    return os.path.isfile(DRIVER) and os.access(DRIVER, os.X_OK)


def write_system(path, fce_ip, fce_in, ngp, ngn, pl, b):
    """fce_ip, fce_in: 1-based cells per face (fce_in = 0 on boundary faces)."""
    with open(path, "wb") as f:
        np.array([len(pl), len(fce_ip)], np.int32).tofile(f)
        np.asarray(fce_ip, np.int32).tofile(f)
        np.asarray(fce_in, np.int32).tofile(f)
        np.asarray(ngp, np.float64).tofile(f)
        np.asarray(ngn, np.float64).tofile(f)
        np.asarray(pl, np.float64).tofile(f)
        np.asarray(b, np.float64).tofile(f)


def run(fce_ip, fce_in, ngp, ngn, pl, b, n_iter, want_x=False, timeout=600, reduc=None):
    """Runs n_iter iterations of the reference's Jacobi PCG from x = 0 (no
    early exit; with reduc, at most n_iter iterations and the reference's
    relative L1 exit test).  Returns (seconds inside solv_cg, x or None)."""
    with tempfile.TemporaryDirectory() as d:
        sys_path = os.path.join(d, "system.bin")
        x_path = os.path.join(d, "x.bin")
        write_system(sys_path, fce_ip, fce_in, ngp, ngn, pl, b)
        # solv_cg keeps its work vectors in automatic arrays (5 x n doubles on
        # the stack): the driver runs under the hard stack limit, from a shell
        # that starts it as its own child
        cmd = ["bash", "-c", 'ulimit -s hard 2>/dev/null; "$0" "$@"', DRIVER, sys_path, str(int(n_iter))]
        if want_x or reduc is not None:
            cmd += [x_path]
        if reduc is not None:
            cmd += [repr(float(reduc))]
        r = subprocess.run(cmd, capture_output=True, text=True, timeout=timeout)
        if r.returncode != 0:
            raise RuntimeError("fvk_ref_cg failed (exit status %d): %s%s" % (r.returncode, r.stdout, r.stderr))
        words = r.stdout.split()
        seconds = float(words[words.index("seconds") + 1])
        x = np.fromfile(x_path, np.float64) if (want_x or reduc is not None) else None
        return seconds, x
