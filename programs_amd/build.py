"""Build libfvk.so (HIP for gfx950 + host C++) in place, next to this file.

    python -m programs_amd.build        # or: python programs_amd/build.py

Every translation unit is compiled without floating-point contraction
(-ffp-contract=off): the parity path reproduces the reference's separately
rounded multiply and add.  Packed FP32 instructions are disabled for the
device code (-Xclang -target-feature -Xclang -packed-fp32-ops).
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "libfvk.so")
BUILD = os.path.join(HERE, "_build")

HIP_SOURCES = ["krylov.hip", "mixed.hip", "learn.hip", "flow.hip", "meshdev.hip", "flowlab.hip", "capi.hip"]
CPP_SOURCES = ["mesh.cpp", "output.cpp", "flmesh.cpp"]
ARCH = os.environ.get("FVK_OFFLOAD_ARCH", "gfx950")

COMMON = ["-O3", "-fPIC", "-std=c++17", "-ffp-contract=off", "-Wall", "-Wno-unused-function",
          "-Wno-unused-variable", "-Wno-unused-but-set-variable"]
HIPFLAGS = ["--offload-arch=" + ARCH, "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
            "-munsafe-fp-atomics"]


def run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(" ".join(cmd) + "\n" + r.stdout + r.stderr)
        raise RuntimeError("compile failed: " + cmd[-1])
    return r


def build(verbose=False):
    """Compile the translation units in parallel (one compiler process each,
    at most as many as there are cores, capped at 8), then link."""
    from concurrent.futures import ThreadPoolExecutor
    os.makedirs(BUILD, exist_ok=True)
    jobs = [(["hipcc", "-c"] + COMMON + HIPFLAGS, s) for s in HIP_SOURCES] + \
           [(["g++", "-c"] + COMMON, s) for s in CPP_SOURCES]
    objs = [os.path.join(BUILD, s + ".o") for _, s in jobs]
    workers = max(1, min(len(jobs), os.cpu_count() or 1, 8))
    with ThreadPoolExecutor(max_workers=workers) as pool:
        for f in [pool.submit(run, cmd + [os.path.join(CSRC, s), "-o", o]) for (cmd, s), o in zip(jobs, objs)]:
            f.result()
    tmp = OUT + ".tmp"
    run(["hipcc", "-shared", "--offload-arch=" + ARCH, "-o", tmp] + objs)
    os.replace(tmp, OUT)
    if verbose:
        print("built", OUT)
    return OUT


if __name__ == "__main__":
    build(verbose=True)
