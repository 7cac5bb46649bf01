"""Benchmark of the pressure solve on the 128^3 jittered cavity mesh (MI355X).

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--n 128] [--no-cpu]

Workload (one "step"): the stand-alone pressure-correction system on the
128^3 hex cavity mesh with every interior vertex jittered by +-20% of the
spacing (SplitMix64 seed 0x5EED0002), face coefficients
a_f = s_f/l_pn * lognormal(0, 1) (seed 0x5EED0003), zero-mean N(0,1)
right-hand side, solved from x = 0 to a relative L1 residual of 1e-8 with
the reference's Jacobi-PCG recurrence in FP64 (solv_cg.f90 with imat=0).
Inputs are resident in HBM when the timed region starts; the timed region
is K complete solves.

Printed: one JSON line with the metric, the solve time, the roofline of the
dominant kernel (the fused p-update + SpMV of the CG iteration, timed with
HIP events on its own stream), the plain SpMV bandwidth, and the CPU oracle's
time for the same solve (single core, extrapolated from a bounded sample of
iterations).
"""
import argparse
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

METRIC = "pressure-solve s to 1e-8 rel. residual (learned vs plain PCG); SpMV HBM GB/s"
PEAK_HBM = 8.0e12


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def cpu_baseline(ctx, fvk, n, jitter, gpu_iters, sample_iters):
    """The oracle's PCG (plain C, one core) on the same system: time a bounded
    number of iterations and scale to the GPU path's iteration count."""
    from oracle import fvo
    t0 = time.time()
    om = fvo.OMesh.box(n, n, n, jitter=jitter, seed=0x5EED0002)
    pl, a, b = fvk.sys_get(ctx)
    A = fvo.OSys(om, pl, -a, -a)
    t_setup = time.time() - t0
    t0 = time.time()
    A.solve("cg", 0, 0.0, sample_iters, b, not_r0=True)
    dt = time.time() - t0
    per_it = dt / sample_iters
    om.free()
    return dict(value=per_it * gpu_iters, unit="s", cores=1, kind="port",
                sample="%d Jacobi-PCG iterations of the same %d^3 system on one host core (%.1f s, "
                       "oracle set-up %.1f s excluded), scaled to the %d iterations the GPU path needs"
                       % (sample_iters, n, dt, t_setup, gpu_iters),
                per_iter_ms=per_it * 1e3)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=128)
    ap.add_argument("--jitter", type=float, default=0.2)
    ap.add_argument("--sigma", type=float, default=1.0)
    ap.add_argument("--tol", type=float, default=1e-8)
    ap.add_argument("--cpu-sample", type=int, default=300)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--methods", default="dc32", help="extra methods timed beside the plain solve")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (the sharding design is "
              "described in DESIGN.md); run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    from programs_amd import fvk
    n = args.n
    t0 = time.time()
    mesh = fvk.Mesh.box(n, n, n, jitter=args.jitter, seed=0x5EED0002)
    log("mesh %d^3: %d cells, %d faces (%.1f s)" % (n, mesh.n_vol, mesh.n_fce, time.time() - t0))
    ctx = fvk.grd_solv(mesh)
    fvk.sys_random(ctx, 0x5EED0003, args.sigma)
    log("context + system ready (%.1f s)" % (time.time() - t0))

    def timed(method, name):
        for _ in range(args.warmup):
            r = fvk.sys_solve(ctx, method, 0, args.tol, 200000)
            log("%s warmup solve: %d iterations, %.3f s device" % (name, r["iters"], r["seconds"]))
        ctx.sync()
        t0 = time.perf_counter()
        res = []
        for _ in range(args.steps):
            res.append(fvk.sys_solve(ctx, method, 0, args.tol, 200000))
        ctx.sync()
        wall = (time.perf_counter() - t0) / args.steps
        r = res[-1]
        log("%s timed: %.4f s per solve (wall), %.4f s device, %d iterations, true L1 %.2e L2 %.2e"
            % (name, wall, r["seconds"], r["iters"], r["res_true_l1"], r["res_true_l2"]))
        return wall, r

    wall, r = timed(fvk.SOLVE_PCG64, "plain_pcg64_jacobi")
    methods = {"plain_pcg64_jacobi": dict(s=wall, iters=r["iters"], res_true_l1=r["res_true_l1"],
                                          res_true_l2=r["res_true_l2"])}
    if "dc32" in args.methods.split(","):
        w2, r2 = timed(fvk.SOLVE_DC32, "dc32_jacobi")
        rk2 = fvk.sys_solve(ctx, fvk.SOLVE_DC32, 0, args.tol, 200000, time_kernel=True)
        methods["dc32_jacobi"] = dict(s=w2, iters=r2["iters"], res_true_l1=r2["res_true_l1"],
                                      res_true_l2=r2["res_true_l2"], speedup_vs_plain=wall / w2,
                                      k_mx_pq_us=rk2["kern_seconds"] * 1e6,
                                      k_mx_pq_gbs=rk2["kern_bytes"] / rk2["kern_seconds"] / 1e9)

    # dominant kernel, timed live per launch with HIP events on the solve stream
    rk = fvk.sys_solve(ctx, fvk.SOLVE_PCG64, 0, args.tol, 200000, time_kernel=True)
    kb, ks = rk["kern_bytes"], rk["kern_seconds"]
    roof = dict(bound="hbm", kernel="k_cg_pq (p = z + beta p; q = A p; p.q)", achieved=kb / ks / 1e9,
                peak=PEAK_HBM / 1e9, unit="GB/s", frac=kb / ks / PEAK_HBM, traffic=None,
                bytes_per_launch=kb, us_per_launch=ks * 1e6, launches=rk["kern_launches"])

    # stand-alone FP64 SpMV, BASELINE's CSR byte model nnz(w+4) + 4(n+1) + 2wn
    t_spmv = fvk.bench_spmv(ctx, 200)
    nnz = mesh.n_vol + 2 * (mesh.n_fce - mesh.n_bf)
    b_spmv = nnz * 12 + 4 * (mesh.n_vol + 1) + 16 * mesh.n_vol

    copy_bytes = 1 << 30
    t_copy = fvk.bench_copy(copy_bytes, 20)

    out = {
        "metric": METRIC,
        "value": wall,
        "unit": "s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall * 1e3,
        "higher_is_better": False,
        "scaling": "strong",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": "pressure-correction system on the %d^3 cavity mesh jittered +-%g%% (%d cells, %d "
                        "faces), a_f = s/l_pn*lognormal(0,%g), zero-mean N(0,1) RHS; plain FP64 Jacobi-PCG "
                        "(reference recurrence) from x=0 to relative L1 residual %g"
                        % (n, 100 * args.jitter, mesh.n_vol, mesh.n_fce, args.sigma, args.tol),
            "method": "plain_pcg64_jacobi",
            "cells": mesh.n_vol,
            "nnz": nnz,
        },
        "iters": r["iters"],
        "device_s_per_solve": r["seconds"],
        "res_true_l1": r["res_true_l1"],
        "res_true_l2": r["res_true_l2"],
        "roofline": roof,
        "spmv": dict(us_per_launch=t_spmv * 1e6, bytes=b_spmv, gbs=b_spmv / t_spmv / 1e9,
                     frac=b_spmv / t_spmv / PEAK_HBM),
        "copy_gbs": 2 * copy_bytes / t_copy / 1e9,
        "methods": methods,
    }
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(ctx, fvk, n, args.jitter, r["iters"], args.cpu_sample)
    print(json.dumps(out), flush=True)
    ctx.destroy()


if __name__ == "__main__":
    main()
