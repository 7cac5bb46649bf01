"""Benchmark of the pressure solve on the 128^3 jittered cavity mesh (MI355X).

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--n 128] [--no-cpu]
                    [--methods dc32,learned]

Workload (one "step"): the stand-alone pressure-correction system on the
128^3 hex cavity mesh with every interior vertex jittered by +-20% of the
spacing (SplitMix64 seed 0x5EED0002), face coefficients
a_f = s_f/l_pn * lognormal(0, 1) (seed 0x5EED0003), zero-mean N(0,1)
right-hand side, solved from x = 0 to an FP64 relative L1 residual of 1e-8.
Inputs are resident in HBM when the timed region starts; the timed region is
K complete solves of each method:

  plain_pcg64_jacobi  the reference's Jacobi-PCG recurrence in FP64
                      (solv_cg.f90 with imat = 0), the baseline
  dc32_jacobi         FP32-inner Jacobi PCG in FP64 defect correction
  dc32_learned        the headline: factor G from the per-face network
                      (inference inside the timed step) and FP32-inner PCG
                      preconditioned by z = G G^T r, in FP64 defect correction

The network is trained beforehand (not timed) on the GPU on 2,000 pressure
systems that differ from the timed one (other mesh and coefficient seeds,
and the pressure-correction systems of a SIMPLE run).

Printed: one JSON line with the metric, the headline solve time and its
speedup over the plain solve, the roofline of the headline's dominant kernel
(the FP32 SpMV q = Ah p with its p.q partials, timed live with HIP events on
the solve's stream; its HBM traffic from the byte counters recorded in
profiles/hbm_traffic.json) and of the plain path's fused p-update + SpMV,
the FP64 SpMV bandwidth, a device copy's bandwidth, SIMPLE outer iterations
on the 128^3 cavity (reference pressure solves, eager and from HIP graphs,
and pressure solves to 1e-8 plain and learned), and CPU baselines on one
host core extrapolated from bounded samples: the reference's own solv_cg
(flang build of solv_cg.f90 with a timing driver, oracle/_ref/cg) on the same
system, the oracle's C restatement of it, and the oracle's SIMPLE outer
iteration.
"""
import argparse
import json
import os
import subprocess
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

METRIC = "pressure-solve s to 1e-8 rel. residual (learned vs plain PCG); SpMV HBM GB/s"
PEAK_HBM = 8.0e12
# HBM bytes per launch of the roofline kernels, from rocprofv3 FETCH_SIZE and
# WRITE_SIZE runs of this script (the commands and the copy calibration are
# in the file)
TRAFFIC = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles", "hbm_traffic.json")


def traffic_bytes(kernel):
    """HBM bytes per launch of `kernel` measured by the byte counters, each
    scaled by what the same counter reads for a device copy of known size;
    None without a measurement."""
    try:
        with open(TRAFFIC) as f:
            t = json.load(f)
        k = t["kernels"][kernel]
        c = t["calibration"]
        return 1024.0 * (k["fetch_kib"] / c["fetch_factor"] + k["write_kib"] / c["write_factor"])
    except (OSError, KeyError, ValueError):
        return None


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def cpu_baseline_reference(ctx, mesh, fvk, n, gpu_iters, sample_iters):
    """The reference's own PCG (spray-simulation/solv_cg.f90 built with flang
    -O2, one core, imat = 0) on the same system, through its timing driver:
    a bounded number of iterations timed inside solv_cg, scaled to the GPU
    path's iteration count (the recurrence is the same, so is the count).
    None where the reference build is absent."""
    from oracle import refcg
    if not refcg.available():
        return None
    pl, a, b = fvk.sys_get(ctx)
    t = mesh.topology()
    try:
        sec, _ = refcg.run(t["fce_ip"], t["fce_in"], -a, -a, pl, b, sample_iters)
    except (RuntimeError, OSError, subprocess.TimeoutExpired) as e:
        log("reference PCG baseline not measured: %s" % e)
        return None
    per_it = sec / sample_iters
    return dict(value=per_it * gpu_iters, unit="s", cores=1, kind="reference",
                sample="%d Jacobi-PCG iterations of the reference's solv_cg (flang -O2, one host core) on the "
                       "same %d^3 system, %.1f s inside solv_cg (writing and reading the system excluded), scaled "
                       "to the %d iterations the GPU path needs" % (sample_iters, n, sec, gpu_iters),
                per_iter_ms=per_it * 1e3)


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


def host_info():
    """Host cores this process may use, and the CPU model."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        usable = len(os.sched_getaffinity(0))
    except AttributeError:
        usable = os.cpu_count()
    return dict(cpu_model=model, nproc=os.cpu_count(), usable_cores=usable)


def cpu_baseline_all_cores(ctx, fvk, n, jitter, gpu_iters, sample_iters):
    """The same Jacobi-PCG recurrence on every usable host core (the
    oracle's OpenMP port: row-gather product, per-thread reductions; at most
    16 threads, the host share of one GPU), a bounded number of iterations
    scaled to the GPU path's count."""
    from oracle import fvo
    info = host_info()
    threads = max(1, min(16, info["usable_cores"]))
    om = fvo.OMesh.box(n, n, n, jitter=jitter, seed=0x5EED0002)
    pl, a, b = fvk.sys_get(ctx)
    A = fvo.OSys(om, pl, -a, -a)
    sec, _ = A.pcg_omp(b, sample_iters, threads)
    om.free()
    per_it = sec / sample_iters
    return dict(value=per_it * gpu_iters, unit="s", cores=threads, kind="port",
                sample="%d Jacobi-PCG iterations of the same %d^3 system on %d threads (OpenMP port of the "
                       "reference recurrence, %.2f s), scaled to the %d iterations the GPU path needs"
                       % (sample_iters, n, threads, sec, gpu_iters),
                per_iter_ms=per_it * 1e3, **info)


def cpu_simple(args, gpu_states):
    """The oracle's SIMPLE (plain C, one core, the reference's algorithm) on
    the C2 cavity: pressure-correction solves to the FP64 relative L1
    residual tol from x = 0 (p_solver 3: the reference's solv_cg with reduc =
    tol), outer iterations 2 and 3 timed (iteration 1 solves nothing: the
    first p' right-hand side is zero).  After iteration 3 each field of the
    GPU flows is compared with the oracle's (max relative error)."""
    from oracle import fvo
    n = args.simple_n
    t0 = time.time()
    om = fvo.OMesh.box(n, n, n, jitter=0.02, seed=0x5EED0002)
    of = fvo.OFlow(om, fvo.default_cfg(simpler=0, imat_u=0, imat_p=0, urf_p=0.3, p_solver=3, p_tol=args.tol))
    of.set_bnd_u(31, [1.0, 0.0, 0.0])
    of.outer(1)
    t_setup = time.time() - t0
    t0 = time.time()
    of.outer(2)
    of.outer(3)
    dt = (time.time() - t0) / 2
    F = of.fields()
    errs = {}
    for name, st in gpu_states.items():
        errs[name] = {k: float(np.max(np.abs(st[k] - F[k])) / np.max(np.abs(F[k])))
                      for k in ("u_cell", "p", "dpdx", "spd")}
    of.free()
    om.free()
    log("oracle SIMPLE %d^3: %.1f s per outer iteration; GPU fields after 3 iterations vs oracle: %s"
        % (n, dt, json.dumps(errs)))
    return dict(value=dt, unit="s per outer iteration", cores=1, kind="port",
                sample="outer iterations 2-3 of the oracle's SIMPLE on the %d^3 cavity with pressure solves "
                       "to %g (set-up and iteration 1, %.1f s, excluded)" % (n, args.tol, t_setup),
                field_rel_err_vs_gpu_after_3=errs)


def train_model(fvk, args):
    """The network of the learned factor, trained on the GPU (untimed) by one
    Adam step on each of args.train_simple pressure-correction systems that a
    SIMPLE cavity run generates (32^3, +-2% jitter, one per outer iteration)
    and, interleaved with them, args.train_lognormal lognormal pressure
    systems (32^3, other mesh and coefficient seeds than the timed system).
    Interleaving was chosen over either set first by the iterations on a
    held-out 64^3 lognormal system (315 against 318 with the SIMPLE systems
    last), and the Adam step 1e-2 over 3e-3 by four held-out 128^3 lognormal
    systems of another mesh seed (2193 against 2228 iterations in all, fewer
    on each)."""
    t0 = time.time()
    model = fvk.Model(seed=0x5EED0004)
    m = fvk.Mesh.box(32, 32, 32, jitter=args.jitter, seed=0x5EED0012)
    c = fvk.grd_solv(m)
    n_simple = max(0, args.train_simple)
    per = args.train_lognormal // max(1, n_simple) if n_simple else args.train_lognormal
    done = [0]
    losses = []

    def lognormal(count):
        for _ in range(count):
            if done[0] >= args.train_lognormal:
                return
            fvk.sys_random(c, 0x7000 + done[0], args.sigma)
            losses.append(model.train(c, 1, lr=args.lr)[0])
            done[0] += 1

    if n_simple > 0:
        ms = fvk.Mesh.box(32, 32, 32, jitter=0.02, seed=0x5EED0013)
        cs = fvk.grd_solv(ms)
        cfg = fvk.flow_cfg(simpler=0, imat_u=0, imat_p=0, urf_p=0.3, p_solver=2, p_tol=args.tol)
        fl = fvk.Flow(cs, cfg)
        fl.set_bnd_u(31, [1.0, 0.0, 0.0])
        fl.set_model(model, train_steps=1, lr=args.lr)
        for it in range(1, n_simple + 1):
            fl.run(it, 1)
            lognormal(per)
            if it % 100 == 0:
                log("  trained on %d SIMPLE and %d lognormal systems, loss %.5f (%.1f s)"
                    % (it, done[0], losses[-1] if losses else 0.0, time.time() - t0))
        fl.destroy()
        cs.destroy()
        ms.free()
    lognormal(args.train_lognormal - done[0])
    c.destroy()
    m.free()
    ctx_s = time.time() - t0
    first = losses[0] if losses else 0.0
    last = losses[-1] if losses else 0.0
    log("trained on %d lognormal + %d SIMPLE pressure systems in %.1f s (loss %.4f -> %.4f on the "
        "lognormal systems)" % (done[0], n_simple, ctx_s, first, last))
    return model, dict(systems=done[0] + n_simple, lognormal=done[0], simple=n_simple, seconds=ctx_s,
                       loss_first=first, loss_last=last, schedule="interleaved")


def simple_leg(fvk, args, model):
    """SIMPLE outer iterations on the 128^3 lid-driven cavity (Re=100, lid u=1,
    +-2% jitter: the reference's cell volumes stay positive there), momentum
    BiCGStab with u, v, w batched (5 iterations, Jacobi), relaxed pressure
    update (urf 0.3):
      * reference pressure solves (PCG recurrence, 20 iterations, Jacobi),
        eager and replayed from HIP graphs: outer iterations per second;
      * pressure-correction solves to an FP64 relative residual of tol, plain
        FP64 PCG against the learned FP32-inner solve: seconds per outer
        iteration."""
    n = args.simple_n
    out = {}
    m = fvk.Mesh.box(n, n, n, jitter=0.02, seed=0x5EED0002)
    c = fvk.grd_solv(m)

    def flow(**kw):
        cfg = fvk.flow_cfg(simpler=0, imat_u=0, imat_p=0, urf_p=0.3, **kw)
        fl = fvk.Flow(c, cfg)
        fl.set_bnd_u(31, [1.0, 0.0, 0.0])
        return fl

    modes = ("eager",) if args.no_graph else ("eager", "graph")
    legs_on = set(args.simple_legs.split(","))
    for mode in modes if "reference" in legs_on else ():
        log("SIMPLE %d^3: reference pressure solves, %s" % (n, mode))
        fl = flow()
        run = fl.run if mode == "eager" else fl.run_graph
        run(1, 2)
        c.sync()
        t0 = time.perf_counter()
        h = run(3, args.simple_its)
        c.sync()
        dt = (time.perf_counter() - t0) / args.simple_its
        out["reference_pressure_" + mode] = dict(outer_its_per_s=1.0 / dt, ms_per_outer=dt * 1e3,
                                                 mx_res_n_last=h[-1]["mx_res_n"])
        fl.destroy()
    # the reference's shipped configuration (cfd_setup.f90:528,581,662,742):
    # SIMPLER pressure, DILU momentum and DIC pressure preconditioners, urf_p
    # 0.1; factor and triangular sweeps in one launch each (the default:
    # blocks owning ranges of rows, values inside a range through LDS), and
    # for comparison one launch with rows taken in level order in chunks and
    # one launch per level.  First one application of the DIC / DILU factor
    # (both sweeps) in the three forms.
    if "shipped" in legs_on:
        fvk.sys_random(c, 0x5EED0003, 1.0)
        sw = {"levels": c.levels()}
        for imat, nvc in ((1, 1), (2, 3)):
            key = "imat%d_nvc%d" % (imat, nvc)
            fvk.set_param(c, "sweep_one_launch", 0)
            sw[key] = {"per_level_launches_us": fvk.bench_sweep(c, imat, nvc, 5) * 1e6}
            fvk.set_param(c, "sweep_one_launch", 1)
            fvk.set_param(c, "sweep_parts", 0)
            sw[key]["one_launch_chunks_us"] = fvk.bench_sweep(c, imat, nvc, 20) * 1e6
            fvk.set_param(c, "sweep_parts", 1)
            sw[key]["one_launch_ranges_us"] = fvk.bench_sweep(c, imat, nvc, 20) * 1e6
        out["sweep_apply"] = sw
        log("SIMPLE %d^3: factor application %s" % (n, json.dumps(sw)))
    shipped = ([(m_, "ranges") for m_ in modes] + [("eager", "chunks"), ("eager", "per_level")]
               if "shipped" in legs_on else [])
    for mode, form in shipped:
        log("SIMPLE %d^3: shipped configuration, %s, sweeps: %s" % (n, mode, form))
        fvk.set_param(c, "sweep_one_launch", int(form != "per_level"))
        fvk.set_param(c, "sweep_parts", int(form == "ranges"))
        cfg = fvk.flow_cfg()
        fl = fvk.Flow(c, cfg)
        fl.set_bnd_u(31, [1.0, 0.0, 0.0])
        run = fl.run if mode == "eager" else fl.run_graph
        run(1, 2)
        c.sync()
        its = args.simple_its
        t0 = time.perf_counter()
        h = run(3, its)
        c.sync()
        dt = (time.perf_counter() - t0) / its
        key = "shipped_" + mode + {"ranges": "", "chunks": "_chunks", "per_level": "_per_level_launches"}[form]
        out[key] = dict(outer_its_per_s=1.0 / dt, ms_per_outer=dt * 1e3, outer_iterations=its,
                        mx_res_n_last=h[-1]["mx_res_n"], levels=c.levels())
        fl.destroy()
    fvk.set_param(c, "sweep_one_launch", 1)
    fvk.set_param(c, "sweep_parts", 1)
    # C2: outer iterations 1..simple_tol_its with every pressure-correction
    # system solved to the FP64 relative L1 residual tol from x = 0, plain
    # FP64 PCG against the learned FP32-inner solve, replayed from HIP
    # graphs; the fields after iteration 3 go to the oracle comparison
    legs = [("plain_pcg64", 3, None)] if "c2" in legs_on else []
    if model is not None and legs:
        legs.append(("dc32_learned", 2, model))
    states3, final = {}, {}
    N = max(args.simple_tol_its, 4)
    for name, ps, mdl in legs:
        fl = flow(p_solver=ps, p_tol=args.tol)
        if mdl is not None:
            fl.set_model(mdl, train_steps=0)
        run = fl.run if args.no_graph else fl.run_graph
        c.sync()
        t0 = time.perf_counter()
        h = run(1, 3)
        c.sync()
        t3 = time.perf_counter() - t0
        states3[name] = fl.state()
        t0 = time.perf_counter()
        h += run(4, N - 3)
        c.sync()
        dt = t3 + time.perf_counter() - t0
        final[name] = fl.state()
        out["c2_" + name] = dict(outer_iterations=N, seconds=dt, ms_per_outer=dt * 1e3 / N,
                                 outer_its_per_s=N / dt, p_iters=[x["it_pp"] for x in h],
                                 mx_res_n_last=h[-1]["mx_res_n"])
        log("SIMPLE %d^3 C2 %s: %d outer iterations in %.2f s, p' iterations %s..%s"
            % (n, name, N, dt, h[1]["it_pp"], h[-1]["it_pp"]))
        fl.destroy()
    if len(final) == 2:
        a, b = final["plain_pcg64"], final["dc32_learned"]
        out["c2_field_rel_diff_learned_vs_plain"] = {
            k: float(np.max(np.abs(a[k] - b[k])) / np.max(np.abs(a[k]))) for k in ("u_cell", "p", "dpdx", "spd")}
        out["c2_speedup_learned_vs_plain"] = (out["c2_plain_pcg64"]["seconds"]
                                              / out["c2_dc32_learned"]["seconds"])
    c.destroy()
    m.free()
    log("SIMPLE %d^3: %s" % (n, json.dumps(out)))
    return dict(workload="SIMPLE outer iterations, %d^3 lid-driven cavity (+-2%% jitter), Re=100, urf_u 0.7, "
                         "urf_p 0.3, momentum BiCGStab (3 RHS batched, Jacobi, 5 its); c2_*: %d outer iterations "
                         "with pressure-correction solves to FP64 relative L1 residual %g" % (n, N, args.tol),
                **out), states3


def channel_leg(fvk, args):
    """SIMPLE on the seeded-Delaunay 2D channel (length 10 x height 1, inlet
    u = 1 (code 11), outlet (21), walls (32), Re = 500), cells in RCM order,
    momentum BiCGStab with u, v batched, relaxed pressure update, outer
    iterations replayed from HIP graphs; pressure-correction systems solved to
    an FP64 relative residual of tol by plain FP64 Jacobi-PCG and by the
    learned FP32-inner solve.  The network is trained (untimed) on small
    channels of other seeds: one Adam step on the pressure-correction system
    of each outer iteration of a SIMPLE run, interleaved with lognormal
    systems; then online on the timed channel's own systems during the
    untimed outer iterations.  Both flows then run the same timed outer
    iterations and their fields are compared."""
    from programs_amd import meshgen
    t0 = time.time()
    model = fvk.Model(seed=0x5EED0007)
    ms = fvk.Mesh.from_cells(2, *meshgen.channel_delaunay(args.channel_train_tri, seed=0x5EED0008))
    cs = fvk.grd_solv(ms, rcm=True)
    ml = fvk.Mesh.from_cells(2, *meshgen.channel_delaunay(args.channel_train_tri, seed=0x5EED0009))
    cl = fvk.grd_solv(ml, rcm=True)
    fl = fvk.Flow(cs, fvk.flow_cfg(simpler=0, imat_u=0, imat_p=0, urf_p=0.3, mu=0.002, p_solver=2, p_tol=args.tol))
    fl.set_bnd_u(11, [1.0, 0.0])
    fl.set_model(model, train_steps=1, lr=args.channel_lr)
    n_log = 0
    for it in range(1, args.channel_pretrain + 1):
        fl.run(it, 1)
        for _ in range(args.channel_pretrain_lognormal):
            fvk.sys_random(cl, 0x9000 + n_log, args.sigma)
            model.train(cl, 1, lr=args.channel_lr)
            n_log += 1
        if it % 100 == 0:
            log("  channel network: %d SIMPLE and %d lognormal systems (%.1f s)" % (it, n_log, time.time() - t0))
    fl.destroy()
    cs.destroy()
    ms.free()
    cl.destroy()
    ml.free()
    train_s = time.time() - t0
    t0 = time.time()
    m = fvk.Mesh.from_cells(2, *meshgen.channel_delaunay(args.channel, seed=0x5EED0005))
    c = fvk.grd_solv(m, rcm=True)
    log("channel: %d triangles, %d faces (%.1f s)" % (m.n_vol, m.n_fce, time.time() - t0))
    out, fields = {}, {}
    warm, its = args.channel_warm, args.channel_its
    seg = args.channel_seg if args.channel_seg > 0 else its
    methods = [(n_, p_) for n_, p_ in (("plain_pcg64", 3), ("dc32_learned", 2)) if n_ in args.channel_methods.split(",")]
    dump = {}
    for name, ps in methods:
        cfg = fvk.flow_cfg(simpler=0, imat_u=0, imat_p=0, urf_p=0.3, mu=0.002, p_solver=ps, p_tol=args.tol)
        fl = fvk.Flow(c, cfg)
        fl.set_bnd_u(11, [1.0, 0.0])
        if ps == 2:
            fl.set_model(model, train_steps=args.channel_train, lr=args.channel_lr)
        h0 = []
        for k in range(warm):
            h0 += fl.run_graph(1 + k, 1)
            log("channel %s: untimed outer iteration %d, p' iterations %d" % (name, 1 + k, h0[-1]["it_pp"]))
        if ps == 2:
            fl.set_model(model, train_steps=0)
        # the timed iterations in segments of seg replays (a log line each)
        c.sync()
        t1 = time.perf_counter()
        h = []
        while len(h) < its:
            h += fl.run_graph(1 + warm + len(h), min(seg, its - len(h)))
            if len(h) < its:
                c.sync()
                log("channel %s: %d of %d timed outer iterations, %.1f s, p' iterations %d..%d"
                    % (name, len(h), its, time.perf_counter() - t1, min(x["it_pp"] for x in h),
                       max(x["it_pp"] for x in h)))
        c.sync()
        dt = (time.perf_counter() - t1) / its
        st = fl.state()
        fields[name] = {k: st[k] for k in ("u_cell", "p")}
        if args.channel_dump:
            dump[name + "_u_sample"] = st["u_cell"][::97]
            dump[name + "_p_sample"] = st["p"][::97]
            dump[name + "_p_iters"] = np.array([x["it_pp"] for x in h0 + h])
            dump[name + "_mx_res_n"] = np.array([x["mx_res_n"] for x in h0 + h])
        out["pressure_to_tol_" + name] = dict(ms_per_outer=dt * 1e3, outer_its_per_s=1.0 / dt,
                                              p_iters_untimed=[x["it_pp"] for x in h0],
                                              p_iters=[x["it_pp"] for x in h])
        log("channel %s: %.1f ms per outer, p' iterations %s (untimed %s)"
            % (name, dt * 1e3, out["pressure_to_tol_" + name]["p_iters"],
               out["pressure_to_tol_" + name]["p_iters_untimed"]))
        fl.destroy()
    model.destroy()
    if args.channel_dump:
        np.savez_compressed(args.channel_dump, **dump)
    if len(fields) == 2:
        a, b = fields["plain_pcg64"], fields["dc32_learned"]
        out["field_rel_diff"] = {k: float(np.max(np.abs(a[k] - b[k])) / np.max(np.abs(a[k]))) for k in a}
        out["speedup_vs_plain"] = (out["pressure_to_tol_plain_pcg64"]["ms_per_outer"]
                                   / out["pressure_to_tol_dc32_learned"]["ms_per_outer"])
    c.destroy()
    m.free()
    out["training"] = dict(seconds=train_s, triangles=args.channel_train_tri, simple=args.channel_pretrain,
                           lognormal=n_log, online_steps_per_system=args.channel_train, online_systems=warm)
    return dict(workload="SIMPLE outer iterations %d..%d, seeded-Delaunay 2D channel (10 x 1, Re=500), %d triangles "
                         "in RCM order, relaxed pressure update, pressure-correction solves to FP64 relative residual "
                         "%g" % (warm + 1, warm + its, args.channel, args.tol), **out)


def large_leg(fvk, args, model):
    """The SpMV kernels on a mesh whose working set exceeds the 256 MB
    Infinity Cache (args.large^3 cells, +-20% jitter, built on the GPU): the
    plain FP64 SpMV and the learned path's k_mx_q, per launch."""
    n = args.large
    fvk.mesh_build_device(0)
    try:
        m = fvk.Mesh.box(n, n, n, jitter=args.jitter, seed=0x5EED00E1)
    finally:
        fvk.mesh_build_device(-1)
    c = fvk.grd_solv(m)
    fvk.sys_random(c, 0x5EED00E2, args.sigma)
    nnz = m.n_vol + 2 * (m.n_fce - m.n_bf)
    t = fvk.bench_spmv(c, 100)
    b = nnz * 12 + 4 * (m.n_vol + 1) + 16 * m.n_vol
    out = dict(cells=m.n_vol, spmv_fp64=dict(us_per_launch=t * 1e6, bytes=b, gbs=b / t / 1e9, frac=b / t / PEAK_HBM))
    if model is not None:
        model.build_factor(c)
        r = fvk.sys_solve(c, fvk.SOLVE_DC32_FACTOR, 0, args.tol, 200000, time_kernel=True)
        out["k_mx_q"] = dict(us_per_launch=r["kern_seconds"] * 1e6, bytes=r["kern_bytes"],
                             gbs=r["kern_bytes"] / r["kern_seconds"] / 1e9,
                             frac=r["kern_bytes"] / r["kern_seconds"] / PEAK_HBM, iters=r["iters"],
                             status=r["status"])
    log("large %d^3: %s" % (n, json.dumps(out)))
    c.destroy()
    m.free()
    return out


def flowlab_leg(fvk, args):
    """flowlab's serial CSR SIMPLE (flowlab/src/navierstokes.f90:515-641) on
    an n^3 lid-driven cavity (+-10% jitter, Re 100, urf 0.7 / 0.3, blending
    0.5): outer iterations per second with the factor and sweeps in one
    launch each (blocks owning row ranges; rows in level order in chunks)
    and with one launch per level, beside one outer iteration of flowlab's
    own Fortran (flang -O2, one core) on the same mesh."""
    from programs_amd import flowlab as fl
    from programs_amd import meshgen
    n = args.flowlab
    t0 = time.perf_counter()
    m = fl.FlMesh.from_records(3, *meshgen.box_records(n, jitter=0.1))
    t_mesh = time.perf_counter() - t0
    t0 = time.perf_counter()
    ctx = fl.FlCtx(m)
    t_ctx = time.perf_counter() - t0
    ub = fl.boundary_values(m, ctx.geometry()["norm"], "Wall", 1, (1, 0, 0, 1))
    par = (1.0, 0.01, 0.7, 0.3, 0.5)
    out = dict(workload="flowlab CSR SIMPLE, %d^3 lid-driven cavity (+-10%% jitter), Re 100, momentum BiCGStab and "
                        "pressure CG (25 iterations, reduction 0.01) with flowlab's factor" % n,
               cells=m.n_dmn, faces=m.n_fce, nnz=m.nnz, mesh_build_s=t_mesh, context_s=t_ctx)
    # one launch per sweep with row ranges owned by blocks (the default), with
    # rows taken in level order in chunks, and one launch per level
    forms = args.flowlab_forms.split(",")
    for key, one, parts in (("one_launch", 1, 1), ("one_launch_chunks", 1, 0), ("per_level_launches", 0, 0)):
        if {"one_launch": "ranges", "one_launch_chunks": "chunks", "per_level_launches": "per_level"}[key] not in forms:
            continue
        ctx.set_param("sweep_one_launch", one)
        if one:
            ctx.set_param("sweep_parts", parts)
        fw = fl.FlFlow(ctx, *par)
        fw.set(vel_bnd=ub)
        fw.run(2, 0.0)
        K = args.flowlab_its if parts else max(1, args.flowlab_its // 2)
        t0 = time.perf_counter()
        res0, gl = fw.run(K, 0.0)
        dt = time.perf_counter() - t0
        out[key] = dict(outer_iterations=K, ms_per_outer=dt * 1e3 / K, outer_its_per_s=K / dt,
                        norm_gl_last=float(gl[-1]))
        log("flowlab %d^3 %s: %.1f ms per outer iteration" % (n, key, dt * 1e3 / K))
        fw.destroy()
    ctx.set_param("sweep_one_launch", 1)
    ctx.set_param("sweep_parts", 1)
    if not args.no_cpu:
        from oracle import flowlab_ref
        if flowlab_ref.available():
            import tempfile
            with tempfile.TemporaryDirectory() as tmp:
                K = args.flowlab_its
                dt, setup = flowlab_ref.time_outer_iterations(m, list(par) + [0.0], ub, K, tmp)
                dt /= K
            out["cpu_baseline"] = dict(value=dt, unit="s per outer iteration", cores=1, kind="reference",
                                       sample="outer iterations 3..%d of flowlab's own Fortran modules (flang -O2, "
                                              "one core) driven by oracle/ref/flowlab_ref_driver.f90 on the same "
                                              "mesh, the iterations the GPU legs time: a %d-iteration run minus a "
                                              "2-iteration run (%.1f s)" % (K + 2, K + 2, setup))
            if "one_launch" in out:
                out["speedup_vs_cpu_reference"] = dt / (out["one_launch"]["ms_per_outer"] * 1e-3)
            log("flowlab %d^3 reference: %.1f s per outer iteration" % (n, dt))
    ctx.destroy()
    return out


def dropin_leg(fvk, args):
    """spray-simulation's SIMPLE in the shipped configuration on an n^3
    cavity (+-2% jitter, Re 100, urf_p 0.1) three ways on the same Gambit
    mesh: the reference's own Fortran (flang -O2, one core: its cfd_setup_eqn
    and cfd_setup_pp, cfd_setup.f90), the same Fortran host program with the
    drop-in modules (programs_amd/fortran: the assembly and solves on the GPU,
    the flow state pushed to and pulled from the host's grd structure around
    every cfd_setup call), and the native GPU flow (fvk.Flow, state resident
    in HBM).  Milliseconds per outer iteration over iterations 3..K+2 (the
    host drivers print each iteration's time of the two cfd_setup calls)."""
    import subprocess
    import tempfile
    from programs_amd import meshgen
    here = os.path.dirname(os.path.abspath(__file__))
    n, K = args.dropin, args.dropin_its
    N = K + 2
    m = fvk.Mesh.box(n, n, n, jitter=0.02, seed=0x5EED0001)
    vx, ty, vp, vr = m.cells()
    out = dict(workload="spray-simulation SIMPLE, shipped configuration (SIMPLER, DILU momentum, DIC pressure), "
                        "%d^3 lid-driven cavity (+-2%% jitter), Re 100, urf_p 0.1" % n,
               cells=m.n_vol, outer_iterations=K)

    def big_stack():
        import resource
        soft, hard = resource.getrlimit(resource.RLIMIT_STACK)
        resource.setrlimit(resource.RLIMIT_STACK, (hard, hard))

    with tempfile.TemporaryDirectory() as tmp:
        neu = os.path.join(tmp, "box.neu")
        meshgen.write_neu(neu, 3, vx, ty, vp, vr, meshgen.box_bnd(n, n, n), np.full(len(ty), 6))
        m.free()
        for key, exe in (("reference_fortran", "oracle/_ref/shipped/fvk_ref_driver"),
                         ("fortran_dropin", "oracle/_ref/dropin/fvk_ref_driver")):
            path = os.path.join(here, exe)
            if not os.path.exists(path):
                log("drop-in leg: %s missing, skipped" % exe)
                continue
            cmd = [path, os.path.join(here, "tests", "golden", "gambit.vrt"), neu, os.path.join(tmp, key + ".bin"),
                   str(N), "1.0", "0.01", "0.1", "31", "1.0", "0.0", "0.0"]
            t0 = time.perf_counter()
            p = subprocess.run(cmd, capture_output=True, text=True, cwd=tmp, timeout=600, preexec_fn=big_stack)
            wall = time.perf_counter() - t0
            if p.returncode != 0:
                raise RuntimeError("%s failed: %s" % (exe, (p.stdout + p.stderr)[-800:]))
            t = [float(ln.split()[2]) for ln in p.stdout.splitlines() if ln.startswith(" TIME17:")]
            ms = 1e3 * sum(t[2:N]) / K
            out[key] = dict(ms_per_outer=ms, process_s=wall)
            log("drop-in leg %d^3 %s: %.2f ms per outer iteration (process %.1f s)" % (n, key, ms, wall))
        m2 = fvk.Mesh.read_neu(neu)
        ctx = fvk.grd_solv(m2)
        fl = fvk.Flow(ctx, fvk.flow_cfg())
        fl.set_bnd_u(31, [1.0, 0.0, 0.0])
        fl.run(1, 2)
        ctx.sync()
        t0 = time.perf_counter()
        fl.run(3, K)
        ctx.sync()
        out["native_gpu"] = dict(ms_per_outer=(time.perf_counter() - t0) * 1e3 / K)
        log("drop-in leg %d^3 native: %.2f ms per outer iteration" % (n, out["native_gpu"]["ms_per_outer"]))
        fl.destroy()
        ctx.destroy()
        m2.free()
    if "reference_fortran" in out:
        r = out["reference_fortran"]["ms_per_outer"]
        for k in ("fortran_dropin", "native_gpu"):
            if k in out:
                out[k]["speedup_vs_reference_fortran"] = r / out[k]["ms_per_outer"]
    return out


def setup_leg(fvk, n, jitter):
    """Mesh connectivity and geometry of the n^3 box: on the host against on
    the GPU (same arrays bit for bit), and the device context build."""
    out = {}
    for dev in (-1, 0):
        fvk.mesh_build_device(dev)
        t0 = time.perf_counter()
        m = fvk.Mesh.box(n, n, n, jitter=jitter, seed=0x5EED0002)
        out["mesh_" + ("host" if dev < 0 else "gpu") + "_s"] = time.perf_counter() - t0
        if dev == 0:
            t0 = time.perf_counter()
            c = fvk.grd_solv(m)
            c.sync()
            out["context_s"] = time.perf_counter() - t0
            c.destroy()
        m.free()
    fvk.mesh_build_device(-1)
    log("set-up %d^3: %s" % (n, json.dumps(out)))
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=128)
    ap.add_argument("--jitter", type=float, default=0.2)
    ap.add_argument("--sigma", type=float, default=1.0)
    ap.add_argument("--tol", type=float, default=1e-8)
    ap.add_argument("--cpu-sample", type=int, default=300, help="iterations of the oracle's PCG timed")
    ap.add_argument("--cpu-ref-sample", type=int, default=500, help="iterations of the reference's PCG timed")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--methods", default="dc32,learned,fsai", help="methods timed beside the plain solve")
    ap.add_argument("--train-lognormal", type=int, default=1600)
    ap.add_argument("--train-simple", type=int, default=400)
    ap.add_argument("--lr", type=float, default=1e-2, help="Adam step of the network's training")
    ap.add_argument("--simple-n", type=int, default=128)
    ap.add_argument("--simple-its", type=int, default=20)
    ap.add_argument("--simple-tol-its", type=int, default=100,
                    help="outer iterations of the C2 leg (pressure solves to tol)")
    ap.add_argument("--no-simple", action="store_true")
    ap.add_argument("--simple-legs", default="reference,shipped,c2",
                    help="SIMPLE legs run: reference (reference pressure solves), shipped (SIMPLER, DILU/DIC), "
                         "c2 (pressure solves to tolerance)")
    ap.add_argument("--no-graph", action="store_true",
                    help="SIMPLE legs eager only (profiling: graph replays under rocprofv3's kernel trace "
                         "ended in a host segmentation fault inside the HIP runtime)")
    ap.add_argument("--large", type=int, default=160, help="cells per side of the beyond-cache SpMV leg (0: off)")
    ap.add_argument("--tight", type=float, default=1e-12, help="tolerance of the solution-agreement leg (0: off)")
    ap.add_argument("--channel", type=int, default=0, help="triangles of the optional channel leg (0: off)")
    ap.add_argument("--channel-warm", type=int, default=6)
    ap.add_argument("--channel-its", type=int, default=3)
    ap.add_argument("--channel-train", type=int, default=5, help="online Adam steps per untimed system")
    ap.add_argument("--channel-lr", type=float, default=3e-3, help="Adam step of the channel network's training")
    ap.add_argument("--channel-train-tri", type=int, default=40000)
    ap.add_argument("--channel-pretrain", type=int, default=300, help="SIMPLE outer iterations of the training run")
    ap.add_argument("--channel-pretrain-lognormal", type=int, default=3, help="lognormal systems per outer")
    ap.add_argument("--channel-only", action="store_true", help="run the channel leg alone")
    ap.add_argument("--channel-methods", default="plain_pcg64,dc32_learned",
                    help="pressure solvers of the channel leg (comma list)")
    ap.add_argument("--channel-seg", type=int, default=0,
                    help="timed outer iterations per graph replay call (0: all in one), a log line after each")
    ap.add_argument("--channel-dump", default="",
                    help="npz of every 97th cell's u and p and the residual histories of each method")
    ap.add_argument("--simple-only", action="store_true", help="run the SIMPLE leg alone (no network)")
    ap.add_argument("--dropin", type=int, default=24,
                    help="cells per side of the Fortran drop-in leg (0: off; the reference's own grd_conn is O(n^2) "
                         "in the faces)")
    ap.add_argument("--dropin-its", type=int, default=5)
    ap.add_argument("--dropin-only", action="store_true", help="run the Fortran drop-in leg alone")
    ap.add_argument("--flowlab", type=int, default=96, help="cells per side of the flowlab CSR SIMPLE leg (0: off)")
    ap.add_argument("--flowlab-its", type=int, default=6)
    ap.add_argument("--flowlab-forms", default="ranges,chunks,per_level",
                    help="forms of the factor and sweeps the flowlab leg times (comma list)")
    ap.add_argument("--flowlab-only", action="store_true", help="run the flowlab leg alone")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (the sharding design is "
              "described with the design notes); run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    from programs_amd import fvk
    if args.channel_only:
        if args.channel <= 0:
            sys.exit("--channel-only needs --channel N (triangles)")
        print(json.dumps({"channel": channel_leg(fvk, args)}), flush=True)
        return
    if args.dropin_only:
        print(json.dumps({"dropin": dropin_leg(fvk, args)}), flush=True)
        return
    if args.flowlab_only:
        print(json.dumps({"flowlab": flowlab_leg(fvk, args)}), flush=True)
        return
    if args.simple_only:
        # the SIMPLE leg alone (for kernel profiles), plain pressure solves
        simple, _ = simple_leg(fvk, args, None)
        # a 1 GiB device copy: the byte counters' calibration in the same run
        simple["copy_gbs"] = 2 * (1 << 30) / fvk.bench_copy(1 << 30, 5) / 1e9
        print(json.dumps({"simple": simple}), flush=True)
        return
    n = args.n
    t0 = time.time()
    mesh = fvk.Mesh.box(n, n, n, jitter=args.jitter, seed=0x5EED0002)
    log("mesh %d^3: %d cells, %d faces (%.1f s)" % (n, mesh.n_vol, mesh.n_fce, time.time() - t0))
    ctx = fvk.grd_solv(mesh)
    fvk.sys_random(ctx, 0x5EED0003, args.sigma)
    log("context + system ready (%.1f s)" % (time.time() - t0))

    def timed(method, name, model=None):
        # one step: (factor from the network +) solve
        def step():
            if model is not None:
                model.build_factor(ctx)
            return fvk.sys_solve(ctx, method, 0, args.tol, 200000)
        for _ in range(args.warmup):
            r = step()
            log("%s warmup solve: %d iterations, %.3f s device" % (name, r["iters"], r["seconds"]))
        ctx.sync()
        t0 = time.perf_counter()
        res = []
        for _ in range(args.steps):
            res.append(step())
        ctx.sync()
        wall = (time.perf_counter() - t0) / args.steps
        r = res[-1]
        log("%s timed: %.4f s per solve (wall), %.4f s device, %d iterations, true L1 %.2e L2 %.2e"
            % (name, wall, r["seconds"], r["iters"], r["res_true_l1"], r["res_true_l2"]))
        # every timed step must have converged on its own (no fallback) to the
        # FP64 true residual, or the time is not the method's
        r["steps_ok"] = all(x["status"] == fvk.FVK_OK and x["fallback"] == 0 and x["res_true_l1"] < args.tol
                            for x in res)
        r["step_iters"] = [x["iters"] for x in res]
        if not r["steps_ok"]:
            log("%s: a timed step did not converge on its own: %s" % (
                name, [(x["status"], x["fallback"], x["res_true_l1"]) for x in res]))
        return wall, r

    wall, r = timed(fvk.SOLVE_PCG64, "plain_pcg64_jacobi")
    methods = {"plain_pcg64_jacobi": dict(s=wall, iters=r["iters"], res_true_l1=r["res_true_l1"],
                                          res_true_l2=r["res_true_l2"], steps_ok=r["steps_ok"])}
    x_plain = fvk.sys_x(ctx)
    model = None
    if "learned" in args.methods.split(","):
        model, tinfo = train_model(fvk, args)
        w3, r3 = timed(fvk.SOLVE_DC32_FACTOR, "dc32_learned", model)
        x_learned = fvk.sys_x(ctx)
        rk3 = fvk.sys_solve(ctx, fvk.SOLVE_DC32_FACTOR, 0, args.tol, 200000, time_kernel=True)
        methods["dc32_learned"] = dict(s=w3, iters=r3["iters"], res_true_l1=r3["res_true_l1"],
                                       res_true_l2=r3["res_true_l2"], fallback=r3["fallback"],
                                       steps_ok=r3["steps_ok"], step_iters=r3["step_iters"],
                                       x_rel_diff_vs_plain=float(np.max(np.abs(x_learned - x_plain))
                                                                 / np.max(np.abs(x_plain))),
                                       inner_iters=r3["inner_iters"], outer_res=r3["outer_res"],
                                       speedup_vs_plain=wall / w3, k_mx_q_us=rk3["kern_seconds"] * 1e6,
                                       k_mx_q_gbs=rk3["kern_bytes"] / rk3["kern_seconds"] / 1e9,
                                       k_mx_q_bytes=rk3["kern_bytes"],
                                       training=tinfo)
    if model is not None and args.tight > 0:
        # the solutions of the plain and the learned solve agree to what the
        # tolerance fixes (cond x tol); at a tighter tolerance, both timed
        tight = {}
        xs = {}
        for name, meth in (("plain_pcg64_jacobi", fvk.SOLVE_PCG64), ("dc32_learned", fvk.SOLVE_DC32_FACTOR)):
            if meth == fvk.SOLVE_DC32_FACTOR:
                model.build_factor(ctx)
            ctx.sync()
            t0 = time.perf_counter()
            rt = fvk.sys_solve(ctx, meth, 0, args.tight, 400000)
            ctx.sync()
            tight[name] = dict(s=time.perf_counter() - t0, iters=rt["iters"], res_true_l1=rt["res_true_l1"],
                               status=rt["status"], fallback=rt["fallback"])
            xs[name] = fvk.sys_x(ctx)
        a, b = xs["plain_pcg64_jacobi"], xs["dc32_learned"]
        tight["tol"] = args.tight
        tight["x_rel_diff"] = float(np.max(np.abs(a - b)) / np.max(np.abs(a)))
        tight["speedup"] = tight["plain_pcg64_jacobi"]["s"] / tight["dc32_learned"]["s"]
        methods["tight"] = tight
        log("tight %g: %s" % (args.tight, json.dumps(tight)))
    if "fsai" in args.methods.split(","):
        # the closed-form FSAI factor on the same pattern: the method the
        # network learns, built without it (factor build inside the step)
        def fsai_step():
            fvk.factor_fsai(ctx)
            return fvk.sys_solve(ctx, fvk.SOLVE_DC32_FACTOR, 0, args.tol, 200000)
        for _ in range(args.warmup):
            fsai_step()
        ctx.sync()
        t0 = time.perf_counter()
        rf = [fsai_step() for _ in range(args.steps)]
        ctx.sync()
        wf = (time.perf_counter() - t0) / args.steps
        methods["dc32_fsai"] = dict(s=wf, iters=rf[-1]["iters"], res_true_l1=rf[-1]["res_true_l1"],
                                    speedup_vs_plain=wall / wf,
                                    steps_ok=all(x["status"] == fvk.FVK_OK and x["fallback"] == 0 and
                                                 x["res_true_l1"] < args.tol for x in rf))
        log("dc32_fsai timed: %.4f s per solve, %d iterations" % (wf, rf[-1]["iters"]))
        if "dc32_learned" in methods:
            methods["dc32_learned"]["iters_vs_fsai"] = methods["dc32_learned"]["iters"] / rf[-1]["iters"]
    if "dc32" in args.methods.split(","):
        w2, r2 = timed(fvk.SOLVE_DC32, "dc32_jacobi")
        rk2 = fvk.sys_solve(ctx, fvk.SOLVE_DC32, 0, args.tol, 200000, time_kernel=True)
        methods["dc32_jacobi"] = dict(s=w2, iters=r2["iters"], res_true_l1=r2["res_true_l1"],
                                      res_true_l2=r2["res_true_l2"], speedup_vs_plain=wall / w2,
                                      inner_iters=r2["inner_iters"], outer_res=r2["outer_res"],
                                      k_mx_pq_us=rk2["kern_seconds"] * 1e6,
                                      k_mx_pq_gbs=rk2["kern_bytes"] / rk2["kern_seconds"] / 1e9)

    # dominant kernel, timed live per launch with HIP events on the solve stream
    rk = fvk.sys_solve(ctx, fvk.SOLVE_PCG64, 0, args.tol, 200000, time_kernel=True)
    kb, ks = rk["kern_bytes"], rk["kern_seconds"]
    roof64 = dict(bound="hbm", kernel="k_cg_pq (p = z + beta p; q = A p; p.q)", achieved=kb / ks / 1e9,
                peak=PEAK_HBM / 1e9, unit="GB/s", frac=kb / ks / PEAK_HBM, traffic=traffic_bytes("k_cg_pq"),
                bytes_per_launch=kb, us_per_launch=ks * 1e6, launches=rk["kern_launches"])

    # stand-alone FP64 SpMV, CSR byte model nnz(w+4) + 4(n+1) + 2wn
    t_spmv = fvk.bench_spmv(ctx, 200)
    nnz = mesh.n_vol + 2 * (mesh.n_fce - mesh.n_bf)
    b_spmv = nnz * 12 + 4 * (mesh.n_vol + 1) + 16 * mesh.n_vol

    copy_bytes = 1 << 30
    t_copy = fvk.bench_copy(copy_bytes, 20)
    # the bandwidth ceiling: streaming probes at 4 GiB (beyond the 256 MB
    # Infinity Cache), four 16-byte loads in flight per thread, and eight
    # non-temporal ones from a grid that is resident at once
    sb = 4 << 30
    ceiling = dict(bytes=sb, copy_gbs=2 * sb / fvk.bench_stream("copy", sb, 10) / 1e9,
                   read_gbs=sb / fvk.bench_stream("read", sb, 10) / 1e9,
                   write_gbs=sb / fvk.bench_stream("write", sb, 10) / 1e9,
                   copy_nt_gbs=2 * sb / fvk.bench_stream("copy_nt", sb, 10) / 1e9,
                   read_nt_gbs=sb / fvk.bench_stream("read_nt", sb, 10) / 1e9)
    ceiling["best_read_gbs"] = max(ceiling["read_gbs"], ceiling["read_nt_gbs"])
    ceiling["best_copy_gbs"] = max(ceiling["copy_gbs"], ceiling["copy_nt_gbs"])
    log("streaming probes at 4 GiB: copy %.0f / %.0f (nt), read %.0f / %.0f (nt), write %.0f GB/s"
        % (ceiling["copy_gbs"], ceiling["copy_nt_gbs"], ceiling["read_gbs"], ceiling["read_nt_gbs"],
           ceiling["write_gbs"]))

    # headline: the learned-preconditioned mixed-precision solve when timed
    head = methods.get("dc32_learned")
    workload = ("pressure-correction system on the %d^3 cavity mesh jittered +-%g%% (%d cells, %d faces), "
                "a_f = s/l_pn*lognormal(0,%g), zero-mean N(0,1) RHS, solved from x=0 to FP64 relative L1 "
                "residual %g" % (n, 100 * args.jitter, mesh.n_vol, mesh.n_fce, args.sigma, args.tol))
    if head:
        value, hr = head["s"], dict(iters=head["iters"], res_true_l1=head["res_true_l1"],
                                    res_true_l2=head["res_true_l2"])
        method = ("dc32_learned: network factor G (per-face MLP, bf16 MFMA) + FP32-inner PCG with z = G G^T r "
                  "in FP64 defect correction; one step = factor build + solve")
        # the factor's L and Delta^2 and the hand-off vector t are FP16, the
        # inner Krylov vectors and the SpMV FP32, the judge (true residual,
        # solution) FP64
        dtype = "f16-factor/f32-inner/f64-outer"
        # bytes per launch by two models: the compressed format k_mx_q reads
        # (half of Ah, 16-bit columns, one-byte slot index) and the FP32 CSR
        # SpMV model nnz (w + 4) + 4 (n + 1) + 2 w n, w = 4
        csr32 = nnz * 8 + 4 * (mesh.n_vol + 1) + 8 * mesh.n_vol
        us = head["k_mx_q_us"]
        roof = dict(bound="hbm", kernel="k_mx_q (FP32: q = Ah p; p.q)",
                    achieved=head["k_mx_q_gbs"], peak=PEAK_HBM / 1e9, unit="GB/s",
                    frac=head["k_mx_q_gbs"] * 1e9 / PEAK_HBM, traffic=traffic_bytes("k_mx_q"),
                    bytes_per_launch=head["k_mx_q_bytes"], us_per_launch=us,
                    bytes_model="compressed format: 8n + 4.5 nnz_off (p, q FP32; upper entries FP32 + 16-bit "
                                "column; lower entries 16-bit column + 1-byte slot)",
                    csr_fp32_bytes=csr32, csr_fp32_achieved=csr32 / (us * 1e-6) / 1e9,
                    csr_fp32_frac=csr32 / (us * 1e-6) / PEAK_HBM)
    else:
        value, hr = wall, dict(iters=r["iters"], res_true_l1=r["res_true_l1"], res_true_l2=r["res_true_l2"])
        method = "plain_pcg64_jacobi"
        dtype = "f64"
        roof = roof64
    out = {
        "metric": METRIC,
        "value": value,
        "unit": "s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": value * 1e3,
        "higher_is_better": False,
        "scaling": "strong",
        "vs_baseline": None,
        "dtype": dtype,
        "data": "synthetic",
        "config": {"workload": workload, "method": method, "cells": mesh.n_vol, "nnz": nnz},
        "plain_pcg64_s": wall,
        "speedup_vs_plain": wall / value,
        **hr,
        "roofline": roof,
        "roofline_plain_pcg64": roof64,
        "spmv": dict(us_per_launch=t_spmv * 1e6, bytes=b_spmv, gbs=b_spmv / t_spmv / 1e9,
                     frac=b_spmv / t_spmv / PEAK_HBM),
        "copy_gbs": 2 * copy_bytes / t_copy / 1e9,
        "ceiling": ceiling,
        "methods": methods,
    }
    if args.large > 0:
        out["large"] = large_leg(fvk, args, model)
    out["setup"] = setup_leg(fvk, n, args.jitter)
    states3 = {}
    if not args.no_simple:
        out["simple"], states3 = simple_leg(fvk, args, model)
    if args.channel > 0:
        out["channel"] = channel_leg(fvk, args)
    if args.flowlab > 0:
        out["flowlab"] = flowlab_leg(fvk, args)
    if args.dropin > 0:
        out["dropin"] = dropin_leg(fvk, args)
    if not args.no_cpu:
        port = cpu_baseline(ctx, fvk, n, args.jitter, r["iters"], args.cpu_sample)
        ref = cpu_baseline_reference(ctx, mesh, fvk, n, r["iters"], args.cpu_ref_sample)
        out["cpu_baseline"] = ref if ref is not None else port
        out["cpu_baseline"].update(host_info())
        out["cpu_baseline_port"] = port
        out["cpu_baseline_all_cores"] = cpu_baseline_all_cores(ctx, fvk, n, args.jitter, r["iters"],
                                                               args.cpu_sample)
        if not args.no_simple and states3:
            out["cpu_baseline_simple"] = cpu_simple(args, states3)
    print(json.dumps(out), flush=True)
    ctx.destroy()


if __name__ == "__main__":
    main()
