#!/usr/bin/env python3
"""Throughput of the ModelG time step (PV2HBSplit23) on one MI355X.

    python bench.py [--steps K] [--warmup W] [--size 256] [--precision fp64|fp32]
                    [--graph] [--headline-only] [--batched] [--drift] [--em]
                    [--no-cpu-baseline] [--calibrate-copy]
                    [--no-autotune | --ideal-tile CAP TZ]

One "step" is PV2HBSplit::step(dt) (3 ideal substeps, 12 phi Metropolis
half-sweeps, 6 charge sub-sweeps; here 3 + 12 + 3 kernel launches) applied to
the whole lattice; the metric is site-updates/s = sites x steps / wall
seconds, inputs resident in HBM.  The default run reports the 256^3 FP64
lattice as `value`, and beside it the FP32-field form, the 96-replica
batched FP32 ensemble (and its one-at-a-time figure), the FP32-vs-FP64
observable drift over paired seeds, and the CPU oracle baselines.  Prints
one JSON line.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK = 8.0e12  # MI355X HBM3E, B/s
# Algorithmic bytes per site per launch, FP64 fields (halve for FP32):
# ideal substep (k_ideal_fused) read 80 + write 80; phi half-sweep
# (k_hb_march for large fields, k_hb_phi otherwise) read all phi 32 + write
# one parity 16; charge direction (k_charge_march, both sub-sweeps of one
# direction) read 48 + write 48.
ALG_BYTES = {"ideal": 160, "phi": 48, "charge": 96}
KERNEL = {"ideal": "k_ideal_fused", "phi": "k_hb_march", "charge": "k_charge_march"}
# Per site-update with one pass per dependency level of the reference
# (12 phi half-sweeps of 48 B, 6 charge sub-sweeps of 96 B): 3*160 + 12*48 + 6*96.
ALG_BYTES_PER_SITE_UPDATE = 1632
# The same for this schedule: 3 ideal, 12 phi and 3 charge launches.
SCHEDULE_BYTES_PER_SITE_UPDATE = 3 * 160 + 12 * 48 + 3 * 96  # 1344
SCHEDULE_BYTES_PER_SITE_UPDATE_FP32 = SCHEDULE_BYTES_PER_SITE_UPDATE
# HBM bytes per launch per site measured with rocprofv3 --pmc FETCH_SIZE and
# WRITE_SIZE (kernel, precision, N) -> (fetch, write), each scaled by the
# copy-kernel calibration (bench.py --calibrate-copy); the summary of those
# runs is profiles/modelg_256_hbm_traffic.md.
MEASURED_TRAFFIC = {
    ("k_ideal_fused", "fp64", 256): (99.71, 80.00),
    ("k_hb_march", "fp64", 256): (37.98, 16.01),
    ("k_charge_march", "fp64", 256): (51.67, 48.25),
    ("k_ideal_fused", "fp32", 256): (53.40, 40.00),
    ("k_hb_march", "fp32", 256): (17.76, 8.01),
    ("k_charge_march", "fp32", 256): (25.75, 24.15),
}


def tables(fp32, march):
    """(algorithmic bytes per site per launch, kernel name, schedule bytes per
    site-update) for the precision's kernel schedule; march: the phi
    half-sweep runs the z-march."""
    sched = SCHEDULE_BYTES_PER_SITE_UPDATE_FP32 if fp32 else SCHEDULE_BYTES_PER_SITE_UPDATE
    if march:
        return ALG_BYTES, KERNEL, sched
    return ALG_BYTES, dict(KERNEL, phi="k_hb_phi"), sched


def traffic(kernel, precision, N):
    """Measured HBM bytes per launch of the kernel at N^3, or None."""
    m = MEASURED_TRAFFIC.get((kernel, precision, N))
    return None if m is None else (m[0] + m[1]) * N ** 3


def calibrate_copy(mib=1024, reps=5):
    """A plain device copy of known size (torch.neg over mib MiB of FP64,
    out of place), run reps times: under rocprofv3 --pmc its FETCH_SIZE and
    WRITE_SIZE per dispatch, against mib MiB read and written, give the
    counters' scale factor."""
    import torch
    n = mib * 2 ** 20 // 8
    x = torch.ones(n, dtype=torch.float64, device="cuda")
    y = torch.empty_like(x)
    for _ in range(reps):
        torch.neg(x, out=y)
    torch.cuda.synchronize()
    t0, t1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    t0.record()
    for _ in range(reps):
        torch.neg(x, out=y)
    t1.record()
    torch.cuda.synchronize()
    sec = t0.elapsed_time(t1) * 1e-3 / reps
    print(json.dumps({"calibration": "copy", "bytes_read": n * 8, "bytes_written": n * 8,
                      "reps": 2 * reps, "kernel": "torch.neg (elementwise, FP64)",
                      "copy_GBps": 2 * n * 8 / sec / 1e9}), flush=True)


PARAMS = dict(mass0=-4.813, lam=1.0, H=0.01, chi=2.0, gamma=1.0)
DT = 0.04


def host_cpu():
    """(model name, nproc) of this host."""
    model = ""
    try:
        with open("/proc/cpuinfo") as f:
            model = next((ln.split(":", 1)[1].strip() for ln in f if ln.startswith("model name")), "")
    except OSError:
        pass
    return model, os.cpu_count()


def _time_oracle(lat, seconds, max_steps):
    lat.step(DT)  # first touch of the pages and thread start-up
    steps, t0 = 0, time.perf_counter()
    while True:
        lat.step(DT)
        steps += 1
        el = time.perf_counter() - t0
        if el > seconds or steps >= max_steps:
            return steps, el


def cpu_baseline(seconds_all=10.0, seconds_one=8.0):
    """The CPU oracle (a C restatement of the reference's stepper) timed on
    this host on bounded samples:
      * all cores (at most 16, the CPU share of one GPU), per-site streams,
        the headline 256^3 FP64 lattice: its loops split over z-planes, as
        MPI ranks split the reference's lattice;
      * one core in the reference's own RNG order (one global engine, the
        sites in k, j, i order: the single-rank reference trajectory) on a
        64^3 lattice -- the stand-in for the reference's CPU build, whose
        PETSc and FFTW dependencies are absent here."""
    from oracle import oracle as orc
    model, nproc = host_cpu()
    cores = max(1, min(16, nproc or 1))
    kw = dict(mass0=PARAMS["mass0"], lam=PARAMS["lam"], H=PARAMS["H"], chi=PARAMS["chi"],
              seed=42)
    orc.set_threads(cores)
    N = 256
    lat = orc.Lattice(N, rng_mode=orc.RNG_PER_SITE, **kw)
    lat.init_gaussian_charges()
    steps, el = _time_oracle(lat, seconds_all, 20)
    del lat
    orc.set_threads(1)
    n1 = 64
    lat = orc.Lattice(n1, rng_mode=orc.RNG_REFERENCE, **kw)
    lat.init_gaussian_charges()
    s1, e1 = _time_oracle(lat, seconds_one, 100)
    del lat
    host = f"{model}, nproc {nproc}"
    return {"value": N ** 3 * steps / el, "unit": "site-updates/s", "cores": cores,
            "kind": "port", "host": host,
            "sample": f"oracle/lgv_oracle.c per-site-stream mode, {steps} PV2HBSplit23 step(s) "
                      f"of a {N}^3 FP64 lattice after one untimed step ({el:.1f} s, {cores} "
                      f"OpenMP threads over z-planes)",
            "single_thread_reference_order": {
                "value": n1 ** 3 * s1 / e1, "unit": "site-updates/s", "cores": 1,
                "kind": "port", "host": host,
                "sample": f"oracle/lgv_oracle.c reference RNG order (one engine, k-j-i site "
                          f"order), {s1} step(s) of a {n1}^3 FP64 lattice after one untimed "
                          f"step ({e1:.1f} s, 1 thread)"}}


def batched_replicas(L):
    """96 independent replicas: 48 x 32^3, 32 x 48^3, 16 x 64^3; seed 1000+r;
    H_r uniform in [0.003, 0.03] from a xoroshiro128+ stream seeded with
    12345 by the reference's seed formula; m^2 = -4.813."""
    from langevina_amd.lgv import ensemble_params
    return ensemble_params([32] * 48 + [48] * 32 + [64] * 16, **PARAMS)


# Model A Euler-Maruyama (lgv_em_*): algorithmic bytes per site-step (FP64,
# 4 components): forward reads phi and writes phi (64); the reverse pass
# recomputes each checkpoint segment (64) and runs the adjoint stencil, which
# reads phi_n and lambda and writes lambda (96).
EM_FWD_BYTES, EM_BWD_BYTES = 64, 64 + 96


def bench_em(args, N=128, nsteps=500, ckpt=50, window=250, reps=3):
    """Forward + reverse of the Euler-Maruyama loss at 128^3 x 500 steps
    (checkpoints every 50): wall time per call with the inputs on the device,
    best of reps after one untimed call."""
    import torch
    from langevina_amd.em import EulerMaruyama
    dev = torch.device("cuda:0")
    g = torch.Generator(device="cpu").manual_seed(42)
    phi0 = (0.1 * torch.randn(4, N, N, N, generator=g, dtype=torch.float64)).to(dev)
    grad = torch.empty_like(phi0)
    em = EulerMaruyama(N, m2=-4.813, lam=1.0, H=0.01, gamma=1.0, dt=0.02, seed=42,
                       nsteps=nsteps, ckpt_every=ckpt)
    em.set_stream(torch.cuda.current_stream().cuda_stream)
    fwd, bwd = [], []
    for i in range(reps + 1):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        loss, _ = em.forward(phi0.data_ptr(), window)
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        gp = em.backward(grad.data_ptr())
        torch.cuda.synchronize()
        t2 = time.perf_counter()
        if i:
            fwd.append(t1 - t0)
            bwd.append(t2 - t1)
    em.close()
    sites = N ** 3
    f, b = min(fwd), min(bwd)
    out = {
        "metric": "Model A Euler-Maruyama site-steps/s (forward; reverse with checkpoints)",
        "value": sites * nsteps / f, "unit": "site-steps/s", "n_gpus": 1,
        "higher_is_better": True, "dtype": "f64",
        "data": "synthetic (phi0 ~ N(0, 0.1^2), seed 42)",
        "config": {"workload": f"Model A EM {N}^3 x {nsteps} steps, checkpoints every {ckpt}, "
                               f"loss window {window}", "m2": -4.813, "lambda": 1.0, "H": 0.01,
                   "dt": 0.02},
        "forward_ms": f * 1e3, "backward_ms": b * 1e3,
        "forward_alg_GBps": sites * nsteps * EM_FWD_BYTES / f / 1e9,
        "backward_alg_GBps": sites * nsteps * EM_BWD_BYTES / b / 1e9,
        "forward_roofline_frac": sites * nsteps * EM_FWD_BYTES / f / HBM_PEAK,
        "backward_roofline_frac": sites * nsteps * EM_BWD_BYTES / b / HBM_PEAK,
        "loss": loss, "grad_params": [float(x) for x in gp],
    }
    print(json.dumps(out), flush=True)


METRIC = "site-updates/sec (ModelG 256^3 FP64 vs FP32/batched) and % HBM roofline, 1 GPU"
T_START = time.perf_counter()


def progress(what):
    """One line per finished part on stderr (stdout carries only the JSON)."""
    print(f"bench: {what} ({time.perf_counter() - T_START:.1f} s)", file=sys.stderr, flush=True)


def kernel_table(tm, alg, kern, sites, esz):
    """Per kernel class: total ms, launches, average launch and algorithmic
    GB/s from the HIP-event timing of a run."""
    return {k: {"kernel": kern[k], "ms_total": round(v[0], 3), "launches": v[1],
                "avg_us": round(v[0] * 1e3 / max(v[1], 1), 2),
                "alg_GBps": round(alg[k] * sites * esz / 8 /
                                  (v[0] * 1e-3 / max(v[1], 1)) / 1e9, 1) if v[1] else None}
            for k, v in tm.items() if k in alg}


def roofline_of(tm, alg, kern, sites, esz, precision, N=None):
    """The roofline object of the dominant kernel class (largest total time)."""
    dom = max(("ideal", "phi", "charge"), key=lambda k: tm[k][0])
    ms, launches = tm[dom]
    per_launch_s = ms * 1e-3 / max(launches, 1)
    bytes_per_launch = alg[dom] * sites * esz // 8
    achieved = bytes_per_launch / per_launch_s
    return {"bound": "hbm", "kernel": kern[dom], "achieved": achieved / 1e9,
            "peak": HBM_PEAK / 1e9, "unit": "GB/s", "frac": achieved / HBM_PEAK,
            "traffic": traffic(kern[dom], precision, N) if N else None,
            "alg_bytes_per_launch": bytes_per_launch, "avg_launch_us": per_launch_s * 1e6}


# the headline's timed run brackets the kernels of one step in this many
# with HIP events (a timed step keeps one stream and records 36 events)
TIMING_EVERY = 4


def pin_phi_kernel(L, ctx, phi_kernel):
    """Sets the phi kernel ('site' / 'march'); returns the autotuner's march
    argument (None: the tuner chooses)."""
    if phi_kernel is None:
        return None
    ctx.set_phi_kernel(L.PHI_MARCH if phi_kernel == "march" else L.PHI_SITE)
    return phi_kernel == "march"


def run_single(L, N, fp32, steps, warmup, graph=False, tune=True, ideal_tile=None,
               measure=False, phi_kernel=None):
    """One N^3 lattice (phi = 0, Gaussian charges): warmup steps, then steps
    timed with a host clock around a synchronised run and with HIP events per
    kernel class.  Returns the measured figures."""
    prec = L.FP32 if fp32 else L.FP64
    esz = 4 if fp32 else 8
    ctx = L.Context(L.model_params(N, seed=42, **PARAMS), precision=prec)
    # synthetic start: phi = 0 (ModelA::initialize), Gaussian charges with the
    # zero mode removed (ModelA::initialize_gaussian_charges)
    ctx.init_gaussian_charges()
    march = pin_phi_kernel(L, ctx, phi_kernel)
    tuned = {}
    if ideal_tile:
        from langevina_amd.autotune import set_ideal_tile
        set_ideal_tile(ctx, 0, *ideal_tile)
        tuned = {N: tuple(ideal_tile)}
    sweeps = {}
    if tune:
        # kernel tiles for this lattice size; the tuning steps are warmup
        from langevina_amd.autotune import autotune_ideal, autotune_sweeps
        if not ideal_tile:
            tuned = autotune_ideal(ctx, DT)
        sweeps = autotune_sweeps(ctx, DT, march=march)
    if graph:
        ctx.use_graph(True)
    ctx.run(DT, warmup)
    ctx.synchronize()
    # stream launches: HIP events bracket the kernels of every TIMING_EVERY-th
    # step of the timed run (the others run as in production); graph replay:
    # the timed run is untimed and a short pass of stream launches after it
    # gives the per-kernel figures
    ctx.timing_enable(0 if graph else TIMING_EVERY)
    ctx.synchronize()
    t0 = time.perf_counter()
    ctx.run(DT, steps)
    ctx.synchronize()
    el = time.perf_counter() - t0
    if graph:
        ctx.timing_enable(True)
        ctx.run(DT, max(5, steps // TIMING_EVERY))
    tm = ctx.timing()
    ctx.timing_enable(False)
    sites = N ** 3
    value = sites * steps / el
    march = ctx.phi_kernel() == L.PHI_MARCH
    alg, kern, sched = tables(fp32, march)
    out = {
        "value": value, "ms_per_step": el * 1e3 / steps, "steps": steps, "warmup": warmup,
        "dtype": "f32-storage/f64-arith" if fp32 else "f64",
        "workload": f"ModelG {N}^3 PV2HBSplit23 {'FP32 fields' if fp32 else 'FP64'}",
        "ideal_tile": list(tuned[N][:2]) if tuned else "default (32, 16)",
        "sweep_tiles": list(sweeps[N]) if sweeps else "default",
        "phi_kernel": kern["phi"],
        "kernel_timing": ("a pass of stream launches after the timed graph replays" if graph else
                          f"HIP events around the kernels of every {TIMING_EVERY}th timed step"),
        "step_roofline_frac": value * ALG_BYTES_PER_SITE_UPDATE * esz / 8 / HBM_PEAK,
        "step_schedule_bytes_frac": value * sched * esz / 8 / HBM_PEAK,
        "roofline": roofline_of(tm, alg, kern, sites, esz, "fp32" if fp32 else "fp64", N),
        "kernels": kernel_table(tm, alg, kern, sites, esz),
    }
    if measure:
        # the measurement pass (Measurer::measure), reported separately
        ctx.measure_full()
        tm0 = time.perf_counter()
        for _ in range(5):
            ctx.measure_full()
        meas_s = (time.perf_counter() - tm0) / 5
        out["measurement"] = {"ms": meas_s * 1e3, "alg_GBps": sites * 10 * esz / meas_s / 1e9,
                              "what": "lgv_measure_full: walls, phase walls, averages, device "
                                      "r2c of every row, rotate_to_vev; host copies included"}
    ctx.close()
    return out


def run_batched(L, fp32, steps, warmup, tune=True, graph=True, graph_steps=10,
                phi_kernel=None):
    """The 96-replica ensemble in one context, one HIP graph per step, against
    the same replicas one context (and one graph) at a time; each graph holds
    graph_steps steps; graph=False launches the kernels on the stream instead
    (for tracing profilers)."""
    from langevina_amd.autotune import (autotune_ideal, autotune_sweeps, set_ideal_tile,
                                        set_sweep_tiles)
    reps = batched_replicas(L)
    sites = sum(p.NX ** 3 for p in reps)
    prec = L.FP32 if fp32 else L.FP64
    esz = 4 if fp32 else 8
    ctx = L.Context(reps, precision=prec)
    for r in range(len(reps)):
        ctx.init_gaussian_charges(r)
    march = pin_phi_kernel(L, ctx, phi_kernel)
    tuned = autotune_ideal(ctx, DT) if tune else {}
    sweeps = {}
    if tune:
        sweeps = autotune_sweeps(ctx, DT, march=march)
    ctx.run(DT, 2)
    ctx.timing_enable(True)  # kernel breakdown from a short direct-launch pass
    ctx.run(DT, 5)
    tm = ctx.timing()
    ctx.timing_enable(False)
    ctx.use_graph(graph_steps if graph else 0)
    ctx.run(DT, warmup)
    ctx.synchronize()
    t0 = time.perf_counter()
    ctx.run(DT, steps)
    ctx.synchronize()
    el = time.perf_counter() - t0
    value = sites * steps / el
    march = ctx.phi_kernel() == L.PHI_MARCH
    ctx.close()
    seq_el = 0.0
    for p in reps:
        c1 = L.Context(p, precision=prec)
        c1.init_gaussian_charges()
        if tuned:
            cap, tz = tuned[p.NX][:2]
            set_ideal_tile(c1, 0, cap, tz)
            set_sweep_tiles(c1, 0, *sweeps[p.NX])
        c1.use_graph(graph_steps if graph else 0)
        c1.run(DT, 2)
        c1.synchronize()
        t0 = time.perf_counter()
        c1.run(DT, steps)
        c1.synchronize()
        seq_el += time.perf_counter() - t0
        c1.close()
    seq_value = sites * steps / seq_el
    alg, kern, _ = tables(fp32, march)
    return {
        "value": value, "ms_per_step": el * 1e3 / steps, "steps": steps, "warmup": warmup,
        "dtype": "f32-storage/f64-arith" if fp32 else "f64",
        "workload": "batched ModelG: 96 replicas (48x32^3, 32x48^3, 16x64^3), one HIP graph "
                    "per step",
        "sites": sites, "ideal_tiles": {str(k): list(v[:2]) for k, v in tuned.items()},
        "sweep_tiles": {str(k): list(v) for k, v in sweeps.items()},
        "phi_kernel": kern["phi"],
        "one_at_a_time": {"value": seq_value, "ms_per_step": seq_el * 1e3 / steps},
        "batched_over_one_at_a_time": value / seq_value,
        "step_roofline_frac": value * ALG_BYTES_PER_SITE_UPDATE * esz / 8 / HBM_PEAK,
        "roofline": roofline_of(tm, alg, kern, sites, esz, "fp32" if fp32 else "fp64"),
        "kernels": kernel_table(tm, alg, kern, sites, esz),
    }


def bench_batched(args):
    import langevina_amd as L
    b = run_batched(L, args.precision == "fp32", args.steps, args.warmup,
                    tune=not args.no_autotune, graph=not args.no_graph,
                    graph_steps=args.graph_steps, phi_kernel=args.phi_kernel)
    out = {"metric": METRIC, "value": b.pop("value"), "unit": "site-updates/s", "n_gpus": 1,
           "steps": args.steps, "warmup": args.warmup, "ms_per_step": b.pop("ms_per_step"),
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
           "dtype": b.pop("dtype"), "data": "synthetic (phi=0, Gaussian charges per replica)",
           "config": {"workload": b.pop("workload"), "dt": DT, "parallelism": "1 GPU",
                      "graph": not args.no_graph}}
    out.update(b)
    print(json.dumps(out), flush=True)


def precision_drift(L, N=256, nsteps=500, every=50, seeds=(42, 43, 44, 45)):
    """FP32 fields against FP64 on the same lattices, seeds and noise streams
    (the FP64 GPU step equals the oracle's to 1e-12): per seed, nsteps
    steps from phi = 0 with a measurement every `every`; per measurement
    <|phi|> = |phi_bar| and the k = 1 wall correlator G1 = V/3 sum_d sum_a
    |W_a(k=1)|^2 (W the device r2c / N of the z/y/x walls).  Per seed the
    series means of <|phi|> and G1 and the susceptibility V (<|phi_bar|^2> -
    <|phi_bar|>^2) over the series; reported as the relative FP32 - FP64
    difference, mean and standard error over the seeds, beside the
    seed-to-seed relative spread of the FP64 values (the statistical error
    of the observable itself)."""
    import numpy as np
    per = {k: {"fp64": [], "fp32": []} for k in ("absphi", "G1", "chi")}
    max_dev = {"absphi": 0.0, "G1": 0.0}
    V = N ** 3
    for seed in seeds:
        ctxs = {}
        for name, prec in (("fp64", L.FP64), ("fp32", L.FP32)):
            c = L.Context(L.model_params(N, seed=seed, **PARAMS), precision=prec)
            c.init_gaussian_charges()
            c.use_graph(True)
            ctxs[name] = c
        series = {k: {"absphi": [], "phi2": [], "G1": []} for k in ctxs}
        for _ in range(nsteps // every):
            for name, c in ctxs.items():
                c.run(DT, every)
                m = c.measure_full()
                pb = m["oavg"][:4]
                wk = m["wall_k"][:, :4, 1]
                series[name]["absphi"].append(float(np.sqrt(np.sum(pb * pb))))
                series[name]["phi2"].append(float(np.sum(pb * pb)))
                series[name]["G1"].append(float(V / 3.0 * np.sum(np.abs(wk) ** 2)))
        for c in ctxs.values():
            c.close()
        for name, ser in series.items():
            per["absphi"][name].append(float(np.mean(ser["absphi"])))
            per["G1"][name].append(float(np.mean(ser["G1"])))
            per["chi"][name].append(float(V * (np.mean(ser["phi2"]) -
                                               np.mean(ser["absphi"]) ** 2)))
        for obs in max_dev:
            a, b = np.array(series["fp64"][obs]), np.array(series["fp32"][obs])
            max_dev[obs] = max(max_dev[obs], float((np.abs(b - a) / np.abs(a)).max()))
        progress(f"drift seed {seed}")
    R = len(seeds)
    out = {"what": "FP32 fields vs FP64, paired seeds", "N": N, "steps": nsteps,
           "measure_every": every, "seeds": list(seeds), "dt": DT}
    for obs, v in per.items():
        a, b = np.array(v["fp64"]), np.array(v["fp32"])
        rel = (b - a) / a
        out[obs] = {"fp64": a.tolist(), "fp32": b.tolist(),
                    "rel_dev_mean": float(rel.mean()),
                    "rel_dev_sem": float(rel.std(ddof=1) / np.sqrt(R)) if R > 1 else None,
                    "fp64_rel_spread": float(a.std(ddof=1) / abs(a.mean())) if R > 1 else None}
        if obs in max_dev:
            out[obs]["max_rel_dev_per_measurement"] = max_dev[obs]
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--size", type=int, default=256)
    ap.add_argument("--precision", choices=["fp64", "fp32"], default="fp64",
                    help="the precision of the headline lattice (fp32: only that form)")
    ap.add_argument("--graph", action="store_true", help="replay a HIP graph per step")
    ap.add_argument("--no-graph", action="store_true",
                    help="--batched: launch on the stream instead of replaying graphs")
    ap.add_argument("--graph-steps", type=int, default=10,
                    help="steps per captured graph of the batched form")
    ap.add_argument("--batched", action="store_true",
                    help="only the 96 mixed-size replicas in one context")
    ap.add_argument("--headline-only", action="store_true",
                    help="only the 256^3 FP64 lattice (no FP32/batched forms, no drift)")
    ap.add_argument("--no-autotune", action="store_true")
    ap.add_argument("--ideal-tile", type=int, nargs=2, metavar=("CAP", "TZ"),
                    help="fix the ideal-kernel tile instead of autotuning it")
    ap.add_argument("--phi-kernel", choices=["site", "march"],
                    help="fix the phi half-sweep kernel instead of letting the tuner pick it")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="only run the copy kernel used to scale HBM counters")
    ap.add_argument("--em", action="store_true",
                    help="Model A Euler-Maruyama forward/reverse at 128^3 x 500 steps")
    ap.add_argument("--drift", action="store_true",
                    help="only the FP32-vs-FP64 observable drift (500 steps, paired seeds)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this build runs on one GPU only (the lattice and every replica "
              "live on a single MI355X); --gpus must be 1", file=sys.stderr)
        sys.exit(2)
    if args.calibrate_copy:
        calibrate_copy()
        return
    import langevina_amd as L
    if args.drift:
        print(json.dumps(precision_drift(L, N=args.size)), flush=True)
        return
    if args.em:
        bench_em(args)
        return
    if args.batched:
        bench_batched(args)
        return

    N = args.size
    fp32 = args.precision == "fp32"
    h = run_single(L, N, fp32, args.steps, args.warmup, graph=args.graph,
                   tune=not args.no_autotune, ideal_tile=args.ideal_tile, measure=True,
                   phi_kernel=args.phi_kernel)
    progress(f"{h['workload']}: {h['value']:.3e} site-updates/s")
    out = {
        "metric": METRIC, "value": h["value"], "unit": "site-updates/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": h["ms_per_step"],
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": h["dtype"],
        "data": "synthetic (phi=0, Gaussian charges, zero mode removed; seed 42)",
        "config": {"workload": h["workload"], "N": N, "dt": DT, "params": PARAMS,
                   "parallelism": "1 GPU", "graph": bool(args.graph),
                   "ideal_tile": h["ideal_tile"], "sweep_tiles": h["sweep_tiles"],
                   "phi_kernel": h["phi_kernel"], "kernel_timing": h["kernel_timing"]},
        "step_roofline_frac": h["step_roofline_frac"],
        "step_schedule_bytes_frac": h["step_schedule_bytes_frac"],
        "roofline": h["roofline"], "kernels": h["kernels"], "measurement": h["measurement"],
    }
    if not fp32 and not args.headline_only:
        # the other two forms of the throughput comparison, FP32 fields and
        # the batched ensemble, against this FP64 form, and what FP32 storage
        # does to the observables
        f = run_single(L, N, True, args.steps, args.warmup, graph=args.graph,
                       tune=not args.no_autotune)
        progress(f"{f['workload']}: {f['value']:.3e} site-updates/s")
        b = run_batched(L, True, args.steps, args.warmup, tune=not args.no_autotune,
                        graph_steps=args.graph_steps)
        progress(f"batched FP32: {b['value']:.3e} site-updates/s")
        for form in (f, b):
            form["over_fp64"] = form["value"] / h["value"]
        out["forms"] = {"fp32_fields": f, "batched_fp32": b}
        out["fp32_drift"] = precision_drift(L, N=N)
        progress("FP32 drift")
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline()
        progress("CPU baseline")
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
