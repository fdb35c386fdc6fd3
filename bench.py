#!/usr/bin/env python3
"""Throughput of the ModelG time step (PV2HBSplit23) on one MI355X.

    python bench.py [--steps K] [--warmup W] [--size 256] [--precision fp64|fp32]
                    [--graph] [--batched] [--drift] [--em] [--no-cpu-baseline]
                    [--calibrate-copy] [--no-autotune | --ideal-tile CAP TZ]

One "step" is PV2HBSplit::step(dt) (3 ideal substeps, 12 phi Metropolis
half-sweeps, 6 charge sub-sweeps; here 3 + 12 + 3 kernel launches) applied to
the whole lattice; the metric
is site-updates/s = sites x steps / wall seconds, inputs resident in HBM.
Prints one JSON line.
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
# (k_hb_march for FP64 fields, k_hb_phi for FP32) read all phi 32 + write
# one parity 16; charge direction (k_charge_march, both sub-sweeps of one
# direction) read 48 + write 48.
ALG_BYTES = {"ideal": 160, "phi": 48, "charge": 96}
KERNEL = {"ideal": "k_ideal_fused", "phi": "k_hb_march", "charge": "k_charge_march"}
ALG_BYTES_FP32 = ALG_BYTES
KERNEL_FP32 = dict(KERNEL, phi="k_hb_phi")
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
    ("k_ideal_fused", "fp64", 256): (95.15, 80.00),
    ("k_hb_march", "fp64", 256): (36.34, 16.01),
    ("k_charge_march", "fp64", 256): (51.91, 48.23),
    ("k_ideal_fused", "fp32", 256): (76.76, 40.00),
    ("k_hb_phi", "fp32", 256): (24.57, 8.12),
    ("k_charge_march", "fp32", 256): (25.70, 24.14),
}


# The library runs the phi z-march for FP64 fields above 1 GiB (one buffer,
# all replicas) and the one-thread-per-site phi kernel otherwise.
LARGE_FIELD_BYTES = 2 ** 30


def tables(fp32, field_bytes=None):
    """(algorithmic bytes per site per launch, kernel name, schedule bytes per
    site-update) for the precision's kernel schedule; field_bytes (one
    buffer) decides the phi kernel."""
    if fp32:
        return ALG_BYTES_FP32, KERNEL_FP32, SCHEDULE_BYTES_PER_SITE_UPDATE_FP32
    if field_bytes is not None and field_bytes <= LARGE_FIELD_BYTES:
        return ALG_BYTES, dict(KERNEL, phi="k_hb_phi"), SCHEDULE_BYTES_PER_SITE_UPDATE
    return ALG_BYTES, KERNEL, SCHEDULE_BYTES_PER_SITE_UPDATE


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


def cpu_baseline(seconds_target=12.0):
    """The CPU oracle (a C restatement of the reference's stepper) timed on
    this host on a bounded sample: 128^3 FP64 PV2HBSplit23 steps until
    ~seconds_target of work, its loops split over z-planes across the host's
    cores (at most 16, the CPU share of one GPU), as MPI ranks split the
    reference's lattice."""
    from oracle import oracle as orc
    N = 128
    cores = max(1, min(16, os.cpu_count() or 1))
    orc.set_threads(cores)
    lat = orc.Lattice(N, mass0=PARAMS["mass0"], lam=PARAMS["lam"], H=PARAMS["H"],
                      chi=PARAMS["chi"], seed=42, rng_mode=orc.RNG_PER_SITE)
    lat.init_gaussian_charges()
    lat.step(DT)  # first touch of the pages and thread start-up
    steps, t0 = 0, time.perf_counter()
    while True:
        lat.step(DT)
        steps += 1
        el = time.perf_counter() - t0
        if el > seconds_target or steps >= 40:
            break
    orc.set_threads(1)
    model = ""
    try:
        with open("/proc/cpuinfo") as f:
            model = next((ln.split(":", 1)[1].strip() for ln in f if ln.startswith("model name")), "")
    except OSError:
        pass
    return {"value": N ** 3 * steps / el, "unit": "site-updates/s", "cores": cores,
            "kind": "port", "host": f"{model}, nproc {os.cpu_count()}",
            "sample": f"oracle/lgv_oracle.c, {steps} PV2HBSplit23 step(s) of a {N}^3 FP64 "
                      f"lattice after one untimed step ({el:.1f} s, {cores} OpenMP threads)"}


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


def bench_batched(args):
    import langevina_amd as L
    from langevina_amd.autotune import autotune_ideal
    reps = batched_replicas(L)
    sites = sum(p.NX ** 3 for p in reps)
    prec = L.FP32 if args.precision == "fp32" else L.FP64
    ctx = L.Context(reps, precision=prec)
    for r in range(len(reps)):
        ctx.init_gaussian_charges(r)
    tuned = autotune_ideal(ctx, DT) if not args.no_autotune else {}
    ctx.timing_enable(True)  # kernel breakdown from a short direct-launch pass
    ctx.run(DT, 5)
    tm = ctx.timing()
    ctx.timing_enable(False)
    ctx.use_graph(True)
    ctx.run(DT, args.warmup)
    ctx.synchronize()
    t0 = time.perf_counter()
    ctx.run(DT, args.steps)
    ctx.synchronize()
    el = time.perf_counter() - t0
    value = sites * args.steps / el
    ctx.close()
    # the same replicas one context (and one graph) at a time
    seq_el = 0.0
    for p in reps:
        c1 = L.Context(p, precision=prec)
        c1.init_gaussian_charges()
        if tuned:
            from langevina_amd.autotune import set_ideal_tile
            cap, tz = tuned[p.NX][:2]
            set_ideal_tile(c1, 0, cap, tz)
        c1.use_graph(True)
        c1.run(DT, 2)
        c1.synchronize()
        t0 = time.perf_counter()
        c1.run(DT, args.steps)
        c1.synchronize()
        seq_el += time.perf_counter() - t0
        c1.close()
    seq_value = sites * args.steps / seq_el
    esz = 8 if prec == L.FP64 else 4
    alg, kern, _ = tables(prec == L.FP32, sites * 10 * esz)
    dom = max(("ideal", "phi", "charge"), key=lambda k: tm[k][0])
    ms, launches = tm[dom]
    per_launch_s = ms * 1e-3 / max(launches, 1)
    bytes_per_launch = alg[dom] * sites * esz // 8
    achieved = bytes_per_launch / per_launch_s
    out = {
        "metric": "site-updates/sec (ModelG 256^3 FP64 vs FP32/batched) and % HBM roofline, 1 GPU",
        "value": value, "unit": "site-updates/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": el * 1e3 / args.steps, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None,
        "dtype": "f32-storage/f64-arith" if prec == L.FP32 else "f64",
        "data": "synthetic (phi=0, Gaussian charges per replica)",
        "config": {"workload": "batched ModelG: 96 replicas (48x32^3, 32x48^3, 16x64^3), "
                               "one HIP graph per step",
                   "sites": sites, "dt": DT, "parallelism": "1 GPU",
                   "ideal_tiles": {str(k): list(v[:2]) for k, v in tuned.items()}},
        "one_at_a_time": {"value": seq_value, "ms_per_step": seq_el * 1e3 / args.steps},
        "batched_over_one_at_a_time": value / seq_value,
        "step_roofline_frac": value * ALG_BYTES_PER_SITE_UPDATE * esz / 8 / HBM_PEAK,
        "roofline": {"bound": "hbm", "kernel": kern[dom],
                     "achieved": achieved / 1e9, "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK, "traffic": None},
        "kernels": {k: {"kernel": kern[k], "launches": v[1],
                        "avg_us": round(v[0] * 1e3 / max(v[1], 1), 2)}
                    for k, v in tm.items() if k in alg and v[1]},
    }
    print(json.dumps(out), flush=True)


def precision_drift(args, steps=500, every=50):
    """FP32 fields against FP64 on the same lattice, seed and noise streams:
    500 steps, a measurement every 50.  Per measurement: <|phi|> = |phi_bar|,
    |phi_bar|^2 and the k = 1 wall correlator G1 = N^3/3 sum_d sum_a
    |phi~_a(k=1)|^2 (phi~ from the device r2c of the z/y/x walls, k in units
    of 2 pi / N).  The susceptibility is N^3 (<|phi_bar|^2> - <|phi_bar|>^2)
    over the measurement series."""
    import numpy as np
    import langevina_amd as L
    N = args.size
    ctxs = {}
    for name, prec in (("fp64", L.FP64), ("fp32", L.FP32)):
        c = L.Context(L.model_params(N, seed=42, **PARAMS), precision=prec)
        c.init_gaussian_charges()
        c.use_graph(True)
        ctxs[name] = c
    series = {k: {"absphi": [], "phi2": [], "G1": []} for k in ctxs}
    for _ in range(steps // every):
        for name, c in ctxs.items():
            c.run(DT, every)
            m = c.measure_full()
            pb = m["oavg"][:4]
            wk = m["wall_k"][:, :4, 1]
            series[name]["absphi"].append(float(np.sqrt(np.sum(pb * pb))))
            series[name]["phi2"].append(float(np.sum(pb * pb)))
            series[name]["G1"].append(float(N ** 3 / 3.0 * np.sum(np.abs(wk) ** 2)))
    for c in ctxs.values():
        c.close()
    out = {"drift": "FP32 fields vs FP64", "N": N, "steps": steps, "measure_every": every,
           "dt": DT, "params": PARAMS, "seed": 42}
    for obs in ("absphi", "G1"):
        a = np.array(series["fp64"][obs])
        b = np.array(series["fp32"][obs])
        rel = np.abs(b - a) / np.abs(a)
        out[obs] = {"fp64": a.tolist(), "fp32": b.tolist(), "max_rel_dev": float(rel.max()),
                    "final_rel_dev": float(rel[-1])}
    chi = {k: float(N ** 3 * (np.mean(v["phi2"]) - np.mean(v["absphi"]) ** 2))
           for k, v in series.items()}
    out["susceptibility"] = {"fp64": chi["fp64"], "fp32": chi["fp32"],
                             "rel_dev": abs(chi["fp32"] - chi["fp64"]) / abs(chi["fp64"])}
    print(json.dumps(out), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--size", type=int, default=256)
    ap.add_argument("--precision", choices=["fp64", "fp32"], default="fp64")
    ap.add_argument("--graph", action="store_true", help="replay a HIP graph per step")
    ap.add_argument("--batched", action="store_true",
                    help="96 mixed-size replicas in one context (use with --precision fp32)")
    ap.add_argument("--no-autotune", action="store_true")
    ap.add_argument("--ideal-tile", type=int, nargs=2, metavar=("CAP", "TZ"),
                    help="fix the ideal-kernel tile instead of autotuning it")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="only run the copy kernel used to scale HBM counters")
    ap.add_argument("--em", action="store_true",
                    help="Model A Euler-Maruyama forward/reverse at 128^3 x 500 steps")
    ap.add_argument("--drift", action="store_true",
                    help="report FP32-field observables against FP64 (500 steps)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this build runs on one GPU only (the lattice and every replica "
              "live on a single MI355X); --gpus must be 1", file=sys.stderr)
        sys.exit(2)
    if args.calibrate_copy:
        calibrate_copy()
        return
    if args.drift:
        precision_drift(args)
        return
    if args.em:
        bench_em(args)
        return
    if args.batched:
        bench_batched(args)
        return

    import numpy as np
    import langevina_amd as L

    N = args.size
    prec = L.FP64 if args.precision == "fp64" else L.FP32
    esz = 8 if prec == L.FP64 else 4
    ctx = L.Context(L.model_params(N, seed=42, **PARAMS), precision=prec)
    # synthetic start: phi = 0 (ModelA::initialize), Gaussian charges with the
    # zero mode removed (ModelA::initialize_gaussian_charges)
    ctx.init_gaussian_charges()
    tuned = {}
    if args.ideal_tile:
        from langevina_amd.autotune import set_ideal_tile
        set_ideal_tile(ctx, 0, *args.ideal_tile)
        tuned = {N: tuple(args.ideal_tile)}
    elif not args.no_autotune:
        # ideal-kernel tile for this lattice size; the tuning steps are warmup
        from langevina_amd.autotune import autotune_ideal
        tuned = autotune_ideal(ctx, DT)
    if args.graph:
        ctx.use_graph(True)
    ctx.run(DT, args.warmup)
    ctx.synchronize()

    ctx.timing_enable(True)
    ctx.synchronize()
    t0 = time.perf_counter()
    ctx.run(DT, args.steps)
    ctx.synchronize()
    el = time.perf_counter() - t0
    tm = ctx.timing()
    ctx.timing_enable(False)

    # the measurement pass (Measurer::measure), reported separately
    ctx.measure_full()
    tm0 = time.perf_counter()
    for _ in range(5):
        ctx.measure_full()
    meas_s = (time.perf_counter() - tm0) / 5

    sites = N ** 3
    value = sites * args.steps / el
    ms_per_step = el * 1e3 / args.steps
    # dominant kernel class: the one with the largest total time
    alg, kern, sched = tables(prec == L.FP32, N ** 3 * 10 * esz)
    dom = max(("ideal", "phi", "charge"), key=lambda k: tm[k][0])
    ms, launches = tm[dom]
    per_launch_s = ms * 1e-3 / max(launches, 1)
    bytes_per_launch = alg[dom] * sites * esz // 8
    achieved = bytes_per_launch / per_launch_s
    kernels = {k: {"kernel": kern[k], "ms_total": round(v[0], 3), "launches": v[1],
                   "avg_us": round(v[0] * 1e3 / max(v[1], 1), 2),
                   "alg_GBps": round(alg[k] * sites * esz / 8 /
                                     (v[0] * 1e-3 / max(v[1], 1)) / 1e9, 1) if v[1] else None}
               for k, v in tm.items() if k in alg}
    out = {
        "metric": "site-updates/sec (ModelG 256^3 FP64 vs FP32/batched) and % HBM roofline, 1 GPU",
        "value": value,
        "unit": "site-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if prec == L.FP64 else "f32-storage/f64-arith",
        "data": "synthetic (phi=0, Gaussian charges, zero mode removed; seed 42)",
        "config": {"workload": f"ModelG {N}^3 PV2HBSplit23 "
                               f"{'FP64' if prec == L.FP64 else 'FP32 fields'}",
                   "N": N, "dt": DT, "params": PARAMS, "parallelism": "1 GPU",
                   "graph": bool(args.graph),
                   "ideal_tile": list(tuned[N][:2]) if tuned else "default (32, 16)"},
        "step_roofline_frac": value * ALG_BYTES_PER_SITE_UPDATE * esz / 8 / HBM_PEAK,
        "step_schedule_bytes_frac": value * sched * esz / 8 / HBM_PEAK,
        "roofline": {"bound": "hbm", "kernel": kern[dom],
                     "achieved": achieved / 1e9, "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK, "traffic": traffic(kern[dom], args.precision, N),
                     "alg_bytes_per_launch": bytes_per_launch, "avg_launch_us": per_launch_s * 1e6},
        "kernels": kernels,
        "measurement": {"ms": meas_s * 1e3, "alg_GBps": N ** 3 * 10 * esz / meas_s / 1e9,
                        "what": "lgv_measure_full: walls, phase walls, averages, device r2c "
                                "of every row, rotate_to_vev; host copies included"},
    }
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline()
    print(json.dumps(out), flush=True)
    ctx.close()


if __name__ == "__main__":
    main()
