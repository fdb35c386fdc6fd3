"""Autotuner of the kernels' tiles, per lattice size, timed with HIP events
on the context's stream: the fused ideal-step kernel (x-y tile width and
z-chunk length), the charge march (tile width and chunk length along the
sweep direction) and the phi z-march (tile width and z-chunk length).

A tile does not change results (every site's arithmetic and random stream
are the same whichever block computes it), only speed.
"""
import ctypes as C
import json
import os
import time

from .lgv import _check, lib

CAPS = (8, 16, 32, 64)
# short z chunks give a small lattice (one 64^3 replica: 32 blocks with the
# default tile) more blocks than the march's depth costs it
TZS = (2, 4, 8, 16, 32, 64)
CHARGE_CAPS = (16, 32, 64)
CHARGE_CHUNKS = (2, 4, 8, 16, 32)
HB_CAPS = (16, 32, 64)
HB_TZS = (8, 16, 32)


def set_ideal_tile(ctx, replica, cap, tz):
    L = lib()
    L.lgv_set_ideal_tile.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int]
    _check(L.lgv_set_ideal_tile(ctx.h, replica, cap, tz))


def charge_eff(N, cap, chunk):
    """The charge tile (TXH, C) a (cap, chunk) request gives (0 = default)."""
    c = cap or 64
    ch = chunk or (16 if N >= 128 else 8)
    return min(c, N // 2), min(ch, N)


def hb_eff(N, cap, tz):
    """The phi-march tile (TXH, TZ) a (cap, tz) request gives (0 = default)."""
    return min(cap or 64, N // 2), min(tz or 16, N)


def set_sweep_tiles(ctx, replica, hb_cap, hb_tz, chg_cap, chg_chunk):
    L = lib()
    L.lgv_set_sweep_tiles.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int]
    _check(L.lgv_set_sweep_tiles(ctx.h, replica, hb_cap, hb_tz, chg_cap, chg_chunk))


# Every candidate is timed in two rounds (the minimum counts), after a few
# untimed warm-up steps, and replaces the incumbent (the defaults first)
# only when faster by more than this fraction: single three-step timings of
# one launch class vary by a few per cent, and tiles picked on such noise
# measured up to 5 % slower in whole steps than the defaults.
MARGIN = 0.01
ROUNDS = 2
WARMUP = 3


def _class_ms(ctx, dt, steps, cls):
    """Mean launch time of one kernel class (HIP events), or (cls None) the
    mean step time of an untimed run on the host clock, with the sweeps in
    the order (sequential or concurrent) the library runs them."""
    ctx.step(dt)  # the first launches after a retile are not timed
    if cls is None:
        ctx.synchronize()
        t0 = time.perf_counter()
        ctx.run(dt, steps)
        ctx.synchronize()
        return (time.perf_counter() - t0) * 1e3 / steps
    ctx.timing_enable(True)
    for _ in range(steps):
        ctx.step(dt)
    tm = ctx.timing()
    ctx.timing_enable(False)
    ms, n = tm[cls]
    return ms / max(n, 1)


def _search(ctx, dt, steps, cls, apply, candidates):
    """Times every candidate (the first is the incumbent) in ROUNDS rounds
    and returns the fastest, keeping the incumbent unless a candidate beats
    it by more than MARGIN."""
    best_ms = {}
    for _ in range(ROUNDS):
        for cand in candidates:
            try:
                apply(cand)
            except Exception:
                continue
            ms = _class_ms(ctx, dt, steps, cls)
            best_ms[cand] = min(ms, best_ms.get(cand, ms))
    inc = candidates[0]
    best = inc
    for cand, ms in best_ms.items():
        if ms < best_ms[best] and ms < best_ms[inc] * (1 - MARGIN):
            best = cand
    apply(best)
    return best, best_ms[best]


def autotune_sweeps(ctx, dt, march=None, steps=3, verbose=False):
    """march None: first the phi half-sweep kernel, the per-site kernel or
    the z-march, timed over whole steps (the march's lower traffic also
    speeds up the kernels that follow it), the size rule's choice being the
    incumbent; march True / False: keep the kernel in use and tune (or not)
    the march's tiles.  Then per size class, as autotune_ideal: the charge
    march (tile width x chunk) and, when the phi half-sweep runs the z-march,
    its tile width x z-chunk.  Returns {N: (hb_cap, hb_tz, chg_cap,
    chg_chunk)} (0 = default); ctx.phi_kernel() reports the kernel."""
    sizes = sorted(set(ctx.N))
    chosen = {N: [0, 0, 0, 0] for N in sizes}

    def apply(N, t):
        reps = [r for r, n in enumerate(ctx.N) if n == N]
        if len(reps) == ctx.nrep:
            set_sweep_tiles(ctx, -1, *t)
        else:
            for r in reps:
                set_sweep_tiles(ctx, r, *t)

    for _ in range(WARMUP):
        ctx.step(dt)
    if march is None:
        from .lgv import PHI_AUTO, PHI_MARCH, PHI_SITE
        ctx.set_phi_kernel(PHI_AUTO)
        inc = ctx.phi_kernel()
        kernel, ms = _search(ctx, dt, 8 * steps, None, ctx.set_phi_kernel,
                             [inc, PHI_SITE if inc == PHI_MARCH else PHI_MARCH])
        march = kernel == PHI_MARCH
        if verbose:
            print(f"autotune phi kernel: {'march' if march else 'site'} {ms:.3f} ms/step",
                  flush=True)
    for N in sizes:
        cands, seen = [(0, 0)], {charge_eff(N, 0, 0)}
        for cap in CHARGE_CAPS:
            for ch in CHARGE_CHUNKS:
                if ch <= N and charge_eff(N, cap, ch) not in seen:
                    seen.add(charge_eff(N, cap, ch))
                    cands.append((cap, ch))
        (cap, ch), ms = _search(ctx, dt, steps, "charge",
                                lambda c: apply(N, chosen[N][:2] + list(c)), cands)
        chosen[N][2:] = [cap, ch]
        if verbose:
            print(f"autotune charge N={N}: cap={cap} chunk={ch} {ms * 1e3:.1f} us", flush=True)
        if not march:
            continue
        cands, seen = [(0, 0)], {hb_eff(N, 0, 0)}
        for cap in HB_CAPS:
            for tz in HB_TZS:
                if hb_eff(N, cap, tz) not in seen:
                    seen.add(hb_eff(N, cap, tz))
                    cands.append((cap, tz))
        (cap, tz), ms = _search(ctx, dt, steps, "phi",
                                lambda c: apply(N, list(c) + chosen[N][2:]), cands)
        chosen[N][:2] = [cap, tz]
        if verbose:
            print(f"autotune phi N={N}: cap={cap} tz={tz} {ms * 1e3:.1f} us", flush=True)
    return {N: tuple(v) for N, v in chosen.items()}


def autotune_ideal(ctx, dt, steps=3, cache_path=None, verbose=False):
    """Greedy per size class: for each distinct N among the context's
    replicas, try every (cap, tz) with the other classes fixed and keep the
    fastest, starting from the default (32, 16).  Steps taken while tuning
    advance the replicas' state like any other steps.
    Returns {N: (cap, tz, ms_per_ideal_launch)}."""
    cache = {}
    if cache_path and os.path.exists(cache_path):
        with open(cache_path) as f:
            cache = json.load(f)
    sizes = sorted(set(ctx.N))
    chosen = {}
    warmed = False
    for N in sizes:
        key = f"N{N}-p{ctx.precision}-r{ctx.nrep}"
        reps = [r for r, n in enumerate(ctx.N) if n == N]
        if key in cache:
            cap, tz = cache[key][:2]
            if len(reps) == ctx.nrep:
                set_ideal_tile(ctx, -1, cap, tz)
            else:
                for r in reps:
                    set_ideal_tile(ctx, r, cap, tz)
            chosen[N] = tuple(cache[key])
            continue

        def apply(c):
            if len(reps) == ctx.nrep:   # one plan rebuild for all replicas
                set_ideal_tile(ctx, -1, *c)
            else:
                for r in reps:
                    set_ideal_tile(ctx, r, *c)

        if not warmed:
            for _ in range(WARMUP):
                ctx.step(dt)
            warmed = True
        cands = [(32, 16)] + [(cap, tz) for cap in CAPS for tz in TZS
                              if cap <= max(N // 2, 2) and tz <= N and (cap, tz) != (32, 16)]
        (cap, tz), ms = _search(ctx, dt, steps, "ideal", apply, cands)
        if verbose:
            print(f"autotune ideal N={N}: cap={cap} tz={tz} {ms * 1e3:.1f} us", flush=True)
        best = (cap, tz, ms)
        chosen[N] = best
        cache[key] = list(best)
    if cache_path:
        with open(cache_path, "w") as f:
            json.dump(cache, f, indent=1)
    return chosen
