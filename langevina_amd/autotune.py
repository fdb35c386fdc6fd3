"""Autotuner of the fused ideal-step kernel's tile (x-y tile width and z-chunk
length), per lattice size, timed with HIP events on the context's stream.

The tile does not change results (every site's rotations and kick are the
same arithmetic whichever block computes them), only speed.
"""
import ctypes as C
import json
import os

from .lgv import _check, lib

CAPS = (8, 16, 32, 64)
TZS = (4, 8, 16, 32, 64)


def set_ideal_tile(ctx, replica, cap, tz):
    L = lib()
    L.lgv_set_ideal_tile.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int]
    _check(L.lgv_set_ideal_tile(ctx.h, replica, cap, tz))


def _ideal_ms(ctx, dt, steps):
    ctx.timing_enable(True)
    for _ in range(steps):
        ctx.step(dt)
    ms, n = ctx.timing()["ideal"]
    ctx.timing_enable(False)
    return ms / max(n, 1)


def autotune_ideal(ctx, dt, steps=2, cache_path=None, verbose=False):
    """Greedy per size class: for each distinct N among the context's
    replicas, try every (cap, tz) with the other classes fixed and keep the
    fastest.  Steps taken while tuning advance the replicas' state like any
    other steps.  Returns {N: (cap, tz, ms_per_ideal_launch)}."""
    cache = {}
    if cache_path and os.path.exists(cache_path):
        with open(cache_path) as f:
            cache = json.load(f)
    sizes = sorted(set(ctx.N))
    chosen = {}
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
        best = None
        for cap in CAPS:
            if cap > max(N // 2, 2):
                continue
            for tz in TZS:
                if tz > N:
                    continue
                try:
                    if len(reps) == ctx.nrep:   # one plan rebuild for all replicas
                        set_ideal_tile(ctx, -1, cap, tz)
                    else:
                        for r in reps:
                            set_ideal_tile(ctx, r, cap, tz)
                except Exception:
                    continue
                ms = _ideal_ms(ctx, dt, steps)
                if verbose:
                    print(f"autotune N={N} cap={cap} tz={tz}: {ms * 1e3:.1f} us", flush=True)
                if best is None or ms < best[2]:
                    best = (cap, tz, ms)
        if len(reps) == ctx.nrep:
            set_ideal_tile(ctx, -1, best[0], best[1])
        else:
            for r in reps:
                set_ideal_tile(ctx, r, best[0], best[1])
        chosen[N] = best
        cache[key] = list(best)
    if cache_path:
        with open(cache_path, "w") as f:
            json.dump(cache, f, indent=1)
    return chosen
