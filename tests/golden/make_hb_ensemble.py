"""Writes hb_ensemble.json: equilibrium observables of the phi Metropolis
sweep (EulerLangevinHB, ModelG/Stepper.cxx:287-389) from the CPU oracle in
reference RNG order (one global xoroshiro128+ engine in the reference's loop
order), for the lattice of ModelG/heatbathtest.json: N = 32, m^2 = -4.813,
lambda = 4, H = 0.01, dt = 0.05, cold start.

Each chain: 4000 thermalisation steps, then one measurement every 50 steps.
Per chain the mean and the reference's autocorrelation-corrected error
(python/Loader.py:194-210); the ensemble value is the mean over chains with
the standard error of the chain means.

    python tests/golden/make_hb_ensemble.py [--chains 14] [--meas 6000]
"""
import argparse
import json
import math
import os
import sys
from multiprocessing import Pool

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
sys.path.insert(0, os.path.dirname(HERE))

PARAMS = dict(N=32, mass0=-4.813, lam=4.0, H=0.01, chi=2.0, dt=0.05, therm=4000, every=50)
# one seed per chain (the fixture was written in two batches of seven)
SEEDS = [900001 + 17 * k for k in range(7)] + [300007 + 31 * k for k in range(7)]


def chain(args):
    seed, nmeas = args
    from oracle.oracle import RNG_REFERENCE, Lattice
    p = PARAMS
    L = Lattice(p["N"], mass0=p["mass0"], lam=p["lam"], H=p["H"], chi=p["chi"], seed=seed,
                rng_mode=RNG_REFERENCE)
    for _ in range(p["therm"]):
        L.hb_phi_step(p["dt"])
    obs = np.zeros((3, nmeas))
    for i in range(nmeas):
        for _ in range(p["every"]):
            L.hb_phi_step(p["dt"])
        o = L.measure()[2]
        obs[0, i] = o[0]
        obs[1, i] = o[10]
        obs[2, i] = math.sqrt(sum(o[a] ** 2 for a in range(4)))
    return seed, obs


def main():
    from test_gpu_ensemble import autocorrelated_err
    ap = argparse.ArgumentParser()
    ap.add_argument("--chains", type=int, default=14)
    ap.add_argument("--meas", type=int, default=6000)
    args = ap.parse_args()
    seeds = SEEDS[:args.chains]
    with Pool(min(args.chains, os.cpu_count() or 1)) as pool:
        res = pool.map(chain, [(s, args.meas) for s in seeds])
    out = {"params": PARAMS, "rng": "reference order", "chains": args.chains,
           "measurements_per_chain": args.meas, "seeds": seeds}
    for k, name in enumerate(("phi0", "phi2", "absphi")):
        per = [autocorrelated_err(obs[k]) for _, obs in res]
        means = np.array([m for m, _, _ in per])
        out[name] = {"mean": float(means.mean()),
                     "err": float(means.std(ddof=1) / math.sqrt(len(means))),
                     "chain_means": means.tolist(),
                     "chain_errs": [float(e) for _, e, _ in per],
                     "chain_tau": [float(t) for _, _, t in per]}
    with open(os.path.join(HERE, "hb_ensemble.json"), "w") as f:
        json.dump(out, f, indent=1)
    print(json.dumps({k: out[k]["mean"] for k in ("phi0", "phi2", "absphi")}))


if __name__ == "__main__":
    main()
