"""Ensemble parity of the phi Metropolis sweep (EulerLangevinHB,
ModelG/Stepper.cxx:287-389): the GPU path samples the same distribution as
the reference algorithm.

Lattice of ModelG/heatbathtest.json:1-14 (which ModelG/heatbathplot.py:6-17
compares with the stored heat-bath result): N = 32, m^2 = -4.813, H = 0.01,
lambda = 4 in the potential lambda/4 (phi^2)^2 of Stepper.cxx:355, dt = 0.05,
cold start; the GPU replicas measure once per 100 sweeps.

* Parity bar: 48 independent GPU replicas (per-site counter streams)
  against the CPU oracle run in the reference's own RNG order
  (tests/golden/hb_ensemble.json, written by make_hb_ensemble.py): <phi_0>,
  <phi^2> and <|phi_bar|> agree within 2 combined standard errors, each
  error from the spread of independent chains.
* The stored reference values (ModelG/hcrit1_N032_m-04813_h010000_l10000_
  averages_ave.out:1): <phi_0> = 0.29771577381 +- 0.00033803, <phi^2> and
  <|phi_bar|> are printed with their distance in standard errors and held
  to 1 % relative; phi_1..3, zero by the O(3) symmetry left by H along
  phi_0, are held within 4 standard errors of the stored values (combined)
  and of zero.  The oracle
  in reference order (14 chains) and the GPU both sit about 3 of its
  standard errors below it: the run that produced it is not in the
  reference tree (its tag reads lambda = 1.0000, a normalisation the current
  code does not use; at lambda = 1 this lattice sits at <phi_0> ~ 1.9, deep
  in the broken phase), so it cannot be repeated exactly.
"""
import json
import math
import os

import numpy as np
import pytest

REF_MEAN, REF_ERR = 0.29771577381, 3.3803070885e-04
# The other columns of that file (mean, error, integrated autocorrelation
# time / 100 per observable): phi_1..3, then 0.99771 +- 0.00005, which is
# the volume average of phi^2, and 0.30236 +- 0.00030, which is |phi_bar|
# (the sixth, 1.2372, matches no quantity the measurer stores).
REF_COLS = {"phi_1": (-6.3950531535e-04, 1.2292241007e-03),
            "phi_2": (-3.5453750383e-04, 1.2086034345e-03),
            "phi_3": (-8.6118122099e-04, 1.1829254410e-03),
            "phi2": (9.9771450882e-01, 4.8322472767e-05),
            "absphi": (3.0235983190e-01, 2.9856553209e-04)}


def autocorrelation(data, tmax):
    """autocorrelationFunction of python/Loader.py:170-191."""
    data = np.asarray(data, dtype=float)
    mean = data.mean()
    d = data - mean
    norm = np.sum(d * d)
    return np.array([np.sum(d[: len(d) - t] * d[t:]) / norm for t in range(tmax)])


def autocorrelated_err(arr, tmax=1000):
    """autocorrelatedErr / intAutocorrrelationTime of python/Loader.py:194-210:
    mean, sqrt(2 tau) * SEM, tau with tau = 1/2 + sum C(t) up to 5 tau."""
    arr = np.asarray(arr, dtype=float)
    ct = autocorrelation(arr, min(tmax, len(arr)))
    res, count = 0.5, 0
    while count < 5 * res and count < len(ct) / 2.0:
        res += ct[count]
        count += 1
    sem = arr.std(ddof=1) / math.sqrt(len(arr))
    return arr.mean(), math.sqrt(2.0 * res) * sem, res


@pytest.mark.parametrize("a", [0.5, 0.8, 0.95])
def test_autocorrelated_err_on_ar1(a):
    """AR(1) x_i = a x_{i-1} + xi (the process of python/testmeans.py:6-32):
    C(t) = a^t, so the reference's sum, which starts at 1/2 and includes
    C(0) = 1, converges to (1+a)/(2(1-a)) + 1; the error it reports is then
    sqrt(2 tau) * SEM."""
    rng = np.random.default_rng(0)
    n = 40000
    xi = rng.normal(size=n)
    x = np.zeros(n)
    for i in range(1, n):
        x[i] = a * x[i - 1] + xi[i]
    mean, err, tau = autocorrelated_err(x)
    tau_exact = (1 + a) / (2 * (1 - a)) + 1.0
    assert abs(tau - tau_exact) < 0.15 * tau_exact
    sem = x.std(ddof=1) / math.sqrt(n)
    assert err == pytest.approx(math.sqrt(2 * tau) * sem)
    assert abs(mean) < 4 * err


GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "hb_ensemble.json")


def replica_ensemble(R=48, therm=4000, nmeas=2000, every=100):
    """Per-replica series of phi_0, phi^2, |phi_bar|, phi_1, phi_2, phi_3."""
    import langevina_amd as L
    g = json.load(open(GOLDEN))["params"]
    reps = [L.model_params(g["N"], mass0=g["mass0"], lam=g["lam"], H=g["H"], chi=g["chi"],
                           seed=7000 + r) for r in range(R)]
    ctx = L.Context(reps)
    for _ in range(therm):
        ctx.hb_phi_step(g["dt"])
    obs = np.zeros((6, R, nmeas))
    for i in range(nmeas):
        for _ in range(every):
            ctx.hb_phi_step(g["dt"])
        for r in range(R):
            o = ctx.measure(r)[2]
            obs[0, r, i] = o[0]
            obs[1, r, i] = o[10]
            obs[2, r, i] = math.sqrt(float(np.sum(o[:4] ** 2)))
            obs[3:6, r, i] = o[1:4]
    ctx.close()
    return obs


@pytest.mark.gpu
def test_heatbath_ensemble_matches_oracle_and_reference(capsys):
    gold = json.load(open(GOLDEN))
    obs = replica_ensemble()
    lines = []
    for k, name in enumerate(("phi0", "phi2", "absphi")):
        means = np.array([autocorrelated_err(obs[k, r])[0] for r in range(obs.shape[1])])
        mean, err = means.mean(), means.std(ddof=1) / math.sqrt(len(means))
        o = gold[name]
        sig = math.hypot(err, o["err"])
        lines.append(f"{name}: GPU {mean:.6f} +- {err:.6f}, oracle {o['mean']:.6f} +- "
                     f"{o['err']:.6f} ({(mean - o['mean']) / sig:+.2f} sigma)")
        assert abs(mean - o["mean"]) < 2 * sig, lines[-1]
        if name == "phi0":
            ref_sig = math.hypot(err, REF_ERR)
            lines.append(f"phi0 vs stored reference {REF_MEAN:.6f} +- {REF_ERR:.6f}: "
                         f"{(mean - REF_MEAN) / ref_sig:+.2f} sigma, "
                         f"{(mean - REF_MEAN) / REF_MEAN:+.3%}")
            assert abs(mean - REF_MEAN) < 0.01 * REF_MEAN, lines[-1]
    for k, name in ((1, "phi2"), (2, "absphi")):
        means = np.array([autocorrelated_err(obs[k, r])[0] for r in range(obs.shape[1])])
        mean, err = means.mean(), means.std(ddof=1) / math.sqrt(len(means))
        rm, re = REF_COLS[name]
        lines.append(f"{name} vs stored reference {rm:.6f} +- {re:.6f}: "
                     f"{(mean - rm) / math.hypot(err, re):+.2f} sigma, {(mean - rm) / rm:+.3%}")
        assert abs(mean - rm) < 0.01 * rm, lines[-1]
    for k, name in ((3, "phi_1"), (4, "phi_2"), (5, "phi_3")):
        means = np.array([autocorrelated_err(obs[k, r])[0] for r in range(obs.shape[1])])
        mean, err = means.mean(), means.std(ddof=1) / math.sqrt(len(means))
        rm, re = REF_COLS[name]
        sig = math.hypot(err, re)
        lines.append(f"{name}: GPU {mean:+.6f} +- {err:.6f}, "
                     f"stored {rm:+.6f} +- {re:.6f} ({(mean - rm) / sig:+.2f} sigma)")
        assert abs(mean - rm) < 4 * sig, lines[-1]
        assert abs(mean) < 4 * err, lines[-1]
    with capsys.disabled():
        print("\n" + "\n".join(lines), flush=True)
