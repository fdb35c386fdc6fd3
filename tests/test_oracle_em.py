"""Pins the oracle's Model A Euler--Maruyama stepper and its adjoint.

* free field (lambda = H = 0, m^2 > 0): the stationary wall two-point
  function of forward Euler is 2 / (w_k^2 (2 - G dt w_k^2)) / N^3 with
  w_k^2 = m^2 + 2 - 2 cos(2 pi k / N)  (ModelG/freeCase.ipynb cells 5, 7,
  free2ptFE) -- dt-dependent, unlike the Metropolis closed form;
* the reverse recurrence against central finite differences with the same
  noise, for the three couplings and a random direction in phi_0.
"""
import math

import numpy as np
import pytest


def test_noise_moments(orc):
    for kind in (0, 1):
        p = orc.em_params(16, 1.0, 0.0, 0.0, 1.0, 0.1, 3, kind)
        xi = np.concatenate([orc.em_noise(p, s).ravel() for s in range(40)])
        n = len(xi)
        var_of_var = (0.8 if kind == 0 else 2.0) / n  # (E x^4 - 1)/n
        assert abs(xi.mean()) < 5 / math.sqrt(n)
        assert abs(xi.var() - 1) < 5 * math.sqrt(var_of_var)
    # uniform kind is the reference's variance1 bounded support
    p = orc.em_params(8, 1.0, 0.0, 0.0, 1.0, 0.1, 3, 0)
    assert np.abs(orc.em_noise(p, 0)).max() < 3 ** 0.5


def test_em_step_matches_numpy(orc):
    N = 6
    p = orc.em_params(N, -0.7, 0.9, 0.2, 1.3, 0.03, 11, 0)
    phi = np.random.default_rng(0).normal(size=(N, N, N, 4))
    out = orc.em_step(p, 5, phi)
    xi = orc.em_noise(p, 5)
    lap = sum(np.roll(phi, s, axis=a) for a in (0, 1, 2) for s in (1, -1)) - 6 * phi
    s2 = (phi ** 2).sum(axis=-1, keepdims=True)
    drift = -lap + p.m2 * phi + p.lam * s2 * phi
    drift[..., 0] -= p.H
    want = phi - p.gamma * p.dt * drift + math.sqrt(2 * p.gamma * p.dt) * xi
    np.testing.assert_allclose(out, want, rtol=1e-12, atol=1e-12)


@pytest.mark.slow
def test_free_forward_euler_two_point(orc):
    # forward Euler is stable for G dt w^2 < 2, w^2 <= m^2 + 12
    N, m2, dt = 8, 1.0, 0.1
    p = orc.em_params(N, m2, 0.0, 0.0, 1.0, dt, 5, 1)
    phi = np.zeros((N, N, N, 4))
    acc = []
    for s in range(4000):
        phi = orc.em_step(p, s, phi)
        if s >= 200 and s % 2 == 0:
            ws = [phi.mean(axis=(1, 2)), phi.mean(axis=(0, 2)), phi.mean(axis=(0, 1))]
            pk = [np.abs(np.fft.rfft(w, axis=0) / N) ** 2 for w in ws]
            acc.append(np.mean(pk, axis=(0, 2)))
    acc = np.array(acc)
    blocks = acc.reshape(20, -1, acc.shape[1]).mean(axis=1)
    mean, err = blocks.mean(axis=0), blocks.std(axis=0) / math.sqrt(20)
    for k in range(3):
        w2 = m2 + 2 - 2 * math.cos(2 * math.pi * k / N)
        exact = 2 / (w2 * (2 - dt * w2)) / N ** 3
        assert abs(mean[k] - exact) < 4 * err[k] + 0.01 * exact, (k, mean[k], exact, err[k])


def test_adjoint_against_finite_differences(orc):
    N, nsteps, window = 6, 12, 6
    base = dict(m2=-0.5, lam=0.8, H=0.1)
    phi0 = np.random.default_rng(2).normal(size=(N, N, N, 4)) * 0.4

    def params(**d):
        q = dict(base)
        q.update(d)
        return orc.em_params(N, q["m2"], q["lam"], q["H"], 1.0, 0.02, 7, 0)

    def loss(pp, ph):
        return orc.em_run(pp, ph, nsteps, window)[0]

    g, gp = orc.em_adjoint(params(), phi0, nsteps, window)
    eps = 1e-6
    for i, name in enumerate(("m2", "H", "lam")):
        fd = (loss(params(**{name: base[name] + eps}), phi0) -
              loss(params(**{name: base[name] - eps}), phi0)) / (2 * eps)
        assert abs(gp[i] - fd) <= 1e-6 * abs(fd), (name, gp[i], fd)
    v = np.random.default_rng(5).normal(size=phi0.shape)
    fd = (loss(params(), phi0 + eps * v) - loss(params(), phi0 - eps * v)) / (2 * eps)
    assert abs((g * v).sum() - fd) <= 1e-6 * abs(fd)


def test_em_series_matches_numpy_means(orc):
    # measure-then-step: row r holds the state after r*every steps
    N, nsteps, every = 6, 23, 5
    p = orc.em_params(N, -0.6, 0.9, 0.05, 1.0, 0.03, 9, 1)
    phi = np.random.default_rng(4).normal(size=(N, N, N, 4)) * 0.2
    rows, fin = orc.em_series(p, phi, nsteps, every)
    assert rows.shape == (5, 6)
    V = N ** 3
    for n in range(nsteps):
        if n % every == 0:
            want = list(phi.reshape(-1, 4).mean(axis=0))
            want.append((phi ** 2).sum(axis=-1).mean())
            want.append(V * sum(w * w for w in want[:4]))
            np.testing.assert_allclose(rows[n // every], want, rtol=1e-13, atol=1e-16)
        phi = orc.em_step(p, n, phi)
    np.testing.assert_array_equal(fin, phi)
    phi0 = np.random.default_rng(4).normal(size=(N, N, N, 4)) * 0.2
    np.testing.assert_array_equal(fin, orc.em_run(p, phi0, nsteps, 1)[2])


def test_em_series_free_field_site_variance(orc):
    # free field: every Fourier mode of forward Euler has the stationary
    # variance 2 / (w_k^2 (2 - G dt w_k^2)), so the mean of |phi|^2 over the
    # lattice is 4/V sum_k of it (four components) and the k = 0 mode gives
    # V <phibar_a^2> = 2 / (m^2 (2 - G dt m^2))  (freeCase.ipynb free2ptFE)
    N, m2, dt = 6, 1.0, 0.1
    p = orc.em_params(N, m2, 0.0, 0.0, 1.0, dt, 13, 1)
    rows, _ = orc.em_series(p, np.zeros((N, N, N, 4)), 30000, 3)
    rows = rows[100:]
    k = 2 * np.pi * np.arange(N) / N
    w2 = m2 + sum(np.meshgrid(*(2 - 2 * np.cos(k),) * 3, indexing="ij"))
    site = 4 * (2 / (w2 * (2 - dt * w2))).mean()
    blocks = rows[:, 4].reshape(20, -1).mean(axis=1)
    err = blocks.std() / math.sqrt(20)
    assert abs(blocks.mean() - site) < 4 * err + 1e-3 * site, (blocks.mean(), site, err)
    V = N ** 3
    zero = 2 / (m2 * (2 - dt * m2))
    b0 = (V * rows[:, :4] ** 2).mean(axis=1).reshape(20, -1).mean(axis=1)
    e0 = b0.std() / math.sqrt(20)
    assert abs(b0.mean() - zero) < 4 * e0 + 0.01 * zero, (b0.mean(), zero, e0)
