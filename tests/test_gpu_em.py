"""GPU tests of the Model A Euler--Maruyama stepper and its adjoint.

Forward: the device trajectory equals the oracle's bit for bit with the
uniform (variance1) noise, and to 1e-12 with Box-Muller noise (device
log/sin/cos); the loss agrees to 1e-12.  Backward: the hand-written adjoint
agrees with the oracle's reverse recurrence to 1e-10 and with same-noise
central finite differences (eps = 1e-6) to 1e-5 relative, including the
128^3, 500-step, checkpoint-every-50 configuration.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def log(*a):
    print(*a, flush=True)


@pytest.fixture(scope="module")
def torch():
    import torch
    assert torch.cuda.is_available()
    return torch


def soa(phi_aos):
    return np.ascontiguousarray(np.moveaxis(phi_aos, -1, 0))


def aos(phi_soa):
    return np.ascontiguousarray(np.moveaxis(phi_soa, 0, -1))


@pytest.mark.parametrize("kind", [0, 1])
def test_em_forward_matches_oracle(torch, orc, kind):
    log("em forward", kind)
    from langevina_amd.em import EulerMaruyama
    N, nsteps, window = 8, 20, 10
    par = dict(m2=-0.7, lam=1.0, H=0.05, gamma=1.0, dt=0.02)
    phi0 = np.random.default_rng(1).normal(size=(N, N, N, 4)) * 0.3
    p = orc.em_params(N, par["m2"], par["lam"], par["H"], par["gamma"], par["dt"], 42, kind)
    loss_o, phibar_o, fin_o = orc.em_run(p, phi0, nsteps, window)
    h = EulerMaruyama(N, par["m2"], par["lam"], par["H"], par["gamma"], par["dt"], 42,
                      nsteps, 7, noise_kind=kind)
    d0 = torch.tensor(soa(phi0), device="cuda")
    fin = torch.empty_like(d0)
    loss, phibar = h.forward(d0.data_ptr(), window, fin.data_ptr())
    torch.cuda.synchronize()
    got = aos(fin.cpu().numpy())
    if kind == 0:
        np.testing.assert_array_equal(got, fin_o)
    else:
        assert np.abs(got - fin_o).max() <= 1e-12 * np.abs(fin_o).max()
    np.testing.assert_allclose(phibar, phibar_o, rtol=1e-12, atol=1e-15)
    assert abs(loss - loss_o) <= 1e-12 * abs(loss_o)
    h.close()


def test_em_adjoint_matches_oracle(torch, orc):
    log("em adjoint vs oracle")
    from langevina_amd.em import EulerMaruyama
    N, nsteps, window = 6, 12, 6
    par = dict(m2=-0.5, lam=0.8, H=0.1, gamma=1.0, dt=0.02)
    phi0 = np.random.default_rng(2).normal(size=(N, N, N, 4)) * 0.4
    p = orc.em_params(N, par["m2"], par["lam"], par["H"], par["gamma"], par["dt"], 7, 0)
    g_o, gp_o = orc.em_adjoint(p, phi0, nsteps, window)
    for ckpt in (1, 5, 12):
        h = EulerMaruyama(N, par["m2"], par["lam"], par["H"], par["gamma"], par["dt"], 7,
                          nsteps, ckpt)
        d0 = torch.tensor(soa(phi0), device="cuda")
        h.forward(d0.data_ptr(), window)
        g = torch.empty_like(d0)
        gp = h.backward(g.data_ptr())
        torch.cuda.synchronize()
        np.testing.assert_allclose(aos(g.cpu().numpy()), g_o, rtol=1e-10, atol=1e-14)
        np.testing.assert_allclose(gp, gp_o, rtol=1e-10)
        h.close()


def _fd_check(torch, N, nsteps, window, ckpt, dt, par, seed, nprobe=3):
    from langevina_amd.autograd import modela_em_loss
    rng = np.random.default_rng(seed)
    phi0 = torch.tensor(rng.normal(size=(4, N, N, N)) * 0.3, device="cuda")
    cfg = dict(dt=dt, gamma=1.0, nsteps=nsteps, window=window, ckpt_every=ckpt, seed=seed)

    def loss_at(m2, H, lam, ph):
        return modela_em_loss(torch.tensor(m2, dtype=torch.float64),
                              torch.tensor(H, dtype=torch.float64),
                              torch.tensor(lam, dtype=torch.float64), ph, **cfg).item()

    m2 = torch.tensor(par["m2"], dtype=torch.float64, requires_grad=True)
    H = torch.tensor(par["H"], dtype=torch.float64, requires_grad=True)
    lam = torch.tensor(par["lam"], dtype=torch.float64, requires_grad=True)
    ph = phi0.clone().requires_grad_(True)
    loss = modela_em_loss(m2, H, lam, ph, **cfg)
    loss.backward()
    eps = 1e-6
    base = dict(par)
    for name, t in (("m2", m2), ("H", H), ("lam", lam)):
        up, dn = dict(base), dict(base)
        up[name] += eps
        dn[name] -= eps
        fd = (loss_at(up["m2"], up["H"], up["lam"], phi0) -
              loss_at(dn["m2"], dn["H"], dn["lam"], phi0)) / (2 * eps)
        ad = t.grad.item()
        log(name, ad, fd)
        assert abs(ad - fd) <= 1e-5 * max(abs(fd), 1e-8), (name, ad, fd)
    # directional derivative along random directions in phi0
    for k in range(nprobe):
        v = torch.tensor(np.random.default_rng(100 + k).normal(size=(4, N, N, N)), device="cuda")
        fd = (loss_at(base["m2"], base["H"], base["lam"], phi0 + eps * v) -
              loss_at(base["m2"], base["H"], base["lam"], phi0 - eps * v)) / (2 * eps)
        ad = float((ph.grad * v).sum())
        log("dir", k, ad, fd)
        assert abs(ad - fd) <= 1e-5 * max(abs(fd), 1e-8), (k, ad, fd)


def test_autograd_against_finite_differences_small(torch):
    log("em autograd FD small")
    _fd_check(torch, 8, 40, 20, 10, 0.02, dict(m2=-0.6, H=0.05, lam=1.0), 3)


def test_autograd_128_cubed_500_steps(torch):
    # the configuration of the differentiable Model A benchmark: 128^3 FP64,
    # dt = 0.02, 500 steps, loss over the last 250, checkpoints every 50
    log("em autograd FD 128^3")
    _fd_check(torch, 128, 500, 250, 50, 0.02, dict(m2=-4.813, H=0.01, lam=1.0), 42, nprobe=1)


C0 = dict(m2=-4.813, lam=1.0, H=0.01, gamma=1.0, dt=0.04)


@pytest.mark.parametrize("kind", [0, 1])
def test_em_series_16_cubed_1000_steps(torch, orc, kind):
    # the plain Model A configuration: 16^3, phi_0 ~ N(0, 0.1^2), seed 42,
    # 1000 steps, the observables measured every 10 steps (before stepping)
    log("em series 16^3 x 1000, noise kind", kind)
    from langevina_amd.em import EulerMaruyama, series_susceptibility
    N, nsteps, every = 16, 1000, 10
    phi0 = np.random.default_rng(42).normal(size=(N, N, N, 4)) * 0.1
    p = orc.em_params(N, C0["m2"], C0["lam"], C0["H"], C0["gamma"], C0["dt"], 42, kind)
    rows_o, fin_o = orc.em_series(p, phi0, nsteps, every)
    h = EulerMaruyama(N, C0["m2"], C0["lam"], C0["H"], C0["gamma"], C0["dt"], 42, nsteps, 50,
                      noise_kind=kind)
    d0 = torch.tensor(soa(phi0), device="cuda")
    fin = torch.empty_like(d0)
    rows = h.series(d0.data_ptr(), every, fin.data_ptr())
    got = aos(fin.cpu().numpy())
    assert rows.shape == (100, 6)
    if kind == 0:
        np.testing.assert_array_equal(got, fin_o)
    else:
        err = np.abs(got - fin_o).max() / np.abs(fin_o).max()
        log("final-state deviation", err)
        assert err <= 1e-12
    np.testing.assert_allclose(rows, rows_o, rtol=1e-12, atol=1e-15)
    chi, chi_o = series_susceptibility(rows, N ** 3), series_susceptibility(rows_o, N ** 3)
    log("susceptibility", chi, chi_o)
    assert abs(chi - chi_o) <= 1e-10 * abs(chi_o)
    # the series leaves the forward state alone: forward still matches
    loss, phibar = h.forward(d0.data_ptr(), 100)
    loss_o, phibar_o, _ = orc.em_run(p, phi0, nsteps, 100)
    np.testing.assert_allclose(phibar, phibar_o, rtol=1e-12, atol=1e-15)
    h.close()


def test_em_step_chain_matches_series(torch, orc):
    log("em single steps vs series")
    from langevina_amd.em import EulerMaruyama
    N, nsteps = 12, 9
    phi0 = np.random.default_rng(3).normal(size=(N, N, N, 4)) * 0.3
    h = EulerMaruyama(N, -0.8, 1.0, 0.02, 1.0, 0.03, 5, nsteps, 3, noise_kind=1)
    d0 = torch.tensor(soa(phi0), device="cuda")
    fin = torch.empty_like(d0)
    rows = h.series(d0.data_ptr(), 1, fin.data_ptr())
    a, b = d0.clone(), torch.empty_like(d0)
    obs = torch.zeros(nsteps, 6, dtype=torch.float64, device="cuda")
    for n in range(nsteps):
        h.step(a.data_ptr(), b.data_ptr(), n, obs[n].data_ptr() if n + 1 < nsteps else None)
        a, b = b, a
    torch.cuda.synchronize()
    assert torch.equal(a, fin)
    np.testing.assert_array_equal(obs[:-1].cpu().numpy(), rows[1:])
    h.close()


def test_em_free_field_closed_form(torch):
    # lambda = H = 0, m^2 = 1, Gaussian noise: forward Euler's stationary
    # variance per Fourier mode is 2 / (w_k^2 (2 - G dt w_k^2))
    # (ModelG/freeCase.ipynb free2ptFE): the k = 0 mode gives
    # V <phibar_a^2>, the sum over all k the mean of |phi|^2
    log("em free field closed form")
    import math
    from langevina_amd.em import EulerMaruyama
    N, m2, dt, nsteps, every = 16, 1.0, 0.1, 200000, 5
    h = EulerMaruyama(N, m2, 0.0, 0.0, 1.0, dt, 17, nsteps, 1000, noise_kind=1)
    d0 = torch.zeros(4, N, N, N, dtype=torch.float64, device="cuda")
    rows = h.series(d0.data_ptr(), every)[200:]
    h.close()
    V = N ** 3
    k = 2 * np.pi * np.arange(N) / N
    w2 = m2 + sum(np.meshgrid(*(2 - 2 * np.cos(k),) * 3, indexing="ij"))
    site = 4 * (2 / (w2 * (2 - dt * w2))).mean()
    nb = 40
    b = rows[: len(rows) // nb * nb, 4].reshape(nb, -1).mean(axis=1)
    err = b.std() / math.sqrt(nb)
    log("<|phi|^2>", b.mean(), "+-", err, "exact", site)
    assert abs(b.mean() - site) < 4 * err
    zero = 2 / (m2 * (2 - dt * m2))
    b0 = (V * rows[: len(rows) // nb * nb, :4] ** 2).mean(axis=1).reshape(nb, -1).mean(axis=1)
    e0 = b0.std() / math.sqrt(nb)
    log("V <phibar_a^2>", b0.mean(), "+-", e0, "exact", zero)
    assert abs(b0.mean() - zero) < 4 * e0
