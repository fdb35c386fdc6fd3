"""Pins the oracle's steppers (ModelG/Stepper.cxx) by invariants, closed forms
and brute force, each chosen so that a plausible slip (a dropped term, a
sign, a wrong neighbour) fails it."""
import math

import numpy as np
import pytest


def random_state(orc, N, seed=0, phi_scale=1.0, charge_scale=1.0, **kw):
    lat = orc.Lattice(N, **kw)
    rng = np.random.default_rng(seed)
    lat.u[...] = rng.normal(size=lat.u.shape)
    lat.u[..., :4] *= phi_scale
    lat.u[..., 4:] *= charge_scale
    return lat


def test_kick_is_a_lattice_divergence(orc):
    # sum_x V_s is conserved and sum_x A_l changes by -dt*H*sum_x phi_{l+1}:
    # every advection term is J(x,x+e) - J(x-e,x) and telescopes
    lat = random_state(orc, 8, H=0.37)
    before = lat.u.sum(axis=(0, 1, 2))
    phisum = lat.u[..., :4].sum(axis=(0, 1, 2))
    dt = 0.013
    lat.ideal_kick(dt)
    after = lat.u.sum(axis=(0, 1, 2))
    np.testing.assert_allclose(after[7:], before[7:], atol=1e-11)
    np.testing.assert_allclose(after[4:7] - before[4:7], -dt * 0.37 * phisum[1:4], atol=1e-11)
    np.testing.assert_array_equal(after[:4], before[:4])


def test_kick_matches_numpy_restatement(orc):
    # the kick written with numpy rolls: A_l += dt (sum_d phi0 (phi_l(x+d) +
    # phi_l(x-d)) - phi_l (phi0(x+d) + phi0(x-d)) - H phi_l), V_s += dt
    # (sum_d phi_{s1}(x) (phi_{s2}(x+d) - ... ) ) with (s,s1,s2) cyclic
    lat = random_state(orc, 6, seed=5, H=0.2)
    u0 = lat.u.copy()
    dt = 0.1
    lat.ideal_kick(dt)
    f = u0[..., :4]
    A = u0[..., 4:7].copy()
    V = u0[..., 7:10].copy()
    for ax in (0, 1, 2):
        p = np.roll(f, -1, axis=ax)
        m = np.roll(f, 1, axis=ax)
        for l in range(3):
            A[..., l] += dt * (-p[..., 0] * f[..., l + 1] + f[..., 0] * p[..., l + 1]
                               + f[..., 0] * m[..., l + 1] - m[..., 0] * f[..., l + 1])
        for s in range(3):
            s1, s2 = (s + 1) % 3, (s + 2) % 3
            V[..., s] += dt * (p[..., 1 + s2] * f[..., 1 + s1] - f[..., 1 + s2] * m[..., 1 + s1]
                               - p[..., 1 + s1] * f[..., 1 + s2] + f[..., 1 + s1] * m[..., 1 + s2])
    for l in range(3):
        A[..., l] -= dt * 0.2 * f[..., l + 1]
    np.testing.assert_allclose(lat.u[..., 4:7], A, rtol=1e-12, atol=1e-12)
    np.testing.assert_allclose(lat.u[..., 7:10], V, rtol=1e-12, atol=1e-12)


def test_rotation_preserves_site_norms(orc):
    lat = random_state(orc, 4, seed=2)
    n0 = np.linalg.norm(lat.u[..., :4], axis=-1)
    lat.rotate_phi(0.3)
    np.testing.assert_allclose(np.linalg.norm(lat.u[..., :4], axis=-1), n0, rtol=1e-14)


def test_pv2_energy_error_is_third_order_per_step(orc):
    # position Verlet: one substep's energy error scales like dt^3
    errs = []
    for dt in (0.02, 0.01, 0.005):
        lat = random_state(orc, 6, seed=9, phi_scale=0.5, charge_scale=0.5, H=0.1)
        lat.ideal_step(dt)
        e0, e1 = lat.c.last_energy
        errs.append(abs(e1 - e0))
    r1, r2 = errs[0] / errs[1], errs[1] / errs[2]
    assert 6.0 < r1 < 10.5 and 6.0 < r2 < 10.5, errs


def test_energy_matches_numpy(orc):
    lat = random_state(orc, 5, seed=3, H=0.3, chi=1.7)
    u = lat.u
    f = u[..., :4]
    grad2 = sum(((np.roll(f, -1, axis=a) - f) ** 2).sum() for a in (0, 1, 2))
    e = 0.5 / 1.7 * (u[..., 4:] ** 2).sum() + 0.5 * grad2 - 0.3 * f[..., 0].sum()
    assert abs(lat.energy() - e) < 1e-11 * abs(e)


def test_charge_sweep_conserves_total_charge(orc):
    lat = random_state(orc, 8, seed=4, chi=2.0)
    before = lat.u[..., 4:].sum(axis=(0, 1, 2))
    phi = lat.u[..., :4].copy()
    for _ in range(3):
        lat.charge_step(0.04)
    np.testing.assert_allclose(lat.u[..., 4:].sum(axis=(0, 1, 2)), before, atol=1e-11)
    np.testing.assert_array_equal(lat.u[..., :4], phi)


def test_charge_pair_kernel_brute_force(orc):
    # modelg_update_charge_pair from (nA, nB): proposals q = rms*U(-sqrt3,sqrt3)
    # accepted with min(1, exp(-dS)), dS = ((nA-q)^2+(nB+q)^2-nA^2-nB^2)/(2 chi);
    # compare the accepted-q histogram with that density (the
    # q_heatbathtest.cxx:121-137 experiment, with the exact kernel)
    chi, rms = 2.0, 1.3
    for nA, nB in ((0.0, 0.0), (2.0, 1.0)):
        e = orc.Engine(seed=99)
        mon = orc.Monitor()
        qs = np.array([e.charge_pair(chi, rms, nA, nB, mon) for _ in range(200000)])
        a = math.sqrt(3) * rms
        edges = np.linspace(-a, a, 41)
        mid = 0.5 * (edges[1:] + edges[:-1])
        dS = ((nA - mid) ** 2 + (nB + mid) ** 2 - nA ** 2 - nB ** 2) / (2 * chi)
        dens = np.minimum(1, np.exp(-dS)) / (2 * a) * (edges[1] - edges[0])
        hist, _ = np.histogram(qs[qs != 0], bins=edges)
        expect = dens * len(qs)
        z = (hist - expect) / np.sqrt(expect)
        assert np.abs(z).max() < 5, (nA, nB, z)
        assert mon.down_counts + mon.up_counts == len(qs)
        assert mon.up_no == (qs == 0).sum()


def test_charge_equilibrium_variance(orc):
    # detailed balance w.r.t. exp(-n^2/(2 chi)) with fixed total charge:
    # <n^2> = chi (1 - 1/V)
    lat = orc.Lattice(6, chi=2.0, seed=11)
    lat.init_gaussian_charges()
    lat.u[..., 4:] = 0.0
    acc = []
    for s in range(400):
        lat.charge_step(0.5)
        if s >= 100:
            acc.append((lat.u[..., 4:] ** 2).mean())
    acc = np.array(acc)
    err = acc.reshape(10, -1).mean(axis=1).std() / math.sqrt(10)
    want = 2.0 * (1 - 1 / 216)
    assert abs(acc.mean() - want) < 5 * err + 1e-3, (acc.mean(), want, err)


def test_two_site_detailed_balance_enumeration(orc):
    # phi Metropolis on a 2x2x2 lattice at lambda = 0, one component moving:
    # the chain must sample exp(-S).  Check the single-site conditional:
    # mean of phi_0 at a site given frozen neighbours equals the Gaussian
    # conditional mean (heff + H)/(2 Lambda2) with variance 1/(2 Lambda2).
    m2, H = 1.5, 0.4
    lat = orc.Lattice(2, mass0=m2, lam=0.0, H=H, seed=5)
    lat.u[...] = 0.0
    lat.u[..., 0] = 0.3  # all sites of parity 1 fixed at 0.3 (they are neighbours)
    xs = []
    for _ in range(62000):
        # proposal width sqrt(2 dt) ~ conditional width: good acceptance
        lat.hb_phi_halfsweep(0.05, 0)
        xs.append(lat.u[0, 0, 0, 0])
    xs = np.array(xs[2000:])
    L2 = 0.5 * (6 + m2)
    heff = 6 * 0.3
    mean = (heff + H) / (2 * L2)
    var = 1 / (2 * L2)
    blocks = xs.reshape(30, -1)
    err = blocks.mean(axis=1).std() / math.sqrt(len(blocks))
    verr = blocks.var(axis=1).std() / math.sqrt(len(blocks))
    assert abs(xs.mean() - mean) < 5 * err, (xs.mean(), mean, err)
    assert abs(xs.var() - var) < 5 * verr, (xs.var(), var, verr)


@pytest.mark.slow
def test_free_field_wall_two_point(orc):
    # lambda = 0, m^2 > 0: <|W(k)|^2> = 1/(N^3 (m^2 + 2 - 2 cos(2 pi k/N)))
    # (freeCase.ipynb cell 7, free2pt), independent of dt
    N, m2 = 8, 1.0
    lat = orc.Lattice(N, mass0=m2, lam=0.0, H=0.0, seed=21)
    samples = []
    for s in range(6000):
        lat.hb_phi_step(0.1)  # proposal width ~ site width: fast mixing
        if s >= 600 and s % 4 == 0:
            f = lat.u[..., :4]
            ws = [f.mean(axis=(1, 2)), f.mean(axis=(0, 2)), f.mean(axis=(0, 1))]
            pk = [np.abs(np.fft.rfft(w, axis=0) / N) ** 2 for w in ws]
            samples.append(np.mean(pk, axis=(0, 2)))
    samples = np.array(samples)
    blocks = samples.reshape(27, -1, samples.shape[1]).mean(axis=1)
    mean, err = blocks.mean(axis=0), blocks.std(axis=0) / math.sqrt(len(blocks))
    for k in range(3):
        exact = 1.0 / (N ** 3 * (m2 + 2 - 2 * math.cos(2 * math.pi * k / N)))
        assert abs(mean[k] - exact) < 4 * err[k] + 0.02 * exact, (k, mean[k], exact, err[k])


def test_reference_and_per_site_modes_share_structure(orc):
    # the same driver sequence consumes counters identically in both modes
    a = orc.Lattice(4, rng_mode=orc.RNG_REFERENCE, seed=3)
    b = orc.Lattice(4, rng_mode=orc.RNG_PER_SITE, seed=3)
    init = []
    for lat in (a, b):
        lat.init_gaussian_charges()
        init.append(lat.u.copy())
        lat.step(0.04)
    # both modes draw the initial charges from the global engine
    # (ModelA.h:463-525): identical bits, zero mode removed
    np.testing.assert_array_equal(init[0], init[1])
    np.testing.assert_allclose(init[0][..., 4:].sum(axis=(0, 1, 2)), 0, atol=1e-12)
    # a step moves charge between sites only: sum V is conserved in both
    # modes (the kick is a lattice divergence, the exchange moves +-q), while
    # the trajectories themselves differ
    for lat in (a, b):
        np.testing.assert_allclose(lat.u[..., 7:].sum(axis=(0, 1, 2)),
                                   init[0][..., 7:].sum(axis=(0, 1, 2)), atol=1e-12)
    assert not np.array_equal(a.u, b.u)
    assert b.c.sweep_ctr == 22  # 3 x (1 ideal + 2 x 2 parities) + 1 shuffle + 6 charge
    assert a.c.sweep_ctr == 0


def test_fp32_mode_stores_float_values(orc):
    lat = random_state(orc, 4, seed=8, precision=orc.FP32)
    lat.round_fp32()
    for _ in range(2):
        lat.step(0.04)
    assert np.array_equal(lat.u, lat.u.astype(np.float32).astype(np.float64))


def test_threaded_per_site_mode_matches_serial(orc):
    """The OpenMP loops of the CPU baseline change only summation orders:
    fields and decision counts equal the serial per-site run."""
    runs = []
    for nt in (1, 4):
        orc.set_threads(nt)
        try:
            lat = orc.Lattice(12, seed=5, rng_mode=orc.RNG_PER_SITE)
            lat.init_gaussian_charges()
            for _ in range(2):
                lat.step(0.04)
            runs.append((lat.u.copy(), lat.monitors()))
        finally:
            orc.set_threads(1)
    (u1, m1), (u4, m4) = runs
    assert np.array_equal(u1, u4)
    for k in ("phi", "charge", "ideal"):
        for f in ("down_counts", "up_counts", "up_yes", "up_no"):
            assert m1[k][f] == m4[k][f]
        # the ideal-step dS is a difference of two volume sums of O(1e3):
        # a different summation order moves it by ~1e-12 absolute
        tol = dict(rel=0, abs=1e-9) if k == "ideal" else dict(rel=1e-12, abs=1e-12)
        for f in ("down_dS", "up_dS"):
            assert m4[k][f] == pytest.approx(m1[k][f], **tol)


@pytest.mark.parametrize("per_site", [0, 1])
def test_one_site_chain_samples_the_exact_distribution(orc, per_site):
    """One phi site in a fixed field h e_0 (the sweep's arithmetic, Stepper.cxx
    :330-383) samples exp(-S), S = -(h+H) phi_0 + Lambda2 phi^2 + lambda/4 phi^4:
    chain means against 4-D quadrature (phi_0 = r cos(theta), measure
    r^3 sin^2(theta)), for one reference-layout engine and for fresh per-site
    streams keyed like the per-site mode."""
    from scipy import integrate
    mass, lam, H, h, dt = -4.813, 4.0, 0.01, 1.8, 0.05
    L2 = 0.5 * (6.0 + mass)

    def w(th, r):
        return r ** 3 * np.sin(th) ** 2 * np.exp((h + H) * r * np.cos(th) - L2 * r * r
                                                 - lam / 4 * r ** 4)
    Z = integrate.dblquad(w, 0, 6, 0, np.pi, epsabs=1e-13, epsrel=1e-12)[0]
    m0 = integrate.dblquad(lambda th, r: w(th, r) * r * np.cos(th), 0, 6, 0, np.pi,
                           epsabs=1e-13, epsrel=1e-12)[0] / Z
    m2 = integrate.dblquad(lambda th, r: w(th, r) * r * r, 0, 6, 0, np.pi,
                           epsabs=1e-13, epsrel=1e-12)[0] / Z
    out = np.zeros(2)
    orc.lib().orc_one_site_chain(per_site, 20_000_000, h, mass, lam, H, dt, 4242, out)
    # 2e7 updates, integrated autocorrelation ~10: standard errors ~5e-4, ~3e-4
    assert abs(out[0] - m0) < 2.5e-3, (out[0], m0)
    assert abs(out[1] - m2) < 1.5e-3, (out[1], m2)


@pytest.mark.parametrize("chi", [2.0, 5.0, 3.0, 0.7, 10.0 / 3.0, 0.123456789])
def test_charge_dS_division_by_reciprocal_is_exact(orc, chi):
    """The GPU charge sweep forms dS = x / (2 chi) with a rounded reciprocal
    and one FMA correction; it must equal the IEEE quotient of the
    reference's expression (Stepper.cxx:426-428)."""
    assert orc.lib().orc_div_by_const_mismatches(chi, 2_000_000, 12345) == 0
