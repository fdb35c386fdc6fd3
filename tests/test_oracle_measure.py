"""Pins of the oracle's measurement functions (oracle/lgv_oracle.c
orc_measure_walls, orc_dft_rows, orc_rotate_to_vev), independent of the
formulas they restate (ModelG/measurer.cxx:10-98, 101-114, 116-221, 231-286,
measurer.h:33-41).

The restatement contracts the charges with n-hat through six scalar
expressions (contract_rho).  The checks below build the antisymmetric 4x4
charge tensor explicitly instead -- rho_i0 = A_i, rho_ij = -eps_ijk V_k
(measurer.cxx:101-114) -- with its Hodge dual rho~_ab = 1/2 eps_abcd rho_cd
computed from the Levi-Civita symbol (not by swapping A and V), and check
what the mathematics fixes: walls are slice sums, the phase observables
rotate covariantly under SO(4), the transforms are numpy's rfft / N, and the
vev rotation is a projection onto n and its complement.  A dropped term, a
flipped sign or a swapped index in any of the six contractions fails one of
them.
"""
import itertools

import numpy as np
import pytest


def levi_civita4():
    e = np.zeros((4, 4, 4, 4))
    for p in itertools.permutations(range(4)):
        inv = sum(1 for i in range(4) for j in range(i + 1, 4) if p[i] > p[j])
        e[p] = -1.0 if inv % 2 else 1.0
    return e


EPS4 = levi_civita4()
EPS3 = EPS4[0, 1:, 1:, 1:]   # eps_ijk = eps_0ijk


def charge_tensor(A, V):
    """[..., 3] axial and vector charges -> [..., 4, 4] antisymmetric rho."""
    M = np.zeros(A.shape[:-1] + (4, 4), dtype=np.result_type(A, V))
    M[..., 1:, 0] = A
    M[..., 0, 1:] = -A
    M[..., 1:, 1:] = -np.einsum("ijk,...k->...ij", EPS3, V)
    return M


def dual(M):
    return 0.5 * np.einsum("abcd,...cd->...ab", EPS4, M)


def charges_of(M):
    """Inverse of charge_tensor: (A, V) of an antisymmetric 4x4 tensor."""
    A = M[..., 1:, 0]
    V = -0.5 * np.einsum("ijk,...ij->...k", EPS3, M[..., 1:, 1:])
    return A, V


def random_so4(rng):
    q, r = np.linalg.qr(rng.normal(size=(4, 4)))
    q = q * np.sign(np.diag(r))
    if np.linalg.det(q) < 0:
        q[:, 0] = -q[:, 0]
    return q


def site_phase_obs(u):
    """The 14 phase observables per site from the explicit tensor:
    |phi|, n (4), A' = rho n (4), V' = rho~ n (4), |phi|^2."""
    phi = u[..., :4]
    M = charge_tensor(u[..., 4:7], u[..., 7:10])
    phi2 = np.einsum("...a,...a->...", phi, phi)
    sigma = np.sqrt(phi2)
    tiny = phi2 <= 100 * np.finfo(np.float64).tiny
    n = np.where(tiny[..., None], np.array([1.0, 0, 0, 0]), phi / np.where(tiny, 1, sigma)[..., None])
    Ap = np.einsum("...ab,...b->...a", M, n)
    Vp = np.einsum("...ab,...b->...a", dual(M), n)
    return np.concatenate([sigma[..., None], n, Ap, Vp, phi2[..., None]], axis=-1)


def slice_means(obs):
    """obs[z, y, x, k] -> walls [3][k][N] (d = x, y, z): means over the other
    two axes (the slice sums / N^2 of measurer.cxx:88-97, 211-220)."""
    wx = obs.mean(axis=(0, 1)).T
    wy = obs.mean(axis=(0, 2)).T
    wz = obs.mean(axis=(1, 2)).T
    return np.stack([wx, wy, wz])


def random_field(rng, N, zero_sites=0):
    u = rng.normal(size=(N, N, N, 10))
    u[..., :4] += 0.4
    flat = u.reshape(-1, 10)
    for s in rng.choice(N ** 3, size=zero_sites, replace=False):
        flat[s, :4] = 0.0        # |phi| = 0: the unit frame n = (1,0,0,0)
    return u


@pytest.mark.parametrize("N", [4, 6, 8])
def test_walls_are_slice_means(orc, N):
    rng = np.random.default_rng(100 + N)
    u = random_field(rng, N, zero_sites=3)
    wall, wallphase, oavg = orc.measure_walls(u)
    plain = np.concatenate([u, np.sum(u[..., :4] ** 2, axis=-1)[..., None]], axis=-1)
    np.testing.assert_allclose(wall, slice_means(plain), rtol=1e-13, atol=1e-15)
    np.testing.assert_allclose(wallphase, slice_means(site_phase_obs(u)), rtol=1e-12,
                               atol=1e-14)
    # OAverage = volume means (computeDerivedObs, measurer.cxx:288-300)
    np.testing.assert_allclose(oavg, plain.reshape(-1, 11).mean(axis=0), rtol=1e-13, atol=1e-15)


def test_phase_walls_identity_frame(orc):
    """phi = (s, 0, 0, 0), s > 0: n = e0, A' = (0, A), V' = (0, V); with
    s < 0, n = -e0 flips both."""
    rng = np.random.default_rng(7)
    N = 4
    u = rng.normal(size=(N, N, N, 10))
    for sign in (1.0, -1.0):
        u[..., 0] = sign * (0.5 + rng.random(size=(N, N, N)))
        u[..., 1:4] = 0.0
        _, wp, _ = orc.measure_walls(u)
        plain = slice_means(u)
        np.testing.assert_allclose(wp[:, 0], sign * plain[:, 0], rtol=1e-14)
        np.testing.assert_allclose(wp[:, 1], sign, rtol=0, atol=0)
        np.testing.assert_allclose(wp[:, 2:5], 0, atol=0)
        for off, src in ((5, 4), (9, 7)):
            np.testing.assert_allclose(wp[:, off], 0, atol=1e-15)
            np.testing.assert_allclose(wp[:, off + 1:off + 4], sign * plain[:, src:src + 3],
                                       rtol=1e-13, atol=1e-15)


def test_phase_walls_so4_covariance(orc):
    """Rotating every site's phi by R and its charge tensor by R rho R^T
    (R in SO(4)) rotates the walls of n, A' and V' by R and leaves |phi| and
    |phi|^2 unchanged; a sign or index slip in one of the contractions (or
    in the dual's A <-> V swap) breaks the covariance."""
    rng = np.random.default_rng(21)
    N = 6
    u = random_field(rng, N)
    R = random_so4(rng)
    M = charge_tensor(u[..., 4:7], u[..., 7:10])
    Mr = np.einsum("ab,...bc,dc->...ad", R, M, R)
    ur = np.empty_like(u)
    ur[..., :4] = u[..., :4] @ R.T
    ur[..., 4:7], ur[..., 7:10] = charges_of(Mr)
    w, wp, _ = orc.measure_walls(u)
    w2, wp2, _ = orc.measure_walls(ur)
    np.testing.assert_allclose(wp2[:, 0], wp[:, 0], rtol=1e-13)
    np.testing.assert_allclose(wp2[:, 13], wp[:, 13], rtol=1e-13)
    np.testing.assert_allclose(w2[:, 10], w[:, 10], rtol=1e-13)
    for blk in (slice(1, 5), slice(5, 9), slice(9, 13)):
        np.testing.assert_allclose(wp2[:, blk], np.einsum("ab,dbn->dan", R, wp[:, blk]),
                                   rtol=1e-11, atol=1e-13)
    np.testing.assert_allclose(w2[:, :4], np.einsum("ab,dbn->dan", R, w[:, :4]), rtol=1e-11,
                               atol=1e-13)
    # A' and V' are orthogonal to n at every site (antisymmetry of rho)
    obs = site_phase_obs(u)
    np.testing.assert_allclose(np.einsum("...a,...a->...", obs[..., 1:5], obs[..., 5:9]), 0,
                               atol=1e-12)
    np.testing.assert_allclose(np.einsum("...a,...a->...", obs[..., 1:5], obs[..., 9:13]), 0,
                               atol=1e-12)


def test_dual_is_the_a_v_swap():
    """The Hodge dual of the tensor of (A, V) is the tensor of (V, A), the
    swap measurer.cxx:206-209 applies; a check of the test's own tensor."""
    rng = np.random.default_rng(3)
    A, V = rng.normal(size=(5, 3)), rng.normal(size=(5, 3))
    np.testing.assert_allclose(dual(charge_tensor(A, V)), charge_tensor(V, A), atol=1e-15)
    np.testing.assert_allclose(dual(dual(charge_tensor(A, V))), charge_tensor(A, V), atol=1e-15)


@pytest.mark.parametrize("N", [4, 8, 10, 16, 32])
def test_dft_rows_is_rfft_over_n(orc, N):
    rng = np.random.default_rng(N)
    rows = rng.normal(size=(7, N))
    out = orc.dft_rows(rows)
    ref = np.fft.rfft(rows, axis=-1) / N
    assert out.shape == (7, N // 2 + 1, 2)
    np.testing.assert_allclose(out[..., 0] + 1j * out[..., 1], ref, rtol=0, atol=1e-15)


def test_dft_rows_closed_form(orc):
    """cos(2 pi k x / N) -> 1/2 at k (1 at k = 0 and k = N/2), sin -> -i/2;
    the FFTW forward sign and the 1/N of measurer.h:33-41."""
    N = 12
    x = np.arange(N)
    rows, want = [], []
    for k in range(N // 2 + 1):
        rows.append(np.cos(2 * np.pi * k * x / N))
        w = np.zeros(N // 2 + 1, dtype=complex)
        w[k] = 1.0 if k in (0, N // 2) else 0.5
        want.append(w)
        if 0 < k < N // 2:
            rows.append(np.sin(2 * np.pi * k * x / N))
            w = np.zeros(N // 2 + 1, dtype=complex)
            w[k] = -0.5j
            want.append(w)
    out = orc.dft_rows(np.array(rows))
    np.testing.assert_allclose(out[..., 0] + 1j * out[..., 1], np.array(want), atol=1e-15)


def rotated_tensor(wk):
    """rotate_to_vev restated with the explicit tensor: wk[11][nk][2] ->
    [14][nk] complex (phi_s, phi_r[4], A'[4], V'[4], phi^2)."""
    z = wk[..., 0] + 1j * wk[..., 1]
    n = z[:4, 0].real.copy()
    if np.linalg.norm(n) < np.finfo(float).tiny:
        n = np.array([1.0, 0, 0, 0])
    n /= np.linalg.norm(n)
    M = charge_tensor(z[4:7].T, z[7:10].T)          # [nk, 4, 4]
    phis = n @ z[:4]
    phir = z[:4] - np.outer(n, phis)
    Ap = np.einsum("kab,b->ak", M, n)
    Vp = np.einsum("kab,b->ak", dual(M), n)
    return np.concatenate([phis[None], phir, Ap, Vp, z[10][None]])


def k_space_walls(rng, nk, vev):
    wk = rng.normal(size=(11, nk, 2))
    wk[:, 0, 1] = 0.0            # the k = 0 mode of a real wall is real
    wk[:4, 0, 0] = vev
    return wk


@pytest.mark.parametrize("vev", [(0.7, -0.2, 0.1, 0.3), (1.0, 0, 0, 0), (0, 0, 0, 0)])
def test_rotate_to_vev(orc, vev):
    rng = np.random.default_rng(5)
    nk = 9
    wk = k_space_walls(rng, nk, vev)
    rot = orc.rotate_to_vev(wk)
    got = rot[..., 0] + 1j * rot[..., 1]
    np.testing.assert_allclose(got, rotated_tensor(wk), rtol=0, atol=1e-14)
    # projection identities: phi = n phi_s + phi_r, n . phi_r = n . A' = n . V' = 0
    z = wk[..., 0] + 1j * wk[..., 1]
    n = np.array(vev, dtype=float)
    n = n / np.linalg.norm(n) if np.linalg.norm(n) > 0 else np.array([1.0, 0, 0, 0])
    np.testing.assert_allclose(np.outer(n, got[0]) + got[1:5], z[:4], atol=1e-14)
    for blk in (slice(1, 5), slice(5, 9), slice(9, 13)):
        np.testing.assert_allclose(n @ got[blk], 0, atol=1e-14)


def test_rotate_to_vev_so4_covariance(orc):
    """Rotating phi(k) by R and rho(k) by R rho R^T leaves phi_s and phi^2
    unchanged and rotates phi_r, A', V' by R."""
    rng = np.random.default_rng(17)
    nk = 6
    wk = k_space_walls(rng, nk, (0.3, 0.5, -0.4, 0.2))
    R = random_so4(rng)
    z = wk[..., 0] + 1j * wk[..., 1]
    M = charge_tensor(z[4:7].T, z[7:10].T)
    Mr = np.einsum("ab,kbc,dc->kad", R, M, R)
    zr = z.copy()
    zr[:4] = R @ z[:4]
    A, V = charges_of(Mr)
    zr[4:7], zr[7:10] = A.T, V.T
    wkr = np.stack([zr.real, zr.imag], axis=-1)
    r1 = orc.rotate_to_vev(wk)
    r2 = orc.rotate_to_vev(wkr)
    c1 = r1[..., 0] + 1j * r1[..., 1]
    c2 = r2[..., 0] + 1j * r2[..., 1]
    np.testing.assert_allclose(c2[0], c1[0], atol=1e-13)
    np.testing.assert_allclose(c2[13], c1[13], atol=0)
    for blk in (slice(1, 5), slice(5, 9), slice(9, 13)):
        np.testing.assert_allclose(c2[blk], R @ c1[blk], atol=1e-13)
