"""The offline post-processing tools langevina_amd/bin/x2k.exe and
x2k_rotated.exe (ModelG/x2k.cxx:43-121, ModelG/x2k_rotated.cxx:35-150, run by
ModelG/x2k.py:39-67 for the wallx/wally/wallz datasets of each event file).

FFTW is absent here, so x2k is pinned by closed forms and by numpy's rfft (a
library r2c transform with the same sign and the 1/N of x2k.cxx:109-110), not
by FFTW's bits.  x2k_rotated is pinned by special cases and by what the
mathematics fixes: with n = (1,0,0,0) the rotation is the identity on the
field labels; phi_r, A' and V' are orthogonal to n; and the map is covariant
under SO(4) (rotating phi and the charge tensor rotates phi_r, A', V' and
leaves phi_s unchanged), which a wrong sign or index in the tensor fails.
"""
import os
import subprocess

import numpy as np
import pytest

from h5tools import h5_read, h5_shape, h5_write

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
X2K = os.path.join(ROOT, "langevina_amd", "bin", "x2k.exe")
X2KR = os.path.join(ROOT, "langevina_amd", "bin", "x2k_rotated.exe")


def run(exe, *args):
    if not os.path.exists(exe):
        pytest.fail(f"{exe} missing: build it with python -c "
                    "'from langevina_amd.build import build_driver; build_driver()'")
    r = subprocess.run([exe, *map(str, args)], capture_output=True, text=True, timeout=120)
    return r.returncode, r.stdout, r.stderr


def cplx(a):
    return a[..., 0] + 1j * a[..., 1]


@pytest.mark.parametrize("size", [8, 12, 16])
def test_x2k_closed_form(tmp_path, size):
    """The signal of ModelG/x2ktest.py:1-21: cos(2 pi k x/N) and sin(2 pi k x/N),
    k = 0, 1, 2, transform to 1 (k = 0), 1/2 and -i/2 at mode k, zero elsewhere."""
    x = np.arange(size, dtype=np.float64)
    data = np.zeros((2, 3, size))
    for k in range(3):
        data[0, k] = np.cos(2 * np.pi * k * x / size)
        data[1, k] = np.sin(2 * np.pi * k * x / size)
    f = tmp_path / "x2ktest.h5"
    h5_write(str(f), "x2ktest", data)
    rc, out, err = run(X2K, f, "x2ktest")
    assert rc == 0, err
    assert f"Creating dataset x2ktest_k in {tmp_path}/x2ktest_out.h5" in out
    wk = h5_read(str(tmp_path / "x2ktest_out.h5"), "/x2ktest_k")
    assert wk.shape == (2, 3, size // 2 + 1, 2)
    expect = np.zeros((2, 3, size // 2 + 1), dtype=complex)
    expect[0, 0, 0] = 1.0
    for k in (1, 2):
        expect[0, k, k] = 0.5
        expect[1, k, k] = -0.5j
    np.testing.assert_allclose(cplx(wk), expect, rtol=0, atol=1e-15)


def test_x2k_matches_rfft_and_keeps_other_datasets(tmp_path):
    rng = np.random.default_rng(3)
    walls = {d: rng.normal(size=(5, 11, 32)) + 0.3 for d in "xyz"}
    f = str(tmp_path / "run_0000.h5")
    for d in "xyz":
        h5_write(f, f"wall{d}", walls[d])
    for d in "xyz":   # ModelG/x2k.py:39-50: one call per direction into one _out file
        rc, _, err = run(X2K, f, f"wall{d}")
        assert rc == 0, err
    out = str(tmp_path / "run_0000_out.h5")
    for d in "xyz":
        wk = h5_read(out, f"/wall{d}_k")
        ref = np.fft.rfft(walls[d], axis=-1) / 32
        np.testing.assert_allclose(cplx(wk), ref, rtol=0, atol=1e-14)
    # a second run replaces the dataset instead of appending rows
    rc, _, err = run(X2K, f, "wallx")
    assert rc == 0, err
    assert h5_shape(out, "/wallx_k") == (5, 11, 17, 2)


def test_x2k_rejects_a_dataset_that_is_not_a_wall(tmp_path):
    f = str(tmp_path / "e.h5")
    h5_write(f, "phi", np.zeros((3, 11)))
    rc, _, err = run(X2K, f, "phi")
    assert rc == 1 and "has rank 2, expected 3" in err


def test_x2k_errors(tmp_path):
    rc, out, _ = run(X2K, "only_one_argument.h5")
    assert rc == 0 and "Usage x2k filename.h5 datasetname" in out
    rc, _, err = run(X2K, tmp_path / "missing.h5", "wallx")
    assert rc == 1 and "Unable to open" in err
    f = str(tmp_path / "a.h5")
    h5_write(f, "wallx", np.zeros((2, 11, 8)))
    rc, _, err = run(X2K, f, "wally")
    assert rc == 1 and "wally does not exist" in err


def rotated_numpy(wk):
    """x2k_rotated restated with an explicit 4x4 charge tensor
    rho_i0 = A_i, rho_ij = -eps_ijk V_k (measurer.cxx:101-114)."""
    z = cplx(wk)                            # [T, F, K]
    T, _, K = z.shape
    n = z[:, :4, 0].real.copy()
    norm = np.linalg.norm(n, axis=1)
    n[norm < np.finfo(float).tiny] = [1, 0, 0, 0]
    n /= np.linalg.norm(n, axis=1)[:, None]
    eps = np.zeros((3, 3, 3))
    for i, j, k in [(0, 1, 2), (1, 2, 0), (2, 0, 1)]:
        eps[i, j, k], eps[j, i, k] = 1, -1

    def tensor(a, v):   # a, v: [T, 3, K] -> [T, 4, 4, K]
        M = np.zeros((T, 4, 4, K), dtype=complex)
        M[:, 1:, 0] = a
        M[:, 0, 1:] = -a
        M[:, 1:, 1:] = -np.einsum("ijk,tkq->tijq", eps, v)
        return M

    phi, A, V = z[:, :4], z[:, 4:7], z[:, 7:10]
    phis = np.einsum("ta,taq->tq", n, phi)
    phir = phi - n[:, :, None] * phis[:, None]
    Ar = np.einsum("tabq,tb->taq", tensor(A, V), n)
    Vr = np.einsum("tabq,tb->taq", tensor(V, A), n)
    return np.concatenate([phis[:, None], phir, Ar, Vr], axis=1)


def make_out_file(tmp_path, wk, name="run_0000_out.h5"):
    f = str(tmp_path / name)
    h5_write(f, "wallz_k", wk)
    return f


def test_x2k_rotated_identity_frame(tmp_path):
    """n = (1,0,0,0): phi_s = phi_0, phi_r = (0, phi_1..3), A' = (0, A), V' = (0, V)."""
    rng = np.random.default_rng(5)
    wk = rng.normal(size=(3, 11, 9, 2))
    wk[:, 0, 0] = [2.5, 0.0]
    wk[:, 1:4, 0, 0] = 0.0
    f = make_out_file(tmp_path, wk)
    rc, out, err = run(X2KR, f, "wallz_k")
    assert rc == 0, err
    r = cplx(h5_read(f, "/wallz_k_rotated"))
    z = cplx(wk)
    assert r.shape == (3, 13, 9)
    np.testing.assert_allclose(r[:, 0], z[:, 0], rtol=0, atol=1e-15)
    np.testing.assert_allclose(r[:, 1], 0, atol=1e-15)
    np.testing.assert_allclose(r[:, 2:5], z[:, 1:4], rtol=0, atol=1e-15)
    for off, src in ((5, 4), (9, 7)):
        np.testing.assert_allclose(r[:, off], 0, atol=1e-15)
        np.testing.assert_allclose(r[:, off + 1:off + 4], z[:, src:src + 3], rtol=0, atol=1e-15)


def random_so4(rng):
    q, r = np.linalg.qr(rng.normal(size=(4, 4)))
    q = q * np.sign(np.diag(r))
    if np.linalg.det(q) < 0:
        q[:, 0] = -q[:, 0]
    return q


def test_x2k_rotated_orthogonality_and_so4_covariance(tmp_path):
    rng = np.random.default_rng(11)
    T, K = 4, 7
    wk = rng.normal(size=(T, 11, K, 2))
    f = make_out_file(tmp_path, wk)
    assert run(X2KR, f, "wallz_k")[0] == 0
    r = cplx(h5_read(f, "/wallz_k_rotated"))
    np.testing.assert_allclose(r, rotated_numpy(wk), rtol=0, atol=1e-13)

    z = cplx(wk)
    n = z[:, :4, 0].real / np.linalg.norm(z[:, :4, 0].real, axis=1)[:, None]
    for block in (slice(1, 5), slice(5, 9), slice(9, 13)):
        np.testing.assert_allclose(np.einsum("ta,taq->tq", n, r[:, block]), 0, atol=1e-13)
    np.testing.assert_allclose(n[:, :, None] * r[:, 0][:, None] + r[:, 1:5], z[:, :4],
                               atol=1e-13)

    # rotate phi -> R phi and the tensor rho -> R rho R^T, then rebuild (A, V)
    R = random_so4(rng)
    eps = np.zeros((3, 3, 3))
    for i, j, k in [(0, 1, 2), (1, 2, 0), (2, 0, 1)]:
        eps[i, j, k], eps[j, i, k] = 1, -1
    M = np.zeros((T, 4, 4, K), dtype=complex)
    M[:, 1:, 0], M[:, 0, 1:] = z[:, 4:7], -z[:, 4:7]
    M[:, 1:, 1:] = -np.einsum("ijk,tkq->tijq", eps, z[:, 7:10])
    Mr = np.einsum("ab,tbcq,dc->tadq", R, M, R)
    zr = z.copy()
    zr[:, :4] = np.einsum("ab,tbq->taq", R, z[:, :4])
    zr[:, 4:7] = Mr[:, 1:, 0]
    zr[:, 7:10] = -0.5 * np.einsum("ijk,tijq->tkq", eps, Mr[:, 1:, 1:])
    wkr = np.stack([zr.real, zr.imag], axis=-1)
    f2 = make_out_file(tmp_path, wkr, "rot_out.h5")
    assert run(X2KR, f2, "wallz_k")[0] == 0
    r2 = cplx(h5_read(f2, "/wallz_k_rotated"))
    np.testing.assert_allclose(r2[:, 0], r[:, 0], atol=1e-12)
    for block in (slice(1, 5), slice(5, 9), slice(9, 13)):
        np.testing.assert_allclose(r2[:, block], np.einsum("ab,tbq->taq", R, r[:, block]),
                                   atol=1e-12)
    # the extra field (phi^2 wall) is not used
    assert r.shape[1] == 13


def test_x2k_rotated_zero_vev_row_uses_unit_frame(tmp_path):
    rng = np.random.default_rng(2)
    wk = rng.normal(size=(2, 10, 5, 2))
    wk[1, :4, 0, 0] = 0.0
    f = make_out_file(tmp_path, wk)
    assert run(X2KR, f, "wallz_k")[0] == 0
    r = cplx(h5_read(f, "/wallz_k_rotated"))
    np.testing.assert_allclose(r, rotated_numpy(wk), rtol=0, atol=1e-13)
    np.testing.assert_allclose(r[1, 0], cplx(wk)[1, 0], atol=1e-15)


def test_x2k_rotated_errors(tmp_path):
    rc, out, _ = run(X2KR, tmp_path / "run_0000.h5", "wallx_k")
    assert rc == 0 and "Usage x2k_rotated filename_out.h5 datasetname" in out
    rc, _, err = run(X2KR, tmp_path / "missing_out.h5", "wallx_k")
    assert rc == 1 and "Unable to open" in err
    f = make_out_file(tmp_path, np.zeros((1, 11, 3, 2)))
    rc, out, _ = run(X2KR, f, "wallx_k")
    assert rc == 1 and "The input dataset wallx_k does not exist" in out
    f = str(tmp_path / "bad_out.h5")
    h5_write(f, "w_k", np.zeros((1, 8, 3, 2)))
    rc, _, err = run(X2KR, f, "w_k")
    assert rc == 1 and "is not [T, F>=10, K, 2]" in err
