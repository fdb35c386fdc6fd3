"""Pins the oracle's O(4) rotation (ModelG/O4AlgebraHelper.h).

* the three known answers of ModelG/O4AlgebraSignTest.cxx:20-38;
* an independent construction: phi -> L phi R with L = exp(-(V+A)/2),
  R = exp((V-A)/2) as unit quaternions equals expm(M_L + M_R) phi, where
  M_L/M_R are the 4x4 matrices of left/right Hamilton multiplication by the
  pure quaternions and expm is scipy's matrix exponential;
* norm preservation and the inverse rotation.
"""
import numpy as np
import scipy.linalg


def hamilton(p, q):
    """Textbook Hamilton product of quaternions (w, x, y, z)."""
    a0, a = p[0], np.asarray(p[1:])
    b0, b = q[0], np.asarray(q[1:])
    return np.concatenate([[a0 * b0 - a @ b], a0 * b + b0 * a + np.cross(a, b)])


def left_matrix(p):
    return np.stack([hamilton(p, e) for e in np.eye(4)], axis=1)


def right_matrix(p):
    return np.stack([hamilton(e, p) for e in np.eye(4)], axis=1)


def test_sign_conventions_known_answers(orc):
    r = 1 / np.sqrt(2)
    out = orc.o4_rotation([0, 0, 0], [np.pi / 4, 0, 0], [1, 0, 0, 0])
    np.testing.assert_allclose(out, [r, -r, 0, 0], atol=1e-15)
    out = orc.o4_rotation([0, 0, np.pi / 4], [0, 0, 0], [0, 1, 0, 0])
    np.testing.assert_allclose(out, [0, r, -r, 0], atol=1e-15)
    out = orc.o4_rotation([0, 0, np.pi / 4], [0, 0, 0], [0, 0, 1, 0])
    np.testing.assert_allclose(out, [0, r, r, 0], atol=1e-15)


def test_quatmul_is_hamilton_product(orc):
    rng = np.random.default_rng(1)
    for _ in range(50):
        p, q = rng.normal(size=4), rng.normal(size=4)
        np.testing.assert_allclose(orc.quatmul(p, q), hamilton(p, q), rtol=1e-14, atol=1e-14)


def test_expmul_is_quaternion_exponential(orc):
    rng = np.random.default_rng(2)
    for _ in range(50):
        v = rng.normal(size=3) * rng.uniform(0.01, 3)
        # exp of the pure quaternion -v via the matrix exponential
        m = scipy.linalg.expm(left_matrix(np.concatenate([[0], -v])))
        np.testing.assert_allclose(orc.expmul(v), m[:, 0], rtol=1e-13, atol=1e-14)
    np.testing.assert_array_equal(orc.expmul([1e-21, 0, 0]), [1, 0, 0, 0])


def test_rotation_equals_so4_matrix_exponential(orc):
    rng = np.random.default_rng(3)
    for _ in range(100):
        V, A = rng.normal(size=3), rng.normal(size=3)
        phi = rng.normal(size=4)
        tmp1 = -0.5 * (V - A)
        tmp2 = 0.5 * (V + A)
        G = left_matrix(np.concatenate([[0], -tmp2])) + right_matrix(np.concatenate([[0], -tmp1]))
        want = scipy.linalg.expm(G) @ phi
        np.testing.assert_allclose(orc.o4_rotation(V, A, phi), want, rtol=1e-12, atol=1e-13)


def test_norm_preserved_and_inverse(orc):
    rng = np.random.default_rng(4)
    for _ in range(200):
        V, A = rng.normal(size=3) * 2, rng.normal(size=3) * 2
        phi = rng.normal(size=4)
        out = orc.o4_rotation(V, A, phi)
        assert abs(np.linalg.norm(out) - np.linalg.norm(phi)) < 1e-14 * np.linalg.norm(phi) * 10
        back = orc.o4_rotation(-V, -A, out)
        np.testing.assert_allclose(back, phi, rtol=1e-13, atol=1e-13)
