"""Pins the oracle's camera models (projection + Jacobians) against
independent computations: scipy's rotation library, the quaternion sandwich
product q (0,M) q* of PSBA/CL_files/compute_exQT.cl:36-69, the pinhole special
cases of every model, Bouguet's published distortion model (kc(1:5)), and
central finite differences."""
import numpy as np
import pytest
from scipy.spatial.transform import Rotation

from conftest import rel

RNG = np.random.default_rng(7)


def qmul(a, b):
    s1, v1 = a[0], a[1:]
    s2, v2 = b[0], b[1:]
    return np.concatenate([[s1 * s2 - v1 @ v2], s1 * v2 + s2 * v1 + np.cross(v1, v2)])


def rand_case(cnp):
    q0 = RNG.standard_normal(4)
    q0 /= np.linalg.norm(q0)
    q0 *= np.sign(q0[0])
    v = 0.2 * RNG.standard_normal(3)
    t = np.array([0.3, -0.2, 8.0]) + 0.5 * RNG.standard_normal(3)
    M = RNG.uniform(-1.5, 1.5, 3)
    K = np.array([800.0, 320.0, 240.0, 1.01, 0.5])
    if cnp == 6:
        cam = np.concatenate([v, t])
    elif cnp == 9:
        cam = np.concatenate([[820.0, -0.07, 0.02], v, t])
    elif cnp == 11:
        cam = np.concatenate([[810.0, 318.0, 244.0, 0.99, 0.3], v, t])
    else:
        cam = np.concatenate([[810.0, 318.0, 244.0, 0.99, 0.3, -0.08, 0.02, 0.001, -0.002, 0.003], v, t])
    return K, q0, cam, M


def camera_point(q0, v, t, M):
    """X' = R(q_v (x) q0) M + t with scipy's rotations (scalar-last)."""
    qv = np.concatenate([[np.sqrt(1 - v @ v)], v])
    Rv = Rotation.from_quat([qv[1], qv[2], qv[3], qv[0]])
    R0 = Rotation.from_quat([q0[1], q0[2], q0[3], q0[0]])
    return (Rv * R0).apply(M) + t


def bouguet(x, y, kc):
    r2 = x * x + y * y
    radial = 1 + kc[0] * r2 + kc[1] * r2 ** 2 + kc[4] * r2 ** 3
    dx = 2 * kc[2] * x * y + kc[3] * (r2 + 2 * x * x)
    dy = kc[2] * (r2 + 2 * y * y) + 2 * kc[3] * x * y
    return radial * x + dx, radial * y + dy


def expected_uv(cnp, K, q0, cam, M):
    v, t = cam[cnp - 6:cnp - 3], cam[cnp - 3:]
    X = camera_point(q0, v, t, M)
    x, y = X[0] / X[2], X[1] / X[2]
    if cnp == 6:
        fx, cx, cy, ar, sk = K
        xd, yd = x, y
    elif cnp == 9:
        fx, (cx, cy, ar, sk) = cam[0], K[1:]
        d = 1 + cam[1] * (x * x + y * y) + cam[2] * (x * x + y * y) ** 2
        xd, yd = d * x, d * y
    elif cnp == 11:
        fx, cx, cy, ar, sk = cam[:5]
        xd, yd = x, y
    else:
        fx, cx, cy, ar, sk = cam[:5]
        xd, yd = bouguet(x, y, cam[5:10])
    return np.array([fx * xd + sk * yd + cx, fx * ar * yd + cy])


@pytest.mark.parametrize("cnp", [6, 9, 11, 16])
def test_projection_matches_independent_rotation(O, cnp):
    for _ in range(50):
        K, q0, cam, M = rand_case(cnp)
        uv = O.project(cnp, K, q0, cam, M)
        assert rel(uv, expected_uv(cnp, K, q0, cam, M)) < 1e-13


def test_pose_model_is_the_quaternion_sandwich(O):
    """compute_exQT.cl composes q = q_v (x) q0 and rotates by q (0,M) q*."""
    for _ in range(50):
        K, q0, cam, M = rand_case(6)
        v = cam[:3]
        q = qmul(np.concatenate([[np.sqrt(1 - v @ v)], v]), q0)
        qc = q * np.array([1, -1, -1, -1])
        X = qmul(qmul(q, np.concatenate([[0.0], M])), qc)[1:] + cam[3:]
        uv_ref = np.array([(K[0] * X[0] + K[4] * X[1] + K[1] * X[2]) / X[2], (K[0] * K[3] * X[1] + K[2] * X[2]) / X[2]])
        assert rel(O.project(6, K, q0, cam, M), uv_ref) < 1e-13


def test_models_reduce_to_pinhole(O):
    for _ in range(20):
        K, q0, cam6, M = rand_case(6)
        v, t = cam6[:3], cam6[3:]
        uv6 = O.project(6, K, q0, cam6, M)
        cam9 = np.concatenate([[K[0], 0.0, 0.0], v, t])
        assert rel(O.project(9, K, q0, cam9, M), uv6) < 1e-14
        cam11 = np.concatenate([K, v, t])
        assert rel(O.project(11, K * 0 + 1, q0, cam11, M), uv6) < 1e-14
        cam16 = np.concatenate([K, np.zeros(5), v, t])
        assert rel(O.project(16, K * 0 + 1, q0, cam16, M), uv6) < 1e-14


def test_identity_rotation_closed_form(O):
    """v = 0, q0 = 1: u = fx X/Z + skew Y/Z + cx, v = fx ar Y/Z + cy."""
    K = np.array([700.0, 300.0, 200.0, 1.2, 0.4])
    M = np.array([0.5, -0.25, 4.0])
    cam = np.array([0, 0, 0, 0.1, 0.2, 1.0])
    uv = O.project(6, K, np.array([1.0, 0, 0, 0]), cam, M)
    X = M + cam[3:]
    assert rel(uv, [700 * X[0] / X[2] + 0.4 * X[1] / X[2] + 300, 700 * 1.2 * X[1] / X[2] + 200]) < 1e-15


@pytest.mark.parametrize("cnp", [6, 9, 11, 16])
def test_jacobian_matches_central_differences(O, cnp):
    for _ in range(20):
        K, q0, cam, M = rand_case(cnp)
        _, JA, JB = O.project(cnp, K, q0, cam, M, jac=True)
        for k in range(cnp):
            h = 1e-6 * max(1.0, abs(cam[k]))
            cp, cm = cam.copy(), cam.copy()
            cp[k] += h
            cm[k] -= h
            fd = (O.project(cnp, K, q0, cp, M) - O.project(cnp, K, q0, cm, M)) / (2 * h)
            assert np.allclose(JA[:, k], fd, rtol=2e-6, atol=1e-6 * np.max(np.abs(JA))), (cnp, k)
        for k in range(3):
            h = 1e-6
            mp, mm = M.copy(), M.copy()
            mp[k] += h
            mm[k] -= h
            fd = (O.project(cnp, K, q0, cam, mp) - O.project(cnp, K, q0, cam, mm)) / (2 * h)
            assert np.allclose(JB[:, k], fd, rtol=2e-6, atol=1e-6 * np.max(np.abs(JB)))


def test_jacobian_at_identity_local_rotation(O):
    """At v = 0 the rotation columns are d/dv of R(q_v) M0: 2 (e_k x M0)."""
    K = np.array([800.0, 320.0, 240.0, 1.0, 0.0])
    q0 = np.array([0.9, 0.1, -0.3, 0.2])
    q0 /= np.linalg.norm(q0)
    M = np.array([0.3, -0.4, 1.2])
    cam = np.array([0, 0, 0, 0.1, 0.2, 6.0])
    _, JA, JB = O.project(6, K, q0, cam, M, jac=True)
    R0 = Rotation.from_quat([q0[1], q0[2], q0[3], q0[0]])
    M0 = R0.apply(M)
    X = M0 + cam[3:]
    G = np.array([[800 / X[2], 0, -800 * X[0] / X[2] ** 2], [0, 800 / X[2], -800 * X[1] / X[2] ** 2]])
    for k in range(3):
        e = np.zeros(3)
        e[k] = 1
        assert rel(JA[:, k], G @ (2 * np.cross(e, M0))) < 1e-13
    assert rel(JA[:, 3:], G) < 1e-14
    assert rel(JB, G @ R0.as_matrix()) < 1e-13
