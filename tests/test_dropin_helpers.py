"""The host helpers of the drop-in misc.h (PSBA/misc.h:5-23), called through
their C++ symbols in libpsba.so: quat2matrix against scipy's rotation matrix,
camsFormatTrans on a packed camera list, vec2quat as the inverse of quat2vec,
and emalloc_.  No device is needed."""
import ctypes as ct
import subprocess

import numpy as np
import pytest
from scipy.spatial.transform import Rotation

from conftest import ROOT

LIB = ROOT + "/psba_amd/libpsba.so"


def cxx_symbols():
    raw = subprocess.run(["nm", "-D", "--defined-only", LIB], stdout=subprocess.PIPE, text=True,
                         check=True).stdout.split("\n")
    mangled = [l.split()[-1] for l in raw if l.strip()]
    dem = subprocess.run(["c++filt"], input="\n".join(mangled), stdout=subprocess.PIPE, text=True,
                         check=True).stdout.split("\n")
    return {d.split("(")[0]: m for m, d in zip(mangled, dem) if "(" in d}


@pytest.fixture(scope="module")
def H(P):
    L = P.lib()
    sym = cxx_symbols()
    dp = ct.POINTER(ct.c_double)
    f = {}
    for name, args, res in [("quat2matrix", [dp, ct.c_int, dp], None),
                            ("camsFormatTrans", [dp, ct.c_int, ct.c_int, ct.c_int, dp], None),
                            ("quat2vec", [dp, ct.c_int, dp, ct.c_int], None),
                            ("vec2quat", [dp, ct.c_int, dp, ct.c_int], None),
                            ("emalloc_", [ct.c_char_p, ct.c_int, ct.c_size_t], ct.c_void_p)]:
        fn = getattr(L, sym[name])
        fn.argtypes = args
        fn.restype = res
        f[name] = fn
    return f


def dptr(a):
    return a.ctypes.data_as(ct.POINTER(ct.c_double))


def unit_quats(rng, n):
    q = rng.standard_normal((n, 4))
    q /= np.linalg.norm(q, axis=1)[:, None]
    q *= np.sign(q[:, :1])
    return q


def scipy_R(q):
    # scipy takes (x, y, z, w)
    return Rotation.from_quat([q[1], q[2], q[3], q[0]]).as_matrix()


def test_quat2matrix_matches_scipy(H):
    rng = np.random.default_rng(3)
    for q in unit_quats(rng, 50):
        out4 = np.empty(9)
        H["quat2matrix"](dptr(np.ascontiguousarray(q)), 4, dptr(out4))
        assert np.allclose(out4.reshape(3, 3), scipy_R(q), rtol=0, atol=1e-14)
        out3 = np.empty(9)
        H["quat2matrix"](dptr(np.ascontiguousarray(q[1:])), 3, dptr(out3))
        # s = sqrt(1 - |v|^2) carries an absolute error of about eps / s
        assert np.allclose(out3.reshape(3, 3), scipy_R(q), rtol=0, atol=1e-14 + 1e-15 / q[0])
    # identity and a half turn about z
    out = np.empty(9)
    H["quat2matrix"](dptr(np.array([1.0, 0, 0, 0])), 4, dptr(out))
    assert np.array_equal(out, np.eye(3).reshape(-1))
    H["quat2matrix"](dptr(np.array([0.0, 0, 0, 1])), 4, dptr(out))
    assert np.array_equal(out, np.diag([-1.0, -1.0, 1.0]).reshape(-1))


@pytest.mark.parametrize("dim", [3, 4])
def test_cams_format_trans(H, dim):
    rng = np.random.default_rng(4)
    n = 7
    q = unit_quats(rng, n)
    t = rng.standard_normal((n, 3))
    packed = np.hstack([q if dim == 4 else q[:, 1:], t])
    out = np.empty((n, 12))
    H["camsFormatTrans"](dptr(np.ascontiguousarray(packed)), n, dim, 3, dptr(out))
    for j in range(n):
        assert np.allclose(out[j, :9].reshape(3, 3), scipy_R(q[j]), rtol=0, atol=1e-14)
        assert np.array_equal(out[j, 9:], t[j])


def test_vec2quat_inverts_quat2vec(H):
    rng = np.random.default_rng(5)
    for lead in (0, 5):
        for q in unit_quats(rng, 20):
            row = np.concatenate([rng.standard_normal(lead), q, rng.standard_normal(3)])
            vec = np.empty(lead + 6)
            H["quat2vec"](dptr(row), lead + 7, dptr(vec), lead + 6)
            back = np.empty(lead + 7)
            H["vec2quat"](dptr(vec), lead + 6, dptr(back), lead + 7)
            assert np.array_equal(back[:lead], row[:lead])
            assert np.allclose(back[lead:lead + 4], q, rtol=0, atol=1e-15 + 1e-15 / q[0])
            assert np.array_equal(back[lead + 4:], row[lead + 4:])
    # a negative scalar part is folded to the same rotation with q0 >= 0
    q = np.array([-0.5, 0.5, 0.5, 0.5])
    vec = np.empty(6)
    H["quat2vec"](dptr(np.concatenate([q, [1.0, 2.0, 3.0]])), 7, dptr(vec), 6)
    assert np.array_equal(vec, [-0.5, -0.5, -0.5, 1.0, 2.0, 3.0])


def test_emalloc(H):
    p = H["emalloc_"](b"test.c", 1, 4096)
    assert p
    ct.memset(p, 0, 4096)
    ct.CDLL(None).free(ct.c_void_p(p))
