"""Dense reduced-system kernels of the oracle: Cholesky solve and the
reference-faithful explicit inverse (PSBA/CL_files/SPD_inv.cl) against
numpy/LAPACK, and the dogleg's modified Cholesky (PSBA/CL_files/cholmod_blk.cl)
against the properties that define it."""
import numpy as np
import pytest

from conftest import rel

RNG = np.random.default_rng(3)


def spd(n, cond=1e3):
    Q, _ = np.linalg.qr(RNG.standard_normal((n, n)))
    d = np.logspace(0, np.log10(cond), n)
    return (Q * d) @ Q.T


@pytest.mark.parametrize("n", [6, 42, 99, 150])
def test_cholesky_solve_matches_lapack(O, n):
    S = spd(n)
    b = RNG.standard_normal(n)
    ok, x, L = O.dense_factor_solve(S, b)
    assert ok
    assert rel(x, np.linalg.solve(S, b)) < 1e-12
    assert rel(L, np.linalg.cholesky(S)) < 1e-12


@pytest.mark.parametrize("n", [6, 42, 99])
def test_explicit_inverse_matches_lapack(O, n):
    S = spd(n)
    b = RNG.standard_normal(n)
    ok, x, _ = O.dense_factor_solve(S, b, solve_mode=O.SOLVE_INVERSE)
    assert ok
    assert rel(x, np.linalg.solve(S, b)) < 1e-11


@pytest.mark.parametrize("mode", [0, 1])
def test_factor_fails_on_indefinite(O, mode):
    S = spd(30)
    S[7, 7] = -5.0
    ok, _, _ = O.dense_factor_solve(S, np.ones(30), solve_mode=mode)
    assert not ok


def test_modchol_spd_needs_no_modification(O):
    n = 60
    A = RNG.standard_normal((n, n)) * 0.1
    S = A @ A.T + n * np.eye(n)
    lam, L, E = O.dense_modchol(S)
    assert lam < 1e-12 * n
    assert np.max(np.abs(E)) < 1e-12 * np.max(np.abs(S))
    assert rel(np.tril(L), np.linalg.cholesky(S)) < 1e-12


@pytest.mark.parametrize("n", [9, 30, 60, 88, 91])
def test_modchol_indefinite_is_S_plus_nonnegative_diagonal(O, n):
    Q, _ = np.linalg.qr(RNG.standard_normal((n, n)))
    d = np.concatenate([-np.linspace(1, 5, n // 3), np.linspace(1, 10, n - n // 3)])
    S = (Q * d) @ Q.T
    lam, L, E = O.dense_modchol(S)
    L = np.tril(L)
    R = L @ L.T - S
    off = R - np.diag(np.diag(R))
    assert np.max(np.abs(off)) < 1e-10 * np.max(np.abs(S))
    assert np.allclose(np.diag(R), E, atol=1e-10 * np.max(np.abs(S)))
    assert np.all(E >= -1e-12)
    assert lam == pytest.approx(abs(E.sum()) / n, rel=1e-12)
    assert np.all(np.linalg.eigvalsh(S + np.diag(E)) > 0)
