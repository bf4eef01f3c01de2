"""Pins the oracle's normal-equation, Schur-complement and back-substitution
stages (PSBA/CL_files/compute_U/V/Wblks/g/Vinv/Yblks/S/ea/eb/dpb.cl) against
a brute-force dense J^T J built from the per-observation Jacobians: blocks of
J^T J, the dense Schur complement, and the full damped normal-equation solve
(J^T J + mu I) dp = J^T e with numpy."""
import numpy as np
import pytest

from conftest import rel


def dense_system(o, sc, coeff=1.0, coeff_g=1.0):
    C, nc, npt, no = sc.cnp, sc.n_cams, sc.n_pts, sc.n_obs
    JA = o.buffer("JA").reshape(no, 2, C)
    JB = o.buffer("JB").reshape(no, 2, 3)
    e = o.buffer("EX").reshape(no, 2)
    n = C * nc
    J = np.zeros((2 * no, n + 3 * npt))
    for k in range(no):
        i, j = sc.iidx[k], sc.jidx[k]
        J[2 * k:2 * k + 2, C * j:C * j + C] = JA[k]
        J[2 * k:2 * k + 2, n + 3 * i:n + 3 * i + 3] = JB[k]
    return J, e.reshape(-1)


SCENES = [(0, dict(n_pts=60)), (1, dict(n_pts=80, n_cams=8)), (1, dict(cnp=16, n_pts=80, n_cams=8)),
          (2, dict(n_pts=120, n_cams=6))]


@pytest.mark.parametrize("config,kw", SCENES)
def test_normal_equations_are_blocks_of_JtJ(P, O, config, kw):
    sc = P.Scene(config, 77, **kw)
    o = O.Oracle.from_scene(sc)
    o.jacobian_normal(2.0, -2.0)  # the dogleg's coefficients
    J, e = dense_system(o, sc)
    C, nc, no = sc.cnp, sc.n_cams, sc.n_obs
    n = C * nc
    H = 2.0 * J.T @ J
    U = o.buffer("U").reshape(nc, C, C)
    for j in range(nc):
        assert rel(U[j], H[C * j:C * j + C, C * j:C * j + C]) < 1e-13
    V = o.buffer("V").reshape(sc.n_pts, 3, 3)
    for i in range(sc.n_pts):
        assert rel(V[i], H[n + 3 * i:n + 3 * i + 3, n + 3 * i:n + 3 * i + 3]) < 1e-13
    W = o.buffer("W").reshape(no, C, 3)
    for k in range(no):
        i, j = sc.iidx[k], sc.jidx[k]
        assert rel(W[k], H[C * j:C * j + C, n + 3 * i:n + 3 * i + 3]) < 1e-12
    assert rel(o.buffer("G"), -2.0 * J.T @ e) < 1e-12
    # fixed camera 0: zero Jacobian columns
    assert np.all(J[:, :C] == 0)


@pytest.mark.parametrize("config,kw", SCENES)
def test_schur_and_step_match_dense_solve(P, O, config, kw):
    sc = P.Scene(config, 78, **kw)
    o = O.Oracle.from_scene(sc)
    o.jacobian_normal(1.0, 1.0)
    J, e = dense_system(o, sc)
    C, nc = sc.cnp, sc.n_cams
    n = C * nc
    mu = 1e-3 * o.max_uvdiag()
    H = J.T @ J + mu * np.eye(J.shape[1])
    g = J.T @ e
    free = np.ones(J.shape[1], bool)
    free[:n][sc.fixmask.astype(bool)] = False
    A, B, D = H[:n, :n], H[:n, n:], H[n:, n:]
    Dinv = np.linalg.inv(D)  # block-diagonal of 3x3 point blocks
    S = A - B @ Dinv @ B.T
    S[~free[:n], :] = 0
    S[:, ~free[:n]] = 0
    S[~free[:n], ~free[:n]] = 1.0
    o.schur(mu)
    So = o.buffer("S").reshape(n, n)
    tl = np.tril_indices(n)
    assert rel(So[tl], S[tl]) < 1e-11
    ea = g[:n] - B @ Dinv @ g[n:]
    assert rel(o.buffer("EA")[free[:n]], ea[free[:n]]) < 1e-10
    Vi = o.buffer("VINV").reshape(sc.n_pts, 3, 3)
    for i in range(sc.n_pts):
        assert rel(Vi[i], np.linalg.inv(D[3 * i:3 * i + 3, 3 * i:3 * i + 3])) < 1e-12
    assert o.factor_solve()
    o.backsub()
    dp = o.buffer("DP")
    Hf = H[np.ix_(free, free)]
    ref = np.linalg.solve(Hf, g[free])
    assert rel(dp[free], ref) < 1e-8
    assert np.all(dp[~free] == 0)
    # the explicit-inverse (reference-faithful) mode gives the same step
    o.schur(mu)
    assert o.factor_solve(solve_mode=O.SOLVE_INVERSE)
    o.backsub()
    assert rel(o.buffer("DP")[free], ref) < 1e-7


def test_jmul_is_J_times_x(P, O):
    sc = P.Scene(0, 79, n_pts=50)
    o = O.Oracle.from_scene(sc)
    o.jacobian_normal()
    J, _ = dense_system(o, sc)
    x = np.random.default_rng(1).standard_normal(J.shape[1])
    assert rel(o.jmul(x), J @ x) < 1e-13
