"""The oracle's robust loss (an extension: the reference minimises the plain
sum of squares).  Pinned against scipy: the loss and its derivative against
scipy's own Huber / Cauchy functions (those of scipy.optimize.least_squares),
the reweighted Jacobian rows against sqrt(rho') times the plain ones, and the
robust LM's end point as a stationary point of sum rho(s) that scipy's
L-BFGS-B cannot improve."""
import numpy as np
import pytest
from scipy.optimize._lsq.least_squares import cauchy, huber

LOSSES = [(1, huber, "huber"), (2, cauchy, "cauchy")]


def scipy_rho(fn, d, s):
    r = np.empty((3, s.size))
    fn(np.asarray(s, float) / (d * d), r, False)
    return d * d * r[0], r[1]


@pytest.mark.parametrize("kind,fn,name", LOSSES)
def test_loss_matches_scipy(O, kind, fn, name):
    d = 1.7
    s = np.concatenate([np.linspace(0, 20, 41), [1e-12, d * d, d * d * (1 + 1e-12), 1e6]])
    rho_ref, drho_ref = scipy_rho(fn, d, s)
    got = np.array([O.loss_eval(kind, d, x) for x in s])
    assert np.allclose(got[:, 0], rho_ref, rtol=1e-14, atol=1e-300)
    assert np.allclose(got[:, 1], drho_ref, rtol=1e-14, atol=0)


@pytest.mark.parametrize("kind,fn,name", LOSSES)
def test_robust_cost_and_weighted_rows(P, O, kind, fn, name):
    sc = P.Scene(0, 21, n_pts=200, outlier_frac=0.05)
    d = 2.0
    plain = O.Oracle.from_scene(sc)
    rob = O.Oracle.from_scene(sc)
    rob.set_loss(kind, d)
    plain.residual()
    ex = plain.buffer("EX").reshape(-1, 2)
    s = (ex ** 2).sum(1)
    rho, drho = scipy_rho(fn, d, s)
    assert rob.residual() == pytest.approx(rho.sum(), rel=1e-13)
    plain.jacobian_normal(1.0, 1.0)
    rob.jacobian_normal(1.0, 1.0)
    w = np.sqrt(drho)
    C = sc.cnp
    assert np.allclose(rob.buffer("JA").reshape(-1, 2 * C), plain.buffer("JA").reshape(-1, 2 * C) * w[:, None],
                       rtol=1e-14, atol=1e-300)
    assert np.allclose(rob.buffer("JB").reshape(-1, 6), plain.buffer("JB").reshape(-1, 6) * w[:, None],
                       rtol=1e-14, atol=1e-300)
    assert np.allclose(rob.buffer("EX").reshape(-1, 2), ex * w[:, None], rtol=1e-14, atol=1e-300)


def objective(O, sc, fn, d):
    """F(p) = sum_k rho(|e_k|^2) with scipy's rho, and its gradient over the
    free parameters, dF/dp = -2 sum_k rho'(s_k) J_k^T e_k (J = dpi/dp)."""
    C, nc = sc.cnp, sc.n_cams
    n = C * nc
    free = np.concatenate([~sc.fixmask.astype(bool), np.ones(3 * sc.n_pts, bool)])

    def F(p, grad=False):
        e = np.empty((sc.n_obs, 2))
        JA = np.empty((sc.n_obs, 2, C))
        JB = np.empty((sc.n_obs, 2, 3))
        for k in range(sc.n_obs):
            i, j = sc.iidx[k], sc.jidx[k]
            uv, JA[k], JB[k] = O.project(C, sc.K[5 * j:5 * j + 5], sc.q0[4 * j:4 * j + 4], p[C * j:C * j + C],
                                         p[n + 3 * i:n + 3 * i + 3], jac=True)
            e[k] = sc.meas[2 * k:2 * k + 2] - uv
        rho, drho = scipy_rho(fn, d, (e ** 2).sum(1))
        if not grad:
            return rho.sum()
        g = np.zeros(p.size)
        for k in range(sc.n_obs):
            i, j = sc.iidx[k], sc.jidx[k]
            g[C * j:C * j + C] += -2 * drho[k] * (e[k] @ JA[k])
            g[n + 3 * i:n + 3 * i + 3] += -2 * drho[k] * (e[k] @ JB[k])
        return rho.sum(), g[free]

    return F, free


@pytest.mark.parametrize("kind,fn,name", LOSSES)
def test_robust_lm_reaches_a_minimum_of_the_robust_cost(P, O, kind, fn, name):
    """The end point is a stationary point of sum rho(s) (scipy's rho): the
    gradient has shrunk by 1e-6 from the start, the reported cost is that
    objective, and scipy's robust trust-region solver, started there, finds
    nothing lower."""
    sc = P.Scene(0, 22, n_pts=40, outlier_frac=0.08)
    d = 2.0
    o = O.Oracle.from_scene(sc)
    o.set_loss(kind, d)
    r = o.levmar(max_iter=300, flags=O.FLAG_NO_HANDOFF, rel_tol=1e-15)
    F, free = objective(O, sc, fn, d)
    p0 = np.concatenate([sc.cams, sc.pts])
    cams, pts = o.read_params()
    p1 = np.concatenate([cams, pts])
    _, g0 = F(p0, grad=True)
    f1, g1 = F(p1, grad=True)
    assert f1 == pytest.approx(r["final_err"], rel=1e-12)
    assert np.linalg.norm(g1) < 1e-6 * np.linalg.norm(g0)

    def scipy_f(x):
        p = p1.copy()
        p[free] = x
        return F(p, grad=True)

    res = minimize_check(scipy_f, p1[free])
    assert res >= f1 * (1 - 1e-9)


def minimize_check(fg, x0):
    from scipy.optimize import minimize
    r = minimize(fg, x0, jac=True, method="L-BFGS-B", options=dict(maxiter=200, ftol=1e-16, gtol=1e-12))
    return float(r.fun)
