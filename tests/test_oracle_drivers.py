"""The oracle's LM (PSBA/levmar.cpp:45-256), dogleg (PSBA/trust_region.cpp:
49-595) and hybrid loop (PSBA/main.cpp:193-208): end results pinned against an
independent solver (MINPACK Levenberg-Marquardt through scipy) and noise-free
convergence; traces checked for the invariants the reference's control logic
implies."""
import numpy as np
import pytest
from scipy.optimize import least_squares


def scipy_minimum(O, sc):
    """Minimise the same residuals with MINPACK (parameters: free ones)."""
    C, nc = sc.cnp, sc.n_cams
    n = C * nc
    free = np.concatenate([~sc.fixmask.astype(bool), np.ones(3 * sc.n_pts, bool)])
    p0 = np.concatenate([sc.cams, sc.pts])

    def res(pf):
        p = p0.copy()
        p[free] = pf
        out = np.empty(2 * sc.n_obs)
        for k in range(sc.n_obs):
            i, j = sc.iidx[k], sc.jidx[k]
            uv = O.project(C, sc.K[5 * j:5 * j + 5], sc.q0[4 * j:4 * j + 4], p[C * j:C * j + C],
                           p[n + 3 * i:n + 3 * i + 3])
            out[2 * k:2 * k + 2] = sc.meas[2 * k:2 * k + 2] - uv
        return out

    r = least_squares(res, p0[free], method="lm", xtol=1e-15, ftol=1e-15, gtol=1e-15, max_nfev=20000)
    return float(2 * r.cost)


@pytest.mark.parametrize("config,kw", [(0, dict(n_pts=40)), (1, dict(n_pts=40, n_cams=6))])
def test_lm_reaches_the_least_squares_minimum(P, O, config, kw):
    sc = P.Scene(config, 11, **kw)
    o = O.Oracle.from_scene(sc)
    r = o.levmar(max_iter=200, flags=O.FLAG_NO_HANDOFF, rel_tol=1e-14)
    ref = scipy_minimum(O, sc)
    assert abs(r["final_err"] - ref) / ref < 1e-7


def test_hybrid_reaches_the_least_squares_minimum(P, O):
    sc = P.Scene(0, 12, n_pts=40)
    o = O.Oracle.from_scene(sc)
    r = o.hybrid()
    ref = scipy_minimum(O, sc)
    # DP_NO_CHANGE, ERR_SMALL_ENOUGH or PASS (the 50-iteration cap)
    assert r["code"] in (5, 6, 7)
    assert abs(r["final_err"] - ref) / ref < 1e-6


@pytest.mark.parametrize("config,kw", [(0, {}), (1, dict(n_pts=400)), (2, dict(n_pts=2000, n_cams=20))])
def test_noise_free_scene_converges_to_zero(P, O, config, kw):
    sc = P.Scene(config, 13, noise=0.0, outlier_frac=0.0, **kw)
    o = O.Oracle.from_scene(sc)
    init = o.residual()
    r = o.levmar(max_iter=100, flags=O.FLAG_NO_HANDOFF)
    assert r["final_err"] < 1e-14 * init + 1e-10


def test_lm_trace_invariants(P, O):
    sc = P.Scene(0, 14)
    o = O.Oracle.from_scene(sc)
    r = o.levmar()
    costs = []
    for t in r["trace"]:
        assert t.kind == 0
        if t.accepted:
            assert t.rho > 0
            costs.append(t.cost)
        elif t.factor_ok and np.isfinite(t.rho):
            assert t.rho <= 0
    assert all(b < a for a, b in zip(costs, costs[1:]))
    # levmar.cpp:200-205: after an accepted step mu <- mu max(1/3, 1-(2 rho-1)^3)
    tr = r["trace"]
    for a, b in zip(tr, tr[1:]):
        if a.accepted:
            f = max(1.0 / 3.0, 1 - (2 * a.rho - 1) ** 3)
            assert b.damp == pytest.approx(a.damp * f, rel=1e-14)


def test_lm_hands_over_after_five_good_steps(P, O):
    """levmar.cpp:215-219: five consecutive |rho-1| < 0.2 -> ITER_TURN_TO_TR."""
    sc = P.Scene(0, 15)
    o = O.Oracle.from_scene(sc)
    r = o.levmar()
    good = 0
    for t in r["trace"]:
        if t.accepted:
            good = good + 1 if abs(t.rho - 1) < 0.2 else 0
    if r["code"] == 2:
        assert good >= 5


def test_dogleg_trace_radius_rules(P, O):
    """trust_region.cpp:223-237: radius /4 on a poor step, x2 (<= 1e4) on a
    very good accepted one, unchanged otherwise."""
    sc = P.Scene(0, 16)
    o = O.Oracle.from_scene(sc)
    o.levmar()
    r = o.trust_region()
    tr = [t for t in r["trace"] if t.kind == 1]
    for a, b in zip(tr, tr[1:]):
        if not np.isfinite(a.rho):
            continue
        if a.rho < 0.25 or not a.accepted:
            assert b.damp == pytest.approx(a.damp / 4) or b.damp == pytest.approx(1.0)
        elif a.rho >= 0.75:
            assert b.damp == pytest.approx(min(2 * a.damp, 1e4))
