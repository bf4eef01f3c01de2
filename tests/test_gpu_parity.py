"""HIP path against the oracle on the same seeded scenes, through the C-ABI.

Stage by stage (residual, Jacobian/normal equations, V^-1, Schur complement,
reduced right-hand side, factor/solve, back-substitution) and whole drivers
(LM per trial, dogleg and the hybrid loop with the scale gauge fixed), for
every camera model (cnp 6, 9, 11, 16).  The plain FP64 path must reproduce
the oracle's per-trial cost and final parameters within 1e-9 relative.
"""
import numpy as np
import pytest

from conftest import rel

pytestmark = pytest.mark.gpu

STAGE_SCENES = [(0, {}), (1, dict(n_pts=600)), (1, dict(cnp=16, n_pts=600)), (2, dict(n_pts=3000, n_cams=20))]
IDS = ["c0-cnp6", "c1-cnp11", "c1-cnp16", "c2-cnp9"]


def pair(P, O, sc, fix=True):
    return P.Context.from_scene(sc, fix=fix), O.Oracle.from_scene(sc, fix=fix)


def fix_gauge(sc):
    """Make the reduced system nonsingular at lambda = 0, which the dogleg
    needs.  Camera 0 is already fixed; hold camera 1's first translation
    coordinate too (the scale).  With all five intrinsics free (cnp >= 11)
    every camera is a general projective camera, so the 15-parameter
    projective ambiguity leaves 3 more flat directions: hold every camera's
    skew (the zero-skew constraint) as well."""
    C = sc.cnp
    sc.fixmask[C + C - 3] = 1
    if C >= 11:
        sc.fixmask.reshape(-1, C)[:, 4] = 1
    return sc


@pytest.mark.parametrize("config,kw", STAGE_SCENES, ids=IDS)
def test_stages_match_oracle(P, O, gpu, config, kw):
    sc = P.Scene(config, 1000 + config, **kw)
    g, o = pair(P, O, sc)
    assert rel(g.residual(), o.residual()) < 1e-13
    for coeff, coeff_g in ((1.0, 1.0), (2.0, -2.0)):
        g.jacobian_normal(coeff, coeff_g)
        o.jacobian_normal(coeff, coeff_g)
        for b in ("EX", "JA", "JB", "U", "V", "W", "G"):
            assert rel(g.buffer(b), o.buffer(b)) < 1e-12, b
    g.jacobian_normal(1.0, 1.0)
    o.jacobian_normal(1.0, 1.0)
    mu = 1e-3 * o.max_uvdiag()
    g.schur(mu)
    o.schur(mu)
    assert rel(g.buffer("VINV"), o.buffer("VINV")) < 1e-12
    assert rel(g.buffer("EA"), o.buffer("EA")) < 1e-11
    n = sc.cnp * sc.n_cams
    tl = np.tril_indices(n)
    assert rel(g.buffer("S").reshape(n, n)[tl], o.buffer("S").reshape(n, n)[tl]) < 1e-11
    assert g.factor_solve() and o.factor_solve()
    g.backsub()
    o.backsub()
    assert rel(g.buffer("DP"), o.buffer("DP")) < 1e-9
    assert rel(g.buffer("EB"), o.buffer("EB")) < 1e-11


def test_schur_is_deterministic(P, gpu):
    sc = P.Scene(2, 7, n_pts=4000, n_cams=24)
    a = P.Context.from_scene(sc)
    b = P.Context.from_scene(sc)
    out = []
    for c in (a, a, b):
        c.jacobian_normal()
        c.schur(1.0)
        out.append((c.buffer("S"), c.buffer("EA")))
    for S, ea in out[1:]:
        assert np.array_equal(S, out[0][0]) and np.array_equal(ea, out[0][1])


def close(a, b, tol):
    """|a - b| <= tol |b|; two NaNs (a trial whose cost was never evaluated,
    e.g. the one that stops on a vanishing step) agree."""
    if np.isnan(a) or np.isnan(b):
        return bool(np.isnan(a) and np.isnan(b))
    return abs(a - b) <= tol * abs(b)


def compare_traces(tg, to, tol=1e-9, damp_tol=1e-6):
    """Per trial: the same decisions (factor ok, accepted).  The cost of every
    accepted trial -- the per-iteration cost -- agrees within tol.  A rejected
    trial's cost only decides the rejection (the damping update after it does
    not use the value), and at the far points such trials reach the cost is
    arbitrarily sensitive to the step (a point pushed towards a camera's focal
    plane), so there only the decision is compared.  The damping is compared
    more loosely: its update uses the gain ratio, whose numerator F_old - F_new
    loses relative accuracy like eps * F / (F_old - F_new) as the solve
    converges (down to ~1e-8 F here), so 1e-6 is what two correct FP64
    evaluations can promise."""
    assert len(tg) == len(to)
    for a, b in zip(tg, to):
        assert (a.itno, a.kind, a.factor_ok, a.accepted) == (b.itno, b.kind, b.factor_ok, b.accepted)
        if a.accepted:
            assert close(a.cost, b.cost, tol), (a.itno, a.cost, b.cost)
        assert close(a.damp, b.damp, damp_tol), (a.itno, a.damp, b.damp)


LM_SCENES = [(0, {}), (1, dict(n_pts=1500)), (1, dict(cnp=16, n_pts=1500)), (2, dict(n_pts=6000, n_cams=30)),
             (1, {}), (1, dict(cnp=16))]


@pytest.mark.parametrize("config,kw", LM_SCENES, ids=IDS + ["c1-cnp11-full", "c1-cnp16-full"])
def test_levmar_trial_by_trial(P, O, gpu, config, kw):
    sc = P.Scene(config, 2000 + config, **kw)
    if not kw and config == 1:
        assert (sc.n_cams, sc.n_pts) == (54, 5200)  # the stated size of the variable-intrinsics scene
    O.set_threads(16)
    g, o = pair(P, O, sc)
    rg = g.levmar()
    ro = o.levmar()
    assert (rg["code"], rg["itno"]) == (ro["code"], ro["itno"])
    compare_traces(rg["trace"], ro["trace"])
    assert rel(rg["final_err"], ro["final_err"]) < 1e-9
    assert rel(np.concatenate(g.read_params()), np.concatenate(o.read_params())) < 1e-9


def test_levmar_pure_with_relative_stop(P, O, gpu):
    """The benchmark's solver (LM only, stopping on a relative cost decrease
    below 1e-8 or 100 iterations) on the sparse-scene geometry without gross
    outliers, through convergence."""
    sc = P.Scene(2, 2100, n_pts=8000, n_cams=40, outlier_frac=0.0)
    g, o = pair(P, O, sc)
    rg = g.levmar(max_iter=100, flags=P.FLAG_NO_HANDOFF, rel_tol=1e-8)
    ro = o.levmar(max_iter=100, flags=O.FLAG_NO_HANDOFF, rel_tol=1e-8)
    assert (rg["code"], rg["itno"]) == (ro["code"], ro["itno"])
    assert rg["code"] in (P.ITER_DP_NO_CHANGE, P.ITER_ERR_SMALL_ENOUGH)
    compare_traces(rg["trace"], ro["trace"])
    assert rel(np.concatenate(g.read_params()), np.concatenate(o.read_params())) < 1e-9


def test_levmar_with_outliers_first_iterations(P, O, gpu):
    """With 1% gross outliers the least-squares trajectory is chaotic: rounding
    differences of 1e-15 grow by orders of magnitude per wild (rejected) step,
    and two correct FP64 evaluation orders part after some 45 iterations.  The
    first 30 iterations agree trial by trial."""
    sc = P.Scene(2, 2100, n_pts=8000, n_cams=40)
    g, o = pair(P, O, sc)
    rg = g.levmar(max_iter=30, flags=P.FLAG_NO_HANDOFF, rel_tol=1e-8)
    ro = o.levmar(max_iter=30, flags=O.FLAG_NO_HANDOFF, rel_tol=1e-8)
    assert (rg["code"], rg["itno"]) == (ro["code"], ro["itno"])
    compare_traces(rg["trace"], ro["trace"])
    assert rel(np.concatenate(g.read_params()), np.concatenate(o.read_params())) < 1e-9


# The dogleg solves the reduced system at lambda = 0.  Its condition number
# (gauge fixed) is ~1e8 for the pose-only scene, ~5e9 for the f,k1,k2 scene and
# ~4e10 with all five intrinsics free; two correct FP64 factorizations give
# steps that agree to about cond * eps, so the free-intrinsics scene is held
# to 1e-7 instead of 1e-9.
# At the stated size of the variable-intrinsics scene (54 cameras, 5,200
# points) the dogleg's reduced systems are worse conditioned still: the
# oracle run with every reduction in descending order (a second valid FP64
# evaluation) differs from itself by 6e-8 to 3e-6 relative in accepted costs
# and up to 3e-6 in final parameters, with identical decisions.  Those cases
# (tol None) are held to ten times the oracle's own spread on the same scene.
TR_SCENES = [(0, {}, 1e-9), (1, dict(n_pts=800, n_cams=20), 1e-7),
             (2, dict(n_pts=3000, n_cams=20, outlier_frac=0.0), 1e-9),
             (1, {}, None), (1, dict(cnp=16), None)]
TR_IDS = ["c0-cnp6", "c1-cnp11", "c2-cnp9", "c1-cnp11-full", "c1-cnp16-full"]


def order_spread(O, sc, run):
    """Relative spread of the oracle against itself with every reduction in
    descending index order: (max over accepted trials of the cost difference,
    final-parameter difference); the decisions must agree."""
    out = []
    try:
        for desc in (False, True):
            O.set_sum_order(desc)
            o = O.Oracle.from_scene(sc, fix=True)
            r = run(o)
            out.append((r, np.concatenate(o.read_params())))
    finally:
        O.set_sum_order(False)
    (ra, pa), (rb, pb) = out
    assert [(a.itno, a.kind, a.factor_ok, a.accepted) for a in ra["trace"]] == \
        [(b.itno, b.kind, b.factor_ok, b.accepted) for b in rb["trace"]]
    dc = max([abs(a.cost - b.cost) / abs(b.cost) for a, b in zip(ra["trace"], rb["trace"]) if a.accepted] + [0.0])
    return dc, rel(pa, pb)


def dogleg_tolerances(O, sc, tol, run):
    if tol is not None:
        return tol, tol
    dc, dp = order_spread(O, sc, run)
    return max(1e-9, 10 * dc), max(1e-9, 10 * dp)


@pytest.mark.parametrize("config,kw,tol", TR_SCENES, ids=TR_IDS)
def test_dogleg_after_lm_with_gauge_fixed(P, O, gpu, config, kw, tol):
    O.set_threads(16)
    sc = fix_gauge(P.Scene(config, 3000 + config, **kw))

    def lm_then_tr(o):
        o.levmar()
        return o.trust_region()

    tc, tp = dogleg_tolerances(O, sc, tol, lm_then_tr)
    g, o = pair(P, O, sc)
    g.levmar()
    o.levmar()
    rg = g.trust_region()
    ro = o.trust_region()
    assert (rg["code"], rg["itno"]) == (ro["code"], ro["itno"])
    compare_traces(rg["trace"], ro["trace"], tol=tc)
    for a, b in zip(rg["trace"], ro["trace"]):
        assert close(a.lambda_, b.lambda_, tc) or a.lambda_ == b.lambda_
    assert rel(np.concatenate(g.read_params()), np.concatenate(o.read_params())) < tp


@pytest.mark.parametrize("config,kw,tol", TR_SCENES, ids=TR_IDS)
def test_hybrid_with_gauge_fixed(P, O, gpu, config, kw, tol):
    O.set_threads(16)
    sc = fix_gauge(P.Scene(config, 4000 + config, **kw))
    tc, tp = dogleg_tolerances(O, sc, tol, lambda o: o.hybrid())
    g, o = pair(P, O, sc)
    rg = g.hybrid()
    ro = o.hybrid()
    assert (rg["code"], rg["itno"]) == (ro["code"], ro["itno"])
    compare_traces(rg["trace"], ro["trace"], tol=tc)
    assert rel(rg["final_err"], ro["final_err"]) < tc
    assert rel(np.concatenate(g.read_params()), np.concatenate(o.read_params())) < tp


def test_dogleg_from_the_start(P, O, gpu):
    """trust_region() entered first (Delta = 1, lambda = 0), scale fixed."""
    sc = fix_gauge(P.Scene(0, 4100))
    g, o = pair(P, O, sc)
    rg = g.trust_region()
    ro = o.trust_region()
    assert (rg["code"], rg["itno"]) == (ro["code"], ro["itno"])
    compare_traces(rg["trace"], ro["trace"])


# ---- dense reduced-system kernels on loaded matrices ----

def spd(n, rng, cond=1e4):
    Q, _ = np.linalg.qr(rng.standard_normal((n, n)))
    return (Q * np.logspace(0, np.log10(cond), n)) @ Q.T


@pytest.mark.parametrize("n_cams", [1, 7, 11, 100, 300])
def test_dense_factor_solve_matches_lapack(P, gpu, n_cams):
    rng = np.random.default_rng(n_cams)
    n = 6 * n_cams
    with P.Context(6, n_cams, 4, 8) as c:
        S = spd(n, rng)
        b = rng.standard_normal(n)
        c.load_S(S, b)
        ok, x = c.factor_only()
        assert ok
        assert rel(x, np.linalg.solve(S, b)) < 1e-9
        L = c.buffer("L").reshape(n, n)
        assert rel(np.tril(L), np.linalg.cholesky(S)) < 1e-11


@pytest.mark.parametrize("n_cams", [7, 40])
def test_dense_factor_reads_the_lower_triangle_only(P, gpu, n_cams):
    """The Schur assembly writes only the lower block positions of S, so the
    factorisation must not read above the diagonal: garbage there changes
    nothing (the same factor and solution, bit for bit)."""
    rng = np.random.default_rng(100 + n_cams)
    n = 6 * n_cams
    S = spd(n, rng, cond=1e8)
    b = rng.standard_normal(n)
    junk = np.tril(S) + np.triu(rng.standard_normal((n, n)) * 1e3, 1)
    with P.Context(6, n_cams, 4, 8) as c:
        c.load_S(S, b)
        ok1, x1 = c.factor_only()
        L1 = np.tril(c.buffer("L").reshape(n, n))
        c.load_S(junk, b)
        ok2, x2 = c.factor_only()
        L2 = np.tril(c.buffer("L").reshape(n, n))
    assert ok1 and ok2
    assert np.array_equal(x1, x2) and np.array_equal(L1, L2)
    assert rel(x1, np.linalg.solve(S, b)) < 1e-7


def test_dense_factor_reports_indefinite(P, gpu):
    rng = np.random.default_rng(9)
    n = 6 * 20
    with P.Context(6, 20, 4, 8) as c:
        S = spd(n, rng)
        S[77, 77] = -1.0
        c.load_S(S, np.ones(n))
        ok, _ = c.factor_only()
        assert not ok
        # the padding must survive a failed factorization: a good matrix after it factors
        S2 = spd(n, rng)
        c.load_S(S2, np.ones(n))
        ok, x = c.factor_only()
        assert ok and rel(x, np.linalg.solve(S2, np.ones(n))) < 1e-9


@pytest.mark.parametrize("cnp,n_cams", [(6, 10), (11, 8), (16, 7)])
def test_modchol_matches_oracle(P, O, gpu, cnp, n_cams):
    rng = np.random.default_rng(cnp)
    n = cnp * n_cams
    Q, _ = np.linalg.qr(rng.standard_normal((n, n)))
    d = np.concatenate([-np.linspace(1, 5, n // 4), np.linspace(1, 10, n - n // 4)])
    S = (Q * d) @ Q.T
    with P.Context(cnp, n_cams, 4, 8) as c:
        c.load_S(S)
        lam = c.modchol()
        lam_o, L_o, E_o = O.dense_modchol(S)
        assert abs(lam - lam_o) <= 1e-9 * lam_o
        assert rel(np.tril(c.buffer("L").reshape(n, n)), np.tril(L_o)) < 1e-9


# ---- edge cases ----

def test_ragged_tracks_and_unseen_camera(P, O, gpu):
    """Points left with a single observation (rank-2 V, damped by mu) and a
    camera with no observation at all (zero U block)."""
    sc = P.Scene(0, 5000, n_pts=300)
    rng = np.random.default_rng(0)
    keep = np.ones(sc.n_obs, bool)
    # drop all of camera 3's observations, and all but one of some tracks
    keep[sc.jidx == 3] = False
    for i in rng.choice(sc.n_pts, 40, replace=False):
        ks = np.nonzero((sc.iidx == i) & keep)[0]
        keep[ks[1:]] = False
    # points still need one observation each
    seen = np.zeros(sc.n_pts, bool)
    seen[sc.iidx[keep]] = True
    assert seen.all()
    sc.iidx, sc.jidx = sc.iidx[keep], sc.jidx[keep]
    sc.meas = sc.meas.reshape(-1, 2)[keep].reshape(-1)
    sc.n_obs = int(keep.sum())
    g, o = pair(P, O, sc)
    rg = g.levmar(max_iter=30, flags=P.FLAG_NO_HANDOFF)
    ro = o.levmar(max_iter=30, flags=O.FLAG_NO_HANDOFF)
    assert (rg["code"], rg["itno"]) == (ro["code"], ro["itno"])
    compare_traces(rg["trace"], ro["trace"])


def test_bad_index_is_rejected(P, gpu):
    sc = P.Scene(0, 1)
    c = P.Context(sc.cnp, sc.n_cams, sc.n_pts, sc.n_obs)
    c.upload(sc.K, sc.meas, sc.q0, sc.cams, sc.pts)
    j = sc.jidx.copy()
    k = int(np.nonzero(sc.iidx[1:] == sc.iidx[:-1])[0][0])
    j[k], j[k + 1] = j[k + 1], j[k]  # camera order within a point broken
    with pytest.raises(P.PsbaError):
        c.set_index(sc.iidx, j)
    i = sc.iidx.copy()
    i[0] = sc.n_pts  # point index out of range
    with pytest.raises(P.PsbaError):
        c.set_index(i, sc.jidx)
    with pytest.raises(P.PsbaError):
        c.levmar()  # no index: state error, nothing launched



# ---- mixed precision (FP32 Jacobian / W / V / Schur products, FP64 factor and cost) ----

@pytest.mark.parametrize("config,kw", [(0, {}), (2, dict(n_pts=3000, n_cams=20))], ids=["c0-cnp6", "c2-cnp9"])
def test_mixed_stages_track_fp64(P, O, gpu, config, kw):
    """FP32 storage and chunk accumulation: the normal-equation blocks and the
    reduced system agree with the FP64 oracle to FP32 accuracy."""
    sc = P.Scene(config, 5100 + config, **kw)
    g = P.Context.from_scene(sc, precision=P.PREC_MIXED)
    o = O.Oracle.from_scene(sc)
    g.residual()
    o.residual()
    g.jacobian_normal()
    o.jacobian_normal()
    for b in ("JA", "JB", "W", "U", "V", "G"):
        assert rel(g.buffer(b), o.buffer(b)) < 2e-5, b
    mu = 1e-3 * o.max_uvdiag()
    g.schur(mu)
    o.schur(mu)
    n = sc.cnp * sc.n_cams
    tl = np.tril_indices(n)
    assert rel(g.buffer("S").reshape(n, n)[tl], o.buffer("S").reshape(n, n)[tl]) < 2e-5
    assert rel(g.buffer("EA"), o.buffer("EA")) < 2e-5


def test_mixed_solve_reaches_fp64_minimum(P, O, gpu):
    """The mixed path converges to the FP64 oracle's minimum: final cost (and so
    RMS reprojection error) within 1e-6 relative, outlier-free sparse scene."""
    sc = P.Scene(2, 5200, n_pts=8000, n_cams=40, outlier_frac=0.0)
    g = P.Context.from_scene(sc, precision=P.PREC_MIXED)
    o = O.Oracle.from_scene(sc)
    rg = g.levmar(max_iter=100, flags=P.FLAG_NO_HANDOFF, rel_tol=1e-8)
    ro = o.levmar(max_iter=100, flags=O.FLAG_NO_HANDOFF, rel_tol=1e-8)
    assert rg["code"] in (P.ITER_DP_NO_CHANGE, P.ITER_ERR_SMALL_ENOUGH)
    assert abs(rg["final_err"] - ro["final_err"]) <= 1e-6 * ro["final_err"]


def test_mixed_context_rejects_dogleg(P, gpu):
    sc = P.Scene(0, 1)
    g = P.Context.from_scene(sc, precision=P.PREC_MIXED)
    with pytest.raises(P.PsbaError, match="FP64 contexts only"):
        g.trust_region()


def test_vinv_small_determinant_branch_on_device(P, O, gpu):
    """k_vinv on hand-built blocks (tests/test_oracle_pins.py derives their
    results by hand): the small-determinant branch of compute_Vinv.cl:31-73
    with its sign quirk and zero pivot, the damped and regular cases, each
    against the oracle's inverse."""
    from test_oracle_pins import HAND, S
    sc = P.Scene(0, 3)
    blocks = [S * M for M, _ in HAND]
    blocks.append(S * np.array([[1.0, 2, 0], [2, 5, 0], [0, 0, 1]]))  # zero pivot: adj / 0
    blocks.append(np.array([[4.0, 1, 0.5], [1, 3, 0.2], [0.5, 0.2, 2]]))  # regular
    V = np.array([blocks[k % len(blocks)] for k in range(sc.n_pts)])
    with P.Context.from_scene(sc) as g:
        for mu in (0.0, 1e-7):
            out = g.vinv_blocks(V, mu)
            for i in range(len(blocks)):
                ref, _ = O.vinv3(V[i], mu)
                fin = np.isfinite(ref)
                assert np.array_equal(np.isfinite(out[i]), fin), i
                assert np.allclose(out[i][fin], ref[fin], rtol=1e-12, atol=0), (i, mu)
                if mu == 0.0 and i < len(HAND):
                    M, sign = HAND[i]
                    assert np.allclose(out[i], sign * np.linalg.inv(V[i]), rtol=1e-12, atol=0)


DEVCTL_CASES = [(0, {}, dict()), (0, {}, dict(max_iter=3)), (1, dict(n_pts=1500), dict()),
                (2, dict(n_pts=6000, n_cams=30), dict(max_iter=100, rel_tol=1e-8)),
                (2, dict(n_pts=6000, n_cams=30), dict(max_iter=100, rel_tol=1e-8, mixed=True))]


@pytest.mark.parametrize("config,kw,opts", DEVCTL_CASES, ids=["c0", "c0-3it", "c1", "c2-rel", "c2-mixed"])
def test_device_control_matches_host_control(P, gpu, config, kw, opts):
    """psba_levmar with the decision on the device (the default: trials queued
    back to back, no read-back per trial) against the same driver deciding on
    the host after each trial: the same trace record for record, bit for bit,
    and the same parameters."""
    opts = dict(opts)
    mixed = opts.pop("mixed", False)
    sc = P.Scene(config, 3000 + config, **kw)
    prec = P.PREC_MIXED if mixed else P.PREC_FP64
    flags = P.FLAG_NO_HANDOFF if "rel_tol" in opts else 0
    out = []
    for extra in (0, P.FLAG_HOST_CONTROL):
        with P.Context.from_scene(sc, precision=prec) as g:
            r = g.levmar(flags=flags | extra, **opts)
            out.append((r, np.concatenate(g.read_params()), g.levmar_entry_cost()))
    (rd, pd, ed), (rh, ph, eh) = out
    assert (rd["code"], rd["itno"], rd["final_err"], ed) == (rh["code"], rh["itno"], rh["final_err"], eh)
    assert len(rd["trace"]) == len(rh["trace"]) > 0
    for a, b in zip(rd["trace"], rh["trace"]):
        fa = [getattr(a, f) for f, _ in a._fields_]
        fb = [getattr(b, f) for f, _ in b._fields_]
        assert np.array_equal(np.array(fa, dtype=float), np.array(fb, dtype=float), equal_nan=True), (fa, fb)
    assert np.array_equal(pd, ph)


def test_device_control_handoff_and_hybrid(P, O, gpu):
    """Without PSBA_FLAG_NO_HANDOFF the device decision hands over to the dogleg
    (ITER_TURN_TO_TR after five good steps) exactly as the host loop does, and
    the hybrid solve through it matches the oracle."""
    sc = fix_gauge(P.Scene(0, 4000))
    with P.Context.from_scene(sc) as g:
        r = g.levmar()
    with P.Context.from_scene(sc) as g:
        rh = g.levmar(flags=P.FLAG_HOST_CONTROL)
    assert r["code"] == rh["code"] == P.ITER_TURN_TO_TR
    assert r["itno"] == rh["itno"] and r["final_err"] == rh["final_err"]
    g, o = pair(P, O, sc)
    rg = g.hybrid()
    ro = o.hybrid()
    assert (rg["code"], rg["itno"]) == (ro["code"], ro["itno"])
    assert abs(rg["final_err"] - ro["final_err"]) <= 1e-9 * ro["final_err"]
