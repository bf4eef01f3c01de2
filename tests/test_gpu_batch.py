"""Batched solves (include/psba_batch.h) through the C-ABI.

* Each problem's result is the same alone (a batch of one) and inside a
  mixed batch, in any order, bit for bit: every reduction of the batched path
  has an order fixed by the problem alone, and stopped problems are skipped.
* Against the single-problem psba_levmar with the same rules (no handoff,
  relative-decrease stop, 100 iterations) and against the oracle's LM, the
  stop code and iteration count agree exactly and the final cost and
  parameters within 1e-9 relative (the reduction orders differ, the
  arithmetic does not), or within ten times the spread of the oracle's own
  solves from starts one ulp apart where that is wider (the free-scale,
  16-parameter scenes).
* Polling, autotuning and re-committing into a used arena change nothing.
* The mixed-precision batch reaches the FP64 batch's final cost within 1e-6
  relative.
"""
import numpy as np
import pytest

from conftest import rel

pytestmark = pytest.mark.gpu

SEED = 1003
OPTS = dict(max_iter=100, rel_tol=1e-8)


def window(P, k, **kw):
    return P.Scene(3, SEED, problem=k, **kw)


def solve_batch(P, scenes, precision=0, **opts):
    b = P.Batch(scenes[0].cnp, precision=precision)
    for sc in scenes:
        b.add_scene(sc)
    b.commit()
    b.solve(**(opts or OPTS))
    out = [b.result(s) for s in range(len(scenes))]
    b.close()
    return out


def ulp_cloud(O, sc, starts=4):
    """Spread of the oracle's converged LM (OPTS, no handoff) over starting
    points one ulp away from the scene's (each parameter moved up or down at
    random): the largest relative differences from the unperturbed solve in
    the cost, the cameras and the points.  Every start must stop at the same
    iteration with the same code."""
    rng = np.random.default_rng(17)
    out = []
    for t in range(starts + 1):
        cams, pts = sc.cams.copy(), sc.pts.copy()
        if t:
            cams = np.nextafter(cams, cams + rng.choice([-1.0, 1.0], cams.size))
            pts = np.nextafter(pts, pts + rng.choice([-1.0, 1.0], pts.size))
        o = O.Oracle(sc.cnp, sc.K, sc.q0, sc.meas, sc.iidx, sc.jidx, cams, pts, sc.fixmask)
        r = o.levmar(max_iter=OPTS["max_iter"], flags=O.FLAG_NO_HANDOFF, rel_tol=OPTS["rel_tol"])
        out.append((r, o.read_params()))
    (r0, (c0, p0)) = out[0]
    assert all((r["code"], r["itno"]) == (r0["code"], r0["itno"]) for r, _ in out)
    return (max(abs(r["final_err"] - r0["final_err"]) / r0["final_err"] for r, _ in out[1:]),
            max(rel(c, c0) for _, (c, _) in out[1:]), max(rel(p, p0) for _, (_, p) in out[1:]))


def same(a, b):
    assert a["code"] == b["code"]
    assert a["itno"] == b["itno"]
    assert a["init_err"] == b["init_err"]
    assert a["final_err"] == b["final_err"]
    assert np.array_equal(a["cams"], b["cams"])
    assert np.array_equal(a["pts"], b["pts"])


def test_alone_equals_in_mix(P, gpu):
    scenes = [window(P, k) for k in range(8)]
    mix = solve_batch(P, scenes)
    rev = solve_batch(P, scenes[::-1])[::-1]
    b = P.Batch(6)
    for k, sc in enumerate(scenes):
        b.reset()
        b.add_scene(sc)
        b.commit()  # re-commits into the same arena
        b.solve(**OPTS)
        alone = b.result(0)
        same(alone, mix[k])
        same(alone, rev[k])
        assert alone["final_err"] < alone["init_err"]
    b.close()


def test_matches_levmar_and_oracle(P, O, gpu):
    scenes = [window(P, k, n_pts=600) for k in (11, 12, 13)]
    res = solve_batch(P, scenes)
    for sc, r in zip(scenes, res):
        c = P.Context.from_scene(sc)
        g = c.levmar(max_iter=100, flags=P.FLAG_NO_HANDOFF, rel_tol=1e-8)
        o = O.Oracle.from_scene(sc).levmar(max_iter=100, flags=O.FLAG_NO_HANDOFF, rel_tol=1e-8)
        for ref in (g, o):
            assert r["code"] == ref["code"]
            assert r["itno"] == ref["itno"]
            assert abs(r["final_err"] - ref["final_err"]) <= 1e-9 * ref["final_err"]
        cams, pts = c.read_params()
        assert rel(r["cams"], cams) < 1e-9
        assert rel(r["pts"], pts) < 1e-9
        c.close()


@pytest.mark.parametrize("config,kw", [(1, dict(n_pts=400, n_cams=12)), (1, dict(cnp=16, n_pts=400, n_cams=10)),
                                       (2, dict(n_pts=1500, n_cams=12, outlier_frac=0.0))],
                         ids=["cnp11", "cnp16", "cnp9"])
def test_camera_models(P, O, gpu, config, kw):
    """Fixed budget (8 iterations, before any stopping test can fire) and
    convergence by the relative-decrease rule.  The reference's own |dp| test
    (|dp| < 1e-12 |p|) is left out: at convergence it compares rounding noise,
    so the iteration it fires on is not a property of the algorithm.

    Stop code and iteration count must agree exactly, cost and parameters
    within 1e-9 -- or within ten times the scene's own noise cloud where that
    is wider: with only camera 0 held, the scale of the reconstruction is
    free, and on the 16-parameter model the converged points move by up to
    1.6e-5 (cameras 6.6e-8) when the oracle starts one ulp away, at the same
    iteration count and cost (1e-15).  Any two correct FP64 solves land
    inside that cloud, so closer agreement would not be a property of the
    algorithm."""
    O.set_threads(16)
    scenes = [P.Scene(config, 77 + k, **kw) for k in range(3)]
    short = solve_batch(P, scenes, max_iter=8, rel_tol=0.0)
    full = solve_batch(P, scenes, **OPTS)
    for sc, r8, r in zip(scenes, short, full):
        c = P.Context.from_scene(sc)
        g8 = c.levmar(max_iter=8, flags=P.FLAG_NO_HANDOFF)
        cams, pts = c.read_params()
        c.close()
        assert (r8["itno"], r8["code"]) == (g8["itno"], g8["code"]) == (8, P.ITER_CONTINUE)
        assert abs(r8["final_err"] - g8["final_err"]) <= 1e-9 * g8["final_err"]
        assert rel(r8["cams"], cams) < 1e-9 and rel(r8["pts"], pts) < 1e-9
        c = P.Context.from_scene(sc)
        g = c.levmar(flags=P.FLAG_NO_HANDOFF, **OPTS)
        c_cams, c_pts = c.read_params()
        c.close()
        dc, dcam, dpt = ulp_cloud(O, sc)
        tc, tcam, tpt = (max(1e-9, 10 * x) for x in (dc, dcam, dpt))
        info = (r["itno"], g["itno"], r["code"], g["code"], r["final_err"], g["final_err"], tc, tcam, tpt)
        assert r["code"] == g["code"], info
        assert r["itno"] == g["itno"], info
        assert abs(r["final_err"] - g["final_err"]) <= tc * g["final_err"], info
        assert rel(r["cams"], c_cams) < tcam and rel(r["pts"], c_pts) < tpt, info


def test_poll_and_autotune(P, gpu):
    scenes = [window(P, k) for k in range(20, 26)]
    ref = solve_batch(P, scenes)
    b = P.Batch(6)
    for sc in scenes:
        b.add_scene(sc)
    b.commit()
    ob, pb = b.autotune()
    assert ob in (64, 128, 256) and pb in (64, 128, 256)
    b.launch(**OPTS)
    spins = 0
    while not b.poll():
        spins += 1
    ms, rounds = b.elapsed()
    assert ms > 0 and rounds >= max(r["itno"] for r in ref)
    for s in range(len(scenes)):
        same(b.result(s), ref[s])
    b.close()


def test_mixed_precision(P, gpu):
    scenes = [window(P, k) for k in range(30, 34)]
    ref = solve_batch(P, scenes)
    mix = solve_batch(P, scenes, precision=P.PREC_MIXED)
    for a, m in zip(ref, mix):
        assert m["code"] in (P.ITER_DP_NO_CHANGE, P.ITER_CONTINUE)
        assert abs(m["final_err"] - a["final_err"]) <= 1e-6 * a["final_err"]


def test_errors(P, gpu):
    b = P.Batch(6)
    with pytest.raises(P.PsbaError):
        b.launch()
    sc = window(P, 40)
    bad = sc.iidx.copy()
    bad[0] = sc.n_pts
    with pytest.raises(P.PsbaError):
        b.add(sc.K, sc.meas, sc.q0, sc.cams, sc.pts, bad, sc.jidx)
    b.add_scene(sc)
    with pytest.raises(P.PsbaError):
        b.result(0)
    b.commit()
    with pytest.raises(P.PsbaError):
        b.result(0)
    b.solve(max_iter=3, rel_tol=0.0)
    r = b.result(0)
    assert r["itno"] == 3 and r["code"] == P.ITER_CONTINUE
    with pytest.raises(P.PsbaError):
        b.result(1)
    b.close()


def test_mixed_sizes(P, gpu):
    """One-, two- and many-tile reduced systems in one batch: the batch then
    runs its dense solve phase by phase (one launch per phase over all
    problems), a problem alone with one or two tiles runs it in one workgroup;
    both do the same tile operations in the same order, so the results agree
    bit for bit.  The large problem also matches its single-problem solve."""
    scenes = [P.Scene(4, 1004, problem=1, size_class=0), P.Scene(3, SEED, problem=5, n_cams=15),
              P.Scene(4, 1004, problem=2, size_class=1, n_pts=2000),
              P.Scene(4, 1004, problem=3, size_class=2, n_pts=2500)]
    mix = solve_batch(P, scenes)
    for sc, r in zip(scenes, mix):
        same(solve_batch(P, [sc])[0], r)
    big = scenes[-1]
    c = P.Context.from_scene(big)
    g = c.levmar(flags=P.FLAG_NO_HANDOFF, **OPTS)
    c.close()
    assert (mix[-1]["code"], mix[-1]["itno"]) == (g["code"], g["itno"])
    assert abs(mix[-1]["final_err"] - g["final_err"]) <= 1e-9 * g["final_err"]


def test_config3_full_batch(P, O, gpu):
    """The stated batch: 1,024 windows (5-30 cameras, 200-4,000 points) in one
    batch.  Sixteen sampled windows solved alone give the same result bit for
    bit, and eight of them match the oracle's LM on stop code, iteration count
    and final cost within 1e-9."""
    scenes = [window(P, k) for k in range(1024)]
    b = P.Batch(6)
    for sc in scenes:
        b.add_scene(sc)
    b.commit()
    b.solve(**OPTS)
    rng = np.random.default_rng(5)
    sample = sorted(rng.choice(1024, 16, replace=False).tolist())
    res = {k: b.result(k) for k in sample}
    b.close()
    assert sum(res[k]["final_err"] < res[k]["init_err"] for k in sample) == 16
    one = P.Batch(6)
    for k in sample:
        one.reset()
        one.add_scene(scenes[k])
        one.commit()
        one.solve(**OPTS)
        same(one.result(0), res[k])
    one.close()
    O.set_threads(16)
    for k in sample[:8]:
        o = O.Oracle.from_scene(scenes[k]).levmar(max_iter=100, flags=O.FLAG_NO_HANDOFF, rel_tol=1e-8)
        r = res[k]
        assert (r["code"], r["itno"]) == (o["code"], o["itno"]), k
        assert abs(r["final_err"] - o["final_err"]) <= 1e-9 * o["final_err"], k


def test_eager_rounds_equal_graph_rounds(P, gpu, monkeypatch):
    """PSBA_BATCH_EAGER=1 (set when the batch is created) enqueues each round
    as plain launches instead of replaying the captured graphs: the same
    kernels and arguments, so the same results bit for bit."""
    scenes = [window(P, k) for k in range(50, 54)]
    graph = solve_batch(P, scenes)
    monkeypatch.setenv("PSBA_BATCH_EAGER", "1")
    eager = solve_batch(P, scenes)
    for a, b in zip(graph, eager):
        same(a, b)
