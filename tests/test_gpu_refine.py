"""FP64 iterative refinement of the mixed path's step (psba_stage_refine,
psba_opts.refine_steps) against the plain FP64 path through the C-ABI.

The mixed path's step solves the damped normal equations only to the
accuracy of its FP32 blocks.  Each refinement sweep recomputes the residual
of those equations in FP64 and solves for a correction with the same blocks
and factor; the error contracts by about the FP32 solve's relative accuracy
per sweep, so after two sweeps the step agrees with the FP64 step to FP64
rounding, and an LM solve with refinement follows the FP64 solve trial by
trial."""
import numpy as np
import pytest

from conftest import rel
from test_gpu_parity import compare_traces

pytestmark = pytest.mark.gpu


def step(P, sc, precision, mu, sweeps=None):
    c = P.Context.from_scene(sc, precision=precision)
    c.jacobian_normal(1.0, 1.0)
    c.schur(mu)
    assert c.factor_solve()
    c.backsub()
    out = [c.buffer("DP").copy()]
    for _ in range(sweeps or 0):
        c.refine(mu, 1)
        out.append(c.buffer("DP").copy())
    return out, c


@pytest.mark.parametrize("config,kw", [(1, dict(n_pts=1500)), (2, dict(n_pts=6000, n_cams=30, outlier_frac=0.0))],
                         ids=["c1", "c2"])
def test_refinement_recovers_the_fp64_step(P, gpu, config, kw):
    sc = P.Scene(config, 4300 + config, **kw)
    c64 = P.Context.from_scene(sc)
    c64.jacobian_normal(1.0, 1.0)
    mu = 1e-3 * float(np.max(np.abs(np.concatenate([c64.buffer("U").reshape(-1, sc.cnp, sc.cnp).diagonal(
        axis1=1, axis2=2).ravel(), c64.buffer("V").reshape(-1, 3, 3).diagonal(axis1=1, axis2=2).ravel()]))))
    (dp64,), _ = step(P, sc, P.PREC_FP64, mu)
    dps, cm = step(P, sc, P.PREC_MIXED, mu, sweeps=3)
    errs = [rel(d, dp64) for d in dps]
    assert errs[0] > 1e-9           # the FP32 blocks show in the plain mixed step
    assert errs[1] < 1e-2 * errs[0]  # one sweep contracts the error
    assert errs[2] < 1e-11 and errs[3] < 1e-11
    # the refinement leaves the FP64 gradient J^T e in place of the FP32 one
    c64b = P.Context.from_scene(sc)
    c64b.jacobian_normal(1.0, 1.0)
    assert rel(cm.buffer("G"), c64b.buffer("G")) < 1e-12


def test_refined_lm_follows_fp64(P, gpu):
    """The 88-camera geometry without outliers: with one refinement sweep per
    step the mixed solve follows the FP64 solve trial by trial and ends at the
    same point.
    (Where the FP32 blocks are too coarse for the conditioning -- small damping
    on a nearly singular reduced system -- the corrections stop contracting
    and the step falls back to the unrefined one; the solve then behaves like
    the plain mixed one.)"""
    sc = P.Scene(2, 1002, outlier_frac=0.0)
    g64 = P.Context.from_scene(sc).levmar(max_iter=100, flags=P.FLAG_NO_HANDOFF, rel_tol=1e-10)
    gm = P.Context.from_scene(sc, precision=P.PREC_MIXED).levmar(max_iter=100, flags=P.FLAG_NO_HANDOFF,
                                                                 rel_tol=1e-10, refine_steps=1)
    assert (gm["code"], gm["itno"]) == (g64["code"], g64["itno"])
    # At the last, tiny dampings the FP32-assembled reduced matrix can lose
    # definiteness: that trial's factorisation fails, the damping is raised and
    # the mixed solve takes one more trial.  Trial by trial up to there, and the
    # same end point.
    tg, to = gm["trace"], g64["trace"]
    k = next((q for q, t in enumerate(tg) if not t.factor_ok), len(tg))
    assert k >= len(to) - 3, "mixed factorisation failed before the last iterations (trial %d)" % k
    compare_traces(tg[:k], to[:k], tol=1e-9, damp_tol=1e-6)
    if k == len(tg):
        compare_traces(tg, to, tol=1e-9, damp_tol=1e-6)
    assert abs(gm["final_err"] - g64["final_err"]) <= 1e-12 * g64["final_err"]


def test_refine_rejects_fp64_context(P, gpu):
    sc = P.Scene(0, 4500)
    c = P.Context.from_scene(sc)
    with pytest.raises(P.PsbaError):
        c.refine(1.0, 1)
