"""Robust loss (an extension; the reference minimises the plain sum of
squares) on the GPU against the oracle, through the C-ABI: the robust cost,
the reweighted Jacobian stages, and LM trial by trial with Huber and Cauchy
on scenes with gross outliers, up to the full 88-camera scene.

The robust solves converge slowly (the reweighted steps converge linearly)
and the robust cost with outliers has many nearby minima, so, as with the
plain cost on outlier scenes, two correct FP64 evaluations drift apart after
a few tens of iterations (first relative difference above 1e-9 at iterations
20-157 on these scenes, measured).  Parity is therefore asserted over the
first 15 iterations: every trial's decision and cost within 1e-9, the
parameters reached within 1e-8 (the points whose observations the loss
down-weights are the least determined and carry the largest rounding drift:
1.0e-9 after 15 Cauchy iterations on the 54-camera scene, measured)."""
import numpy as np
import pytest

from conftest import rel
from test_gpu_parity import compare_traces, pair

pytestmark = pytest.mark.gpu

LOSSES = [(1, "huber"), (2, "cauchy")]


@pytest.mark.parametrize("kind,name", LOSSES)
def test_robust_stages(P, O, gpu, kind, name):
    sc = P.Scene(2, 3100, n_pts=3000, n_cams=20, outlier_frac=0.03)
    g, o = pair(P, O, sc)
    g.set_loss(kind, 2.0)
    o.set_loss(kind, 2.0)
    assert rel(g.residual(), o.residual()) < 1e-13
    g.jacobian_normal(1.0, 1.0)
    o.jacobian_normal(1.0, 1.0)
    for b in ("EX", "JA", "JB", "U", "V", "W", "G"):
        assert rel(g.buffer(b), o.buffer(b)) < 1e-12, b


@pytest.mark.parametrize("kind,name", LOSSES)
@pytest.mark.parametrize("config,kw", [(0, dict(outlier_frac=0.05)), (1, dict(n_pts=1500, outlier_frac=0.03)),
                                       (2, dict(n_pts=6000, n_cams=30))], ids=["c0", "c1", "c2"])
def test_robust_levmar_trial_by_trial(P, O, gpu, kind, name, config, kw):
    sc = P.Scene(config, 3200 + config, **kw)
    g, o = pair(P, O, sc)
    g.set_loss(kind, 2.0)
    o.set_loss(kind, 2.0)
    rg = g.levmar(max_iter=15, flags=P.FLAG_NO_HANDOFF)
    ro = o.levmar(max_iter=15, flags=O.FLAG_NO_HANDOFF)
    assert (rg["code"], rg["itno"]) == (ro["code"], ro["itno"])
    compare_traces(rg["trace"], ro["trace"])
    gc, gp = g.read_params()
    oc, op = o.read_params()
    assert rel(gc, oc) < 1e-8 and rel(gp, op) < 1e-8


def test_robust_full_size(P, O, gpu):
    """The 88-camera / 64,298-point scene with its 1% outliers, Huber loss."""
    sc = P.Scene(2, 1002)
    g, o = pair(P, O, sc)
    g.set_loss(1, 3.0)
    o.set_loss(1, 3.0)
    rg = g.levmar(max_iter=15, flags=P.FLAG_NO_HANDOFF)
    ro = o.levmar(max_iter=15, flags=O.FLAG_NO_HANDOFF)
    assert (rg["code"], rg["itno"]) == (ro["code"], ro["itno"])
    compare_traces(rg["trace"], ro["trace"])


def test_loss_argument_checks(P, gpu):
    sc = P.Scene(0, 1)
    c = P.Context.from_scene(sc)
    with pytest.raises(P.PsbaError):
        c.set_loss(7, 1.0)
    with pytest.raises(P.PsbaError):
        c.set_loss(1, 0.0)
    c.set_loss(0, 0.0)  # NONE ignores the scale
