"""The hybrid driver's report (PSBA/main.cpp:209-225 and additions) through
the C-ABI: the same solve as psba_hybrid, the reference's error measure
sqrt(SSE)/n_obs, the RMS per observation, the LM / dogleg split and the
device time per stage."""
import math

import pytest

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("config,kw", [(0, {}), (1, dict(n_pts=600))], ids=["c0", "c1"])
def test_hybrid_report(P, gpu, config, kw):
    sc = P.Scene(config, 4100 + config, **kw)
    a = P.Context.from_scene(sc)
    rep = a.hybrid_report()
    b = P.Context.from_scene(sc)
    ref = b.hybrid()
    assert (rep["code"], rep["itno"]) == (ref["code"], ref["itno"])
    assert rep["final_sse"] == ref["final_err"] and rep["init_sse"] == ref["init_err"]
    n = sc.n_obs
    assert rep["final_error_ref"] == pytest.approx(math.sqrt(rep["final_sse"]) / n, rel=1e-15)
    assert rep["final_rms_px"] == pytest.approx(math.sqrt(rep["final_sse"] / n), rel=1e-15)
    assert rep["init_rms_px"] > rep["final_rms_px"]
    # iterations entered; a loop that stops inside an iteration (before the
    # shared counter moves on) adds one to the count
    assert rep["lm_iterations"] + rep["tr_iterations"] in (rep["itno"], rep["itno"] + 1)
    assert rep["lm_trials"] + rep["tr_steps"] == len(ref["trace"])
    assert rep["handoffs"] == sum(1 for x, y in zip(ref["trace"], ref["trace"][1:]) if x.kind != y.kind)
    st = rep["stages"]
    assert st["jacobians"] >= 1 and st["trials"] >= 1
    assert st["jacobian_normal"] > 0 and st["schur"] > 0 and st["factor"] > 0 and rep["wall_ms"] > 0
