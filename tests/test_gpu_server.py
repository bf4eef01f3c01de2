"""The request loop (include/psba_server.h) through the C-ABI.

* Each request's result is the same alone (a batch of one) and in the mix,
  bit for bit, with requests of all three size classes arriving while others
  are on the device.
* Requests overlap: small requests submitted after a large one complete
  before it (continuous batching, not a queue).
* Admission keeps the device memory held within the HBM budget; a request
  larger than the whole budget is rejected, the rest still complete.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SEED = 1004


def req(P, k, size_class=0, **kw):
    return P.Scene(4, SEED, problem=k, size_class=size_class, **kw)


def alone(P, sc, precision=0):
    b = P.Batch(sc.cnp, precision=precision)
    b.add_scene(sc)
    b.commit()
    b.solve(max_iter=100, rel_tol=1e-8)
    r = b.result(0)
    b.close()
    return r


def same(a, b):
    assert (a["code"], a["itno"]) == (b["code"], b["itno"])
    assert a["init_err"] == b["init_err"] and a["final_err"] == b["final_err"]
    assert np.array_equal(a["cams"], b["cams"]) and np.array_equal(a["pts"], b["pts"])


def collect(P, srv, ids, out):
    for rid in ids:
        if rid not in out:
            r = srv.result(rid)
            if r is not None:
                out[rid] = r


@pytest.mark.parametrize("precision", [0, 1], ids=["fp64", "mixed"])
def test_same_alone_and_in_mix(P, gpu, precision):
    scenes = [req(P, k) for k in range(10)]
    scenes += [req(P, 100 + k, 1, n_pts=3000) for k in range(2)]
    scenes += [req(P, 200, 2, n_pts=3000)]
    order = [0, 10, 1, 2, 12, 3, 4, 11, 5, 6, 7, 8, 9]
    srv = P.Server(6, precision=precision, lanes=3, max_batch=4)
    ids, out = {}, {}
    for n, k in enumerate(order):
        ids[k] = srv.submit_scene(scenes[k])
        if n % 3 == 2:  # arrivals interleaved with work already on the device
            srv.step()
            collect(P, srv, ids.values(), out)
    srv.drain()
    collect(P, srv, ids.values(), out)
    st = srv.stats()
    srv.close()
    assert st["completed"] == len(scenes) and st["rejected"] == 0 and st["batches"] >= 4
    for k, sc in enumerate(scenes):
        r = out[ids[k]]
        assert r["t_done"] >= r["t_submit"]
        same(r, alone(P, sc, precision))


def test_small_requests_overtake_a_large_one(P, gpu):
    big = req(P, 300, 2)
    small = [req(P, 310 + k) for k in range(6)]
    srv = P.Server(6, lanes=2, max_batch=8)
    bid = srv.submit_scene(big)
    srv.step()  # the large request is on the device
    sids = [srv.submit_scene(sc) for sc in small]
    out = {}
    order = []
    while len(out) < 7:
        srv.step()
        for rid in [bid] + sids:
            if rid not in out:
                r = srv.result(rid, params=False)
                if r is not None:
                    out[rid] = r
                    order.append(rid)
    srv.close()
    assert order[-1] == bid
    assert max(out[s]["t_done"] for s in sids) < out[bid]["t_done"]


def test_budget(P, gpu):
    small = [req(P, 400 + k) for k in range(6)]
    one = max(P.problem_bytes(6, sc.n_cams, sc.n_pts, sc.n_obs, P.pair_items(sc.iidx)) for sc in small)
    budget = int(1.5 * one)
    huge = req(P, 410, 1)
    srv = P.Server(6, lanes=4, max_batch=8, hbm_budget=budget)
    ids = [srv.submit_scene(sc) for sc in small[:3]] + [srv.submit_scene(huge)] + \
        [srv.submit_scene(sc) for sc in small[3:]]
    srv.drain()
    st = srv.stats()
    res = [srv.result(i) for i in ids]
    srv.close()
    assert st["bytes_peak"] <= budget
    assert st["completed"] == 6 and st["rejected"] == 1 and st["batches"] == 6
    assert res[3]["code"] == -3  # PSBA_ERR_NOMEM
    for sc, r in zip(small, res[:3] + res[4:]):
        same(r, alone(P, sc))


@pytest.mark.parametrize("precision", [0, 1], ids=["fp64", "mixed"])
def test_full_size_classes_alone_and_in_mix(P, gpu, precision):
    """The stream's medium (50 cameras, 20,000 points) and large (138 cameras,
    19,878 points, heavy-tailed tracks) requests at their stated sizes,
    arriving among small ones while those run: each result equals the same
    request solved alone, bit for bit."""
    scenes = [req(P, 500 + k) for k in range(6)] + [req(P, 600, 1), req(P, 601, 1), req(P, 700, 2)]
    assert (scenes[6].n_cams, scenes[6].n_pts) == (50, 20000)
    assert (scenes[8].n_cams, scenes[8].n_pts) == (138, 19878)
    srv = P.Server(6, precision=precision, lanes=3, max_batch=4)
    ids, out = {}, {}
    for n, k in enumerate([0, 6, 1, 8, 2, 3, 7, 4, 5]):
        ids[k] = srv.submit_scene(scenes[k])
        if n % 2 == 1:
            srv.step()
            collect(P, srv, ids.values(), out)
    srv.drain()
    collect(P, srv, ids.values(), out)
    st = srv.stats()
    srv.close()
    assert st["completed"] == len(scenes) and st["rejected"] == 0
    for k, sc in enumerate(scenes):
        same(out[ids[k]], alone(P, sc, precision))
