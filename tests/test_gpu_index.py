"""The index built on the device (psba_set_index_device,
psba_generate_idxs_device) against the host construction, bit for bit:
every index array, the scene statistics, and whole LM traces run on both."""
import os

import numpy as np
import pytest

from conftest import GOLDEN
import sbafmt

pytestmark = pytest.mark.gpu

SCENES = [(0, {}), (1, dict(n_pts=800)), (1, dict(cnp=16, n_pts=500)), (2, dict(n_pts=4000, n_cams=30)),
          (4, dict(size_class=2, n_pts=3000))]
IDS = ["c0", "c1-cnp11", "c1-cnp16", "c2-small", "c4-large"]


def ragged(P):
    """A scene with a camera that sees nothing and points without observations."""
    sc = P.Scene(0, 1234, n_pts=300)
    keep = sc.jidx != 3
    keep &= ~np.isin(sc.iidx, [0, 7, 299])
    return sc, keep


def build_pair(P, sc, iidx, jidx):
    a = P.Context(sc.cnp, sc.n_cams, sc.n_pts, int(iidx.size))
    b = P.Context(sc.cnp, sc.n_cams, sc.n_pts, int(iidx.size))
    a.set_index(iidx, jidx)
    b.set_index_device(iidx, jidx)
    return a, b


def same_index(a, b):
    for name in P_INDEX:
        x, y = a.read_index(name), b.read_index(name)
        assert x.shape == y.shape and np.array_equal(x, y), name
    ia, ib = a.info(), b.info()
    for k in ("n_pairs", "n_pair_items", "sum_t2"):
        assert ia[k] == ib[k], k


P_INDEX = ["OBS_PT", "OBS_CAM", "PT_PTR", "CAM_PTR", "CAM_OBS", "CCHUNK_BEGIN", "CCHUNK_END", "CAM_CCHUNK_PTR",
           "PAIR_ITEMS", "PAIR_K", "PAIR_L", "SCHUNK_BEGIN", "SCHUNK_END", "PAIR_SCHUNK_PTR", "LBLOCK_PAIR",
           "SCHUNK_ORDER", "PT_ORDER"]


@pytest.mark.parametrize("config,kw", SCENES, ids=IDS)
def test_device_index_bit_exact(P, gpu, config, kw):
    sc = P.Scene(config, 2000 + config, **kw)
    a, b = build_pair(P, sc, sc.iidx, sc.jidx)
    same_index(a, b)
    # the same solve on both
    for c in (a, b):
        c.upload(sc.K, sc.meas, sc.q0, sc.cams, sc.pts)
        c.set_fixed(sc.fixmask)
    ra = a.levmar(max_iter=8, flags=P.FLAG_NO_HANDOFF)
    rb = b.levmar(max_iter=8, flags=P.FLAG_NO_HANDOFF)
    assert [(t.cost, t.damp, t.accepted) for t in ra["trace"]] == [(t.cost, t.damp, t.accepted) for t in rb["trace"]]
    ca, pa = a.read_params()
    cb, pb = b.read_params()
    assert np.array_equal(ca, cb) and np.array_equal(pa, pb)


@pytest.mark.parametrize("config,kw", SCENES, ids=IDS)
def test_point_order(P, gpu, config, kw):
    """The back substitution's point order: a permutation of the points by
    descending track length, ties in ascending point id (numpy's stable sort
    of the negated lengths gives the same sequence)."""
    sc = P.Scene(config, 2100 + config, **kw)
    a, _ = build_pair(P, sc, sc.iidx, sc.jidx)
    ptr = a.read_index("PT_PTR")
    want = np.argsort(-np.diff(ptr), kind="stable")
    assert np.array_equal(a.read_index("PT_ORDER"), want)


def test_device_index_ragged(P, gpu):
    sc, keep = ragged(P)
    a, b = build_pair(P, sc, sc.iidx[keep], sc.jidx[keep])
    same_index(a, b)
    assert b.read_index("CAM_PTR")[3] == b.read_index("CAM_PTR")[4]  # camera 3 sees nothing


def test_device_index_rejects_bad_order(P, gpu):
    sc = P.Scene(0, 99)
    c = P.Context(sc.cnp, sc.n_cams, sc.n_pts, sc.n_obs)
    j = sc.jidx.copy()
    j[[0, 1]] = j[[1, 0]]
    with pytest.raises(P.PsbaError):
        c.set_index_device(sc.iidx, j)


@pytest.mark.parametrize("pts_name,n_cams", [("3pts.txt", 3), ("7pts.txt", 7), ("9pts.txt", 9)])
def test_generate_idxs_device_reference_data(P, gpu, pts_name, n_cams):
    _, tracks = sbafmt.read_points(os.path.join(GOLDEN, pts_name))
    vm = sbafmt.vmask_of(tracks, n_cams)
    n_pts = vm.shape[0]
    n_obs = int(vm.astype(bool).sum())
    h = P.generate_idxs(n_cams, n_pts, n_obs, vm)
    d = P.generate_idxs_device(n_cams, n_pts, n_obs, vm)
    for x, y in zip(h[:4], d):
        assert np.array_equal(x, y)


def test_generate_idxs_device_scene(P, gpu):
    sc = P.Scene(2, 1002, n_pts=20000)
    vm = np.zeros((sc.n_pts, sc.n_cams), np.int8)
    vm[sc.iidx, sc.jidx] = 1
    h = P.generate_idxs(sc.n_cams, sc.n_pts, sc.n_obs, vm)
    d = P.generate_idxs_device(sc.n_cams, sc.n_pts, sc.n_obs, vm)
    for x, y in zip(h[:4], d):
        assert np.array_equal(x, y)
    assert np.array_equal(d[0], sc.iidx) and np.array_equal(d[1], sc.jidx)
