"""Integer indexing of PSBA/misc.cpp:178-218 (generate_idxs): observation
order, blk_idx, camera-pair shared-point lists and counts.  The oracle's
restatement and the product's host implementation (psba_generate_idxs,
C-ABI) must both equal a brute-force construction, bit for bit, on the
reference's own data files and on generated scenes."""
import os

import numpy as np
import pytest

from conftest import GOLDEN
import sbafmt


def brute(vm):
    n_pts, n_cams = vm.shape
    iidx, jidx = [], []
    blk = -np.ones((n_pts, n_cams), np.int32)
    for i in range(n_pts):
        for j in range(n_cams):
            if vm[i, j]:
                blk[i, j] = len(iidx)
                iidx.append(i)
                jidx.append(j)
    cnt = np.zeros((n_cams, n_cams), np.int32)
    lists = {}
    for j in range(n_cams):
        for k in range(n_cams):
            common = [i for i in range(n_pts) if vm[i, j] and vm[i, k]]
            cnt[j, k] = len(common)
            lists[(j, k)] = common
    return np.array(iidx, np.int32), np.array(jidx, np.int32), blk, cnt, lists


CASES = [("3pts.txt", "7camsvarK.txt", 3), ("5pts.txt", "7camsvarK.txt", 5), ("7pts.txt", "7camsvarK.txt", 7),
         ("9pts.txt", "9camsvarK.txt", 9)]


def load_case(pts_name, n_cams):
    _, tracks = sbafmt.read_points(os.path.join(GOLDEN, pts_name))
    return sbafmt.vmask_of(tracks, n_cams)


@pytest.mark.parametrize("pts_name,cams_name,n_cams", CASES)
def test_oracle_generate_idxs_reference_data(O, pts_name, cams_name, n_cams):
    vm = load_case(pts_name, n_cams)
    n_pts = vm.shape[0]
    iidx, jidx, blk, cnt, comm = O.generate_idxs(n_cams, n_pts, vm)
    bi, bj, bb, bc, lists = brute(vm)
    assert np.array_equal(iidx, bi) and np.array_equal(jidx, bj)
    assert np.array_equal(blk.reshape(n_pts, n_cams), bb)
    assert np.array_equal(cnt.reshape(n_cams, n_cams), bc)
    comm = comm.reshape(n_cams, n_cams, n_pts)
    for (j, k), common in lists.items():
        assert list(comm[j, k, :len(common)]) == common
        assert np.all(comm[j, k, len(common):] == -7)  # untouched tail


@pytest.mark.parametrize("pts_name,cams_name,n_cams", CASES)
def test_product_generate_idxs_bit_exact(P, O, pts_name, cams_name, n_cams):
    vm = load_case(pts_name, n_cams)
    n_pts = vm.shape[0]
    a = O.generate_idxs(n_cams, n_pts, vm)
    b = P.generate_idxs(n_cams, n_pts, int(vm.sum()), vm, with_comm=True)
    for x, y in zip(a, b):
        assert np.array_equal(x, y)


def test_reference_files_are_camera_ascending():
    """generate_idxs assigns observation numbers camera-ascending within a
    point while the reader keeps file order; the shipped files agree."""
    for name in ("3pts.txt", "5pts.txt", "7pts.txt", "9pts.txt"):
        _, tracks = sbafmt.read_points(os.path.join(GOLDEN, name))
        for fr in tracks:
            js = [j for j, _, _ in fr]
            assert js == sorted(js) and len(set(js)) == len(js)


@pytest.mark.parametrize("config,kw", [(0, {}), (1, dict(n_pts=300)), (3, dict(problem=5)),
                                       (4, dict(size_class=0, problem=2))])
def test_scene_observation_order_is_generate_idxs_order(P, O, config, kw):
    sc = P.Scene(config, 1234, **kw)
    vm = np.zeros((sc.n_pts, sc.n_cams), np.int8)
    vm[sc.iidx, sc.jidx] = 1
    iidx, jidx, _, _, _ = O.generate_idxs(sc.n_cams, sc.n_pts, vm, with_comm=False)
    assert np.array_equal(iidx, sc.iidx) and np.array_equal(jidx, sc.jidx)


def test_product_generate_idxs_rejects_count_mismatch(P):
    vm = np.ones((4, 3), np.int8)
    with pytest.raises(P.PsbaError):
        P.generate_idxs(3, 4, 5, vm)
