"""The sba text reader (psba_read_sba / readInitialSBAEstimate, the reference's
PSBA/readparams.cpp:444-541) on the reference's own data files and on small
files written here: counts, projections in ascending frame order per point,
visibility mask, the quat2vec camera filter, covariance forms, the error
paths, and the writer of solved parameters (psba_write_sba)."""
import os

import numpy as np
import pytest

from conftest import GOLDEN
import sbafmt

CASES = [("7camsvarK.txt", "3pts.txt"), ("7camsvarK.txt", "5pts.txt"), ("7camsvarK.txt", "7pts.txt"),
         ("9camsvarK.txt", "9pts.txt"), ("54camsvarK.txt", "54pts.txt"),
         ("Trafalgar-21-11315-cams.txt", "Trafalgar-21-11315-pts.txt")]


def quat_vec(q):
    q = np.asarray(q, float)
    return np.sign(q[0] if q[0] != 0 else 1.0) * q[1:] / np.linalg.norm(q)


@pytest.mark.parametrize("cams_name,pts_name", CASES)
def test_reads_reference_files(P, cams_name, pts_name):
    d = P.read_sba(os.path.join(GOLDEN, cams_name), os.path.join(GOLDEN, pts_name))
    rows = sbafmt.read_cams(os.path.join(GOLDEN, cams_name))
    X, tracks = sbafmt.read_points(os.path.join(GOLDEN, pts_name))
    assert d["n_cams"] == len(rows) and d["n_pts"] == len(tracks)
    assert d["n_obs"] == sum(len(t) for t in tracks)
    cams = d["cams"].reshape(-1, 11)
    for j, r in enumerate(rows):
        assert np.array_equal(cams[j, :5], r[:5])
        assert np.allclose(cams[j, 5:8], quat_vec(r[5:9]), rtol=0, atol=1e-15)
        assert np.array_equal(cams[j, 8:], r[9:12])
        q = d["initrot"][4 * j:4 * j + 4]
        assert abs(np.linalg.norm(q) - 1) < 1e-12 and q[0] >= 0
        assert np.allclose(q[1:], cams[j, 5:8], atol=0)
    assert np.array_equal(d["pts"].reshape(-1, 3), X)
    # projections in ascending frame order per point (the shipped files are
    # already sorted); the mask marks every (point, frame)
    flat = np.array([[x, y] for fr in tracks for _, x, y in sorted(fr, key=lambda r: r[0])]).reshape(-1)
    assert np.array_equal(d["imgpts"], flat)
    assert np.array_equal(d["vmask"], sbafmt.vmask_of(tracks, len(rows)))
    assert d["cov"] is None


def test_quat2vec_matches_formula(P):
    import ctypes as ct
    L = P.lib()
    rng = np.random.default_rng(5)
    for _ in range(20):
        q = rng.standard_normal(4) * rng.uniform(0.5, 2.0)
        row = np.concatenate([rng.standard_normal(5), q, rng.standard_normal(3)])
        out = np.empty(11)
        L.psba_quat2vec(row.ctypes.data_as(ct.POINTER(ct.c_double)), 12, out.ctypes.data_as(ct.POINTER(ct.c_double)), 11)
        assert np.array_equal(out[:5], row[:5])
        assert np.allclose(out[5:8], quat_vec(q), rtol=1e-15, atol=1e-16)
        assert np.array_equal(out[8:], row[9:])


def write(tmp_path, name, text):
    p = tmp_path / name
    p.write_text(text)
    return str(p)


CAMS6 = "# pose only\n0 0 0 0 0 0\n0.01 0 0 1 0 0\n\n0 0.02 0 0 1 0\n"


def test_pose_only_without_filter(P, tmp_path):
    c = write(tmp_path, "c.txt", CAMS6)
    p = write(tmp_path, "p.txt", "1 2 3 2 0 10 11 2 12 13\n4 5 6  2 1 14 15 2 16 17\n")
    d = P.read_sba(c, p, cnp=6, quat_filter=False)
    assert d["n_cams"] == 3 and d["n_pts"] == 2 and d["n_obs"] == 4
    assert np.array_equal(d["cams"], np.array([0, 0, 0, 0, 0, 0, .01, 0, 0, 1, 0, 0, 0, .02, 0, 0, 1, 0.]))
    assert np.array_equal(d["imgpts"], np.arange(10, 18, dtype=float))
    assert np.array_equal(d["vmask"], np.array([[1, 0, 1], [0, 1, 1]], np.int8))
    assert d["initrot"][4] == pytest.approx(np.sqrt(1 - 1e-4), abs=0)


def test_full_and_triangular_covariance(P, tmp_path):
    c = write(tmp_path, "c.txt", CAMS6)
    full = write(tmp_path, "pf.txt", "1 2 3 2 0 10 11 4 1 1 9 2 12 13 1 0 0 1\n4 5 6 1 1 14 15 2 -1 -1 3\n")
    d = P.read_sba(c, full, cnp=6, quat_filter=False)
    assert np.array_equal(d["cov"].reshape(-1, 2, 2),
                          np.array([[[4, 1], [1, 9]], [[1, 0], [0, 1]], [[2, -1], [-1, 3]]], float))
    tri = write(tmp_path, "pt.txt", "1 2 3 2 0 10 11 4 1 9 2 12 13 1 0 1\n4 5 6 1 1 14 15 2 -1 3\n")
    d2 = P.read_sba(c, tri, cnp=6, quat_filter=False)
    assert np.array_equal(d2["cov"], d["cov"])
    assert np.array_equal(d2["imgpts"], d["imgpts"])


@pytest.mark.parametrize("pts,msg", [
    ("1 2 3 1 5 10 11\n", "frame 5"),             # frame beyond the camera count
    ("1 2 3 2 0 10 11\n", "fewer projections"),   # frame count larger than the list
    ("1 2 3 1 0 10 11 7\n", "more values"),       # trailing values
    ("1 2 3 2 0 10 11 0 12 13\n", "twice"),        # same frame twice
    ("1 2 x\n", "coordinates"),
])
def test_malformed_point_files_are_rejected(P, tmp_path, pts, msg):
    c = write(tmp_path, "c.txt", CAMS6)
    p = write(tmp_path, "p.txt", pts)
    with pytest.raises(P.PsbaError, match=msg):
        P.read_sba(c, p, cnp=6, quat_filter=False)


def test_camera_line_length_and_missing_files(P, tmp_path):
    c = write(tmp_path, "c.txt", "0 0 0 0 0\n")
    p = write(tmp_path, "p.txt", "1 2 3 1 0 10 11\n")
    with pytest.raises(P.PsbaError, match="expected 6"):
        P.read_sba(c, p, cnp=6, quat_filter=False)
    with pytest.raises(P.PsbaError, match="cannot open"):
        P.read_sba(str(tmp_path / "none.txt"), p, cnp=6, quat_filter=False)


def test_unsorted_frames_are_put_in_camera_order(P, tmp_path):
    """generate_idxs numbers a point's observations camera-ascending
    (PSBA/misc.cpp:189-216); a file listing frames out of order must not
    shift projections onto the wrong cameras (covariances move with them)."""
    c = write(tmp_path, "c.txt", CAMS6)
    p = write(tmp_path, "p.txt", "1 2 3 3  2 20 21 5 0 0 6   0 10 11 1 0 0 1   1 14 15 2 0 0 3\n"
                                 "4 5 6 2  1 30 31 7 0 0 8   0 40 41 9 0 0 10\n")
    d = P.read_sba(c, p, cnp=6, quat_filter=False)
    assert np.array_equal(d["imgpts"], np.array([10, 11, 14, 15, 20, 21, 40, 41, 30, 31], float))
    cov = d["cov"].reshape(-1, 4)
    assert np.array_equal(cov[:, 0], np.array([1, 2, 5, 9, 7], float))
    assert np.array_equal(d["vmask"], np.array([[1, 1, 1], [1, 1, 0]], np.int8))
    # the sorted file reads identically
    p2 = write(tmp_path, "p2.txt", "1 2 3 3  0 10 11 1 0 0 1  1 14 15 2 0 0 3  2 20 21 5 0 0 6\n"
                                  "4 5 6 2  0 40 41 9 0 0 10  1 30 31 7 0 0 8\n")
    d2 = P.read_sba(c, p2, cnp=6, quat_filter=False)
    for k in ("imgpts", "cov", "vmask", "pts", "cams"):
        assert np.array_equal(d[k], d2[k])


def _qmul(a, b):
    s1, v1 = a[0], np.asarray(a[1:])
    s2, v2 = b[0], np.asarray(b[1:])
    return np.concatenate([[s1 * s2 - v1 @ v2], s1 * v2 + s2 * v1 + np.cross(v1, v2)])


@pytest.mark.parametrize("with_k", [True, False])
def test_writer_round_trip(P, tmp_path, with_k):
    """Solved parameters (local rotation v relative to initrot) written and read
    back: the read rotation is q_v (x) q0 (independently composed here), every
    other value and every projection comes back as the same double."""
    src = P.read_sba(os.path.join(GOLDEN, "7camsvarK.txt"), os.path.join(GOLDEN, "7pts.txt"))
    nc = src["n_cams"]
    rng = np.random.default_rng(11)
    c11 = src["cams"].reshape(nc, 11)
    K = c11[:, :5].copy()
    solved = c11[:, 5:].copy()  # cnp 6: [v t]
    solved[:, :3] = rng.uniform(-0.05, 0.05, (nc, 3))
    solved[:, 3:] += rng.standard_normal((nc, 3)) * 0.1
    pts = src["pts"] + rng.standard_normal(src["pts"].shape) * 0.01
    cp, pp = str(tmp_path / "c.txt"), str(tmp_path / "p.txt")
    P.write_sba(cp, pp, 6, solved.reshape(-1), pts, src["initrot"], src["imgpts"], src["vmask"],
                K=K.reshape(-1) if with_k else None)
    back = P.read_sba(cp, pp, cnp=11 if with_k else 6)
    cb = back["cams"].reshape(nc, -1)
    for j in range(nc):
        v = solved[j, :3]
        qv = np.concatenate([[np.sqrt(1 - v @ v)], v])
        q = _qmul(qv, src["initrot"][4 * j:4 * j + 4])
        q = q / np.linalg.norm(q) * np.sign(q[0])
        assert np.allclose(back["initrot"][4 * j:4 * j + 4], q, rtol=0, atol=2e-15)
        assert np.array_equal(cb[j, -3:], solved[j, 3:])
        if with_k:
            assert np.array_equal(cb[j, :5], K[j])
    assert np.array_equal(back["pts"], pts)
    assert np.array_equal(back["imgpts"], src["imgpts"])
    assert np.array_equal(back["vmask"], src["vmask"])
