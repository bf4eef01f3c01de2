"""The drop-in boundary end to end on the device, and the full-size sparse
scene.

* tests/dropin/main_dropin follows PSBA/main.cpp:88-208 against the
  reference's headers (readInitialSBAEstimate, setup_cl, fill_initBuffer2,
  generate_idxs, fill_idxBuffer, levmar <-> trust_region, read_params) on the
  reference's own data files; its result must equal the C-ABI hybrid driver
  bit for bit and reach the oracle's minimum.
* The 88-camera / 64,298-point scene: one full LM iteration's stages against
  the oracle at full size, and size-independent properties of a whole solve
  (determinism, monotone accepted costs).
"""
import os
import subprocess

import numpy as np
import pytest

from conftest import GOLDEN, ROOT, rel
from test_gpu_parity import compare_traces

pytestmark = pytest.mark.gpu

DROPIN = os.path.join(ROOT, "tests", "dropin", "main_dropin")


def load_reference_problem(P, cams_name, pts_name):
    """What PSBA/main.cpp:102-150 hands to the drivers: cnp 6, K split off,
    the vector part of the pose cleared (the rotation lives in initrot)."""
    d = P.read_sba(os.path.join(GOLDEN, cams_name), os.path.join(GOLDEN, pts_name))
    nc = d["n_cams"]
    c11 = d["cams"].reshape(nc, 11)
    K = c11[:, :5].reshape(-1).copy()
    cams = c11[:, 5:].copy()
    cams[:, :3] = 0.0
    vm = d["vmask"]
    iidx, jidx = np.nonzero(vm)  # point-major, camera-ascending (generate_idxs order)
    return dict(K=K, cams=cams.reshape(-1), pts=d["pts"], q0=d["initrot"], meas=d["imgpts"],
                iidx=iidx.astype(np.int32), jidx=jidx.astype(np.int32), n_cams=nc, n_pts=d["n_pts"],
                n_obs=d["n_obs"])


def run_dropin(cams, pts, tmp_path):
    out = str(tmp_path / "params.bin")
    r = subprocess.run([DROPIN, cams, pts, out], stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stdout
    vals = dict(l.split("=", 1) for l in r.stdout.splitlines() if "=" in l)
    # the reference's own report lines (PSBA/main.cpp:210-218)
    rep = dict(l.split(":", 1) for l in r.stdout.splitlines() if l.startswith(("initial error", "final error",
                                                                                "total iteration")))
    assert int(rep["total iteration"]) == int(vals["itno"])
    return dict(flag=int(vals["iter_flag"]), itno=int(vals["itno"]), final=float(vals["final_err"]),
                init=float(vals["init_err"]), init_ref=float(rep["initial error"]),
                final_ref=float(rep["final error"]), params=np.fromfile(out, np.float64))


REFERENCE_FILES = [("7camsvarK.txt", "7pts.txt"), ("9camsvarK.txt", "9pts.txt"), ("54camsvarK.txt", "54pts.txt"),
                   ("Trafalgar-21-11315-cams.txt", "Trafalgar-21-11315-pts.txt")]


@pytest.mark.parametrize("cams_name,pts_name", REFERENCE_FILES)
def test_reference_program_on_reference_data(P, O, gpu, tmp_path, cams_name, pts_name):
    """main_dropin (PSBA/main.cpp's sequence, PSBA/main.cpp:61-62 runs
    Trafalgar-21-11315) against the C-ABI hybrid on the same files: the same
    stop code, iteration count, final SSE and parameters bit for bit; initErr
    is the SSE at the entry of the last levmar call (PSBA/levmar.cpp:95).  The
    oracle's hybrid solve stops with the same code after the same number of
    iterations at the same final error (1e-12; measured <= 1.6e-14).  As in the
    reference no camera is held, so the similarity gauge of the parameters is
    free and they are compared through the cost, not one by one (the
    oracle's and the GPU's parameters differ along it by up to 1.3e-3)."""
    d = run_dropin(os.path.join(GOLDEN, cams_name), os.path.join(GOLDEN, pts_name), tmp_path)
    pr = load_reference_problem(P, cams_name, pts_name)
    with P.Context(6, pr["n_cams"], pr["n_pts"], pr["n_obs"]) as g:
        g.upload(pr["K"], pr["meas"], pr["q0"], pr["cams"], pr["pts"])
        g.set_index(pr["iidx"], pr["jidx"])
        r = g.hybrid()
        assert (r["code"], r["itno"]) == (d["flag"], d["itno"])
        assert r["final_err"] == d["final"]
        assert np.array_equal(np.concatenate(g.read_params()), d["params"])
        assert d["init"] == g.levmar_entry_cost()
        if not any(t.kind == 1 for t in r["trace"]):  # one levmar call: its entry is the initial SSE
            assert d["init"] == r["init_err"]
    no = pr["n_obs"]
    assert abs(d["init_ref"] - np.sqrt(d["init"]) / no) <= 1e-14 * d["init_ref"]
    assert abs(d["final_ref"] - np.sqrt(d["final"]) / no) <= 1e-14 * d["final_ref"]
    # the oracle on the same inputs (no camera fixed, as in the reference)
    O.set_threads(16)
    o = O.Oracle(6, pr["K"], pr["q0"], pr["meas"], pr["iidx"], pr["jidx"], pr["cams"], pr["pts"], None)
    ro = o.hybrid()
    assert (d["flag"], d["itno"]) == (ro["code"], ro["itno"])
    assert abs(d["final"] - ro["final_err"]) <= 1e-12 * ro["final_err"]


@pytest.mark.parametrize("cams_name,pts_name", REFERENCE_FILES)
def test_reference_data_lm_phase_trial_by_trial(P, O, gpu, cams_name, pts_name):
    """The LM phase the reference program starts with, on the reference's data
    files, trial by trial against the oracle: every trial's decision and, for
    accepted trials, the cost within 1e-9."""
    pr = load_reference_problem(P, cams_name, pts_name)
    O.set_threads(16)
    with P.Context(6, pr["n_cams"], pr["n_pts"], pr["n_obs"]) as g:
        g.upload(pr["K"], pr["meas"], pr["q0"], pr["cams"], pr["pts"])
        g.set_index(pr["iidx"], pr["jidx"])
        rg = g.levmar()
        o = O.Oracle(6, pr["K"], pr["q0"], pr["meas"], pr["iidx"], pr["jidx"], pr["cams"], pr["pts"], None)
        ro = o.levmar()
        assert (rg["code"], rg["itno"]) == (ro["code"], ro["itno"])
        compare_traces(rg["trace"], ro["trace"])
        assert rel(rg["final_err"], ro["final_err"]) < 1e-9
        assert rel(np.concatenate(g.read_params()), np.concatenate(o.read_params())) < 1e-9


@pytest.fixture(scope="module")
def c2_full(P):
    return P.Scene(2, 2024)


def test_full_size_iteration_matches_oracle(P, O, gpu, c2_full):
    sc = c2_full
    assert (sc.n_cams, sc.n_pts) == (88, 64298)
    O.set_threads(16)
    g = P.Context.from_scene(sc)
    o = O.Oracle.from_scene(sc)
    assert rel(g.residual(), o.residual()) < 1e-13
    g.jacobian_normal()
    o.jacobian_normal()
    for b in ("U", "V", "G"):
        assert rel(g.buffer(b), o.buffer(b)) < 1e-12, b
    mu = 1e-3 * o.max_uvdiag()
    g.schur(mu)
    o.schur(mu)
    n = sc.cnp * sc.n_cams
    tl = np.tril_indices(n)
    assert rel(g.buffer("S").reshape(n, n)[tl], o.buffer("S").reshape(n, n)[tl]) < 1e-11
    assert rel(g.buffer("EA"), o.buffer("EA")) < 1e-11
    assert g.factor_solve() and o.factor_solve()
    g.backsub()
    o.backsub()
    assert rel(g.buffer("DP"), o.buffer("DP")) < 1e-9


def test_full_size_solve_matches_oracle_without_outliers(P, O, gpu):
    """The 88-camera / 64,298-point geometry without gross outliers: the whole
    LM solve (stop on relative decrease < 1e-8) trial by trial against the
    oracle, and the final parameters."""
    sc = P.Scene(2, 2024, outlier_frac=0.0)
    O.set_threads(16)
    g = P.Context.from_scene(sc)
    o = O.Oracle.from_scene(sc)
    rg = g.levmar(max_iter=100, flags=P.FLAG_NO_HANDOFF, rel_tol=1e-8)
    ro = o.levmar(max_iter=100, flags=O.FLAG_NO_HANDOFF, rel_tol=1e-8)
    assert (rg["code"], rg["itno"]) == (ro["code"], ro["itno"])
    compare_traces(rg["trace"], ro["trace"])
    assert rel(rg["final_err"], ro["final_err"]) < 1e-9
    assert rel(np.concatenate(g.read_params()), np.concatenate(o.read_params())) < 1e-9


def test_full_size_solve_is_deterministic_and_monotone(P, gpu, c2_full):
    sc = c2_full
    g = P.Context.from_scene(sc)
    init = g.residual()
    g.snapshot()
    r1 = g.levmar(max_iter=100, flags=P.FLAG_NO_HANDOFF, rel_tol=1e-8)
    p1 = np.concatenate(g.read_params())
    g.restore()
    g.itno = 0
    r2 = g.levmar(max_iter=100, flags=P.FLAG_NO_HANDOFF, rel_tol=1e-8)
    p2 = np.concatenate(g.read_params())
    assert r1["final_err"] == r2["final_err"] and np.array_equal(p1, p2)
    acc = [t.cost for t in r1["trace"] if t.accepted]
    assert len(acc) > 5 and all(b < a for a, b in zip(acc, acc[1:]))
    assert r1["final_err"] < init


def test_full_size_outlier_solve_trial_by_trial_to_horizon(P, O, gpu):
    """The headline 88-camera scene with its 1% gross outliers, trial by trial
    against the oracle over the first 30 iterations.  Two valid FP64 orders of
    the oracle itself agree to 1e-9 on this scene until iteration 41 and part
    after (tests/test_oracle_chaos.py shows the same on a smaller scene), so
    30 iterations is the horizon over which any correct implementation can be
    held to 1e-9 per trial (the benchmark's scene, seed 1002)."""
    sc = P.Scene(2, 1002)
    O.set_threads(16)
    g = P.Context.from_scene(sc)
    o = O.Oracle.from_scene(sc)
    rg = g.levmar(max_iter=30, flags=P.FLAG_NO_HANDOFF, rel_tol=1e-8)
    ro = o.levmar(max_iter=30, flags=O.FLAG_NO_HANDOFF, rel_tol=1e-8)
    assert (rg["code"], rg["itno"]) == (ro["code"], ro["itno"]) == (P.ITER_CONTINUE, 30)
    compare_traces(rg["trace"], ro["trace"])
    assert rel(rg["final_err"], ro["final_err"]) < 1e-9
