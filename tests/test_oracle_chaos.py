"""How far two valid FP64 evaluation orders of the same LM solve stay together.

The oracle is run twice on one scene, once with every reduction (cost, U, V,
g, S, e_a, e_b, the gain-ratio denominator) in ascending index order -- the
reference's -- and once in descending order.  Both are correct FP64 programs.

* Without gross outliers the two solves agree trial by trial to the end.
* With 1% gross outliers (the sparse-scene workload) the least-squares
  trajectory amplifies the 1e-16 differences: the accepted costs agree to
  1e-9 for the first 40 iterations and then part, and the two solves stop at
  different iterations with final costs apart by far more than 1e-6.  A GPU
  solve of the outlier scene can therefore be held to the oracle trial by
  trial only over that horizon (tests/test_gpu_parity.py), and its final RMS
  cannot be held to 1e-6 of the oracle's by any FP64 implementation.
"""
import numpy as np


def solve_both(P, O, sc):
    O.set_threads(8)
    out = []
    try:
        for desc in (False, True):
            O.set_sum_order(desc)
            out.append(O.Oracle.from_scene(sc).levmar(max_iter=100, flags=O.FLAG_NO_HANDOFF, rel_tol=1e-8))
    finally:
        O.set_sum_order(False)
    return out


def first_iteration_apart(ta, tb, tol):
    for a, b in zip(ta, tb):
        if (a.accepted, a.factor_ok) != (b.accepted, b.factor_ok):
            return a.itno
        if a.accepted and abs(a.cost - b.cost) > tol * abs(b.cost):
            return a.itno
    return None


def test_outlier_free_orders_agree(P, O):
    sc = P.Scene(2, 2100, n_pts=8000, n_cams=40, outlier_frac=0.0)
    ra, rb = solve_both(P, O, sc)
    assert (ra["code"], ra["itno"]) == (rb["code"], rb["itno"])
    assert first_iteration_apart(ra["trace"], rb["trace"], 1e-12) is None
    assert abs(ra["final_err"] - rb["final_err"]) <= 1e-13 * ra["final_err"]


def test_outlier_orders_part_after_a_horizon(P, O):
    sc = P.Scene(2, 2100, n_pts=8000, n_cams=40)
    ra, rb = solve_both(P, O, sc)
    k9 = first_iteration_apart(ra["trace"], rb["trace"], 1e-9)
    k6 = first_iteration_apart(ra["trace"], rb["trace"], 1e-6)
    # together (1e-9) for the first 40 iterations, apart (1e-6) before the end
    assert k9 is not None and k9 >= 40
    assert k6 is not None and k6 > k9
    rel = abs(ra["final_err"] - rb["final_err"]) / ra["final_err"]
    assert rel > 1e-6
    rms_a = np.sqrt(ra["final_err"] / (2 * sc.n_obs))
    rms_b = np.sqrt(rb["final_err"] / (2 * sc.n_obs))
    assert abs(rms_a - rms_b) / rms_a > 5e-7
