"""Benchmark: time-to-converge of the 88-camera / 64,298-point synthetic scene.

One step = one complete bundle-adjustment solve of scene config 2 (88 cameras,
cnp 9 = f, k1, k2 + pose; 64,298 points; Zipf track lengths fitted to ~384k
observations; 1 px noise + 1% outliers) on one MI355X: parameters are reset on
the device to the same perturbed start, then Levenberg-Marquardt runs until an
accepted step lowers the cost by less than 1e-8 relative (or 100 iterations).
Inputs are resident in HBM before the timed region.

    python bench.py [--gpus 1] [--steps K] [--warmup W]

Prints one JSON line (see the "metric" and "config" fields).
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "time-to-converge vs plain FP64 at equal RMS; solves/s + p99 latency (stream)"
FP64_PEAK_TFLOPS = 78.6  # MI355X dense FP64 (matrix and vector)


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def profiled_traffic(kernel):
    """HBM bytes per launch of `kernel` from its committed rocprofv3 PMC summary
    (profiles/<kernel>_c2_traffic.md, FETCH_SIZE + WRITE_SIZE scaled by a copy
    calibration), or None when no summary exists."""
    path = os.path.join(ROOT, "profiles", "%s_c2_traffic.md" % kernel)
    if not os.path.exists(path):
        return None, None
    for line in open(path):
        if line.startswith("bytes_per_launch"):
            return float(line.split("=")[1]), os.path.relpath(path, ROOT)
    return None, None


def cpu_baseline(scene, kw):
    """The oracle (the CPU FP64 restatement of the reference's LM) on this
    host's cores: one complete solve of the same scene with the same stopping
    rule, timed end to end."""
    from oracle import binding as O
    threads = int(os.environ.get("OMP_NUM_THREADS") or min(16, os.cpu_count() or 1))
    O.set_threads(threads)
    orc = O.Oracle.from_scene(scene)
    t0 = time.perf_counter()
    r = orc.levmar(max_iter=kw["max_iter"], flags=O.FLAG_NO_HANDOFF, rel_tol=kw["rel_tol"])
    dt = time.perf_counter() - t0
    return {"value": 1.0 / dt, "unit": "solves/s", "cores": threads, "kind": "port",
            "sample": "one full LM solve of the same scene and stopping rule (%d iterations, %d trials, "
                      "%.1f s, final cost %.10g)" % (r["itno"], len(r["trace"]), dt, r["final_err"])}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--seed", type=int, default=1002)
    ap.add_argument("--no-cpu", action="store_true")
    a = ap.parse_args()
    if a.gpus != 1:
        print("bench.py measures one MI355X only: multi-GPU runs are not part of this build "
              "(run with --gpus 1)", file=sys.stderr)
        sys.exit(2)
    import psba_amd as P

    sc = P.Scene(2, a.seed)
    log("scene: cams=%d cnp=%d pts=%d obs=%d sum_t2=%.0f zipf_s=%.3f outliers=%d" % (
        sc.n_cams, sc.cnp, sc.n_pts, sc.n_obs, sc.sum_t2, sc.zipf_s, sc.n_outliers))
    ctx = P.Context.from_scene(sc)
    ctx.snapshot()
    kw = dict(max_iter=100, flags=P.FLAG_NO_HANDOFF, rel_tol=1e-8)

    def solve_on(c):
        c.restore()
        c.itno = 0
        return c.levmar(**kw)

    def solve():
        return solve_on(ctx)

    def timed(c, steps, warmup):
        c.snapshot()
        for _ in range(warmup):
            solve_on(c)
        t0 = time.perf_counter()
        rr = [solve_on(c) for _ in range(steps)]
        return (time.perf_counter() - t0) / steps, rr[-1]

    for _ in range(a.warmup):
        solve()
    # HIP events around every Schur-assembly launch on the context's stream
    ctx.probe(P.PROBE_SCHUR)
    ctx.probe_read()
    res = []
    t0 = time.perf_counter()
    for _ in range(a.steps):
        res.append(solve())  # returns after its final device read-back (stream drained)
    dt = time.perf_counter() - t0
    schur_ms, schur_n = ctx.probe_read()
    ctx.probe(P.PROBE_OFF)
    info = ctx.info()
    r = res[-1]
    trials = len(r["trace"])
    final = r["final_err"]
    rms = float(np.sqrt(final / (2 * sc.n_obs)))
    log("solve: code=%d iters=%d trials=%d final_sse=%.6e rms=%.6f px, %.3f ms/solve" % (
        r["code"], r["itno"], trials, final, rms, 1e3 * dt / a.steps))
    ms = 1e3 * dt / a.steps
    out = {
        "metric": METRIC,
        "value": a.steps / dt,
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded scene generator, config 2)",
        "config": {"workload": "config 2: 88 cams (f,k1,k2+pose), %d pts, %d obs, LM to rel. decrease < 1e-8, "
                               "plain FP64" % (sc.n_pts, sc.n_obs),
                   "n_obs": sc.n_obs, "sum_t2": sc.sum_t2, "lm_iterations": r["itno"], "lm_trials": trials,
                   "final_rms_px": rms},
    }
    # Schur assembly S = U - sum_i W_ij V_i^-1 W_ik^T over the lower camera-pair
    # blocks: per (point, pair j >= k of its cameras) one C x 3 by 3 x C product,
    # 3 C^2 FMAs = 6 C^2 flops; sum over points of t(t+1)/2 such products.
    C = sc.cnp
    flops = 6.0 * C * C * info["n_pair_items"]
    avg_s = schur_ms * 1e-3 / max(schur_n, 1)
    achieved = flops / avg_s / 1e12
    traffic, traffic_src = profiled_traffic("k_schur_chunks")
    out["roofline"] = {"bound": "mfma", "achieved": achieved, "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                       "frac": achieved / FP64_PEAK_TFLOPS, "traffic": traffic,
                       "kernel": "k_schur_chunks", "avg_launch_us": avg_s * 1e6, "launches": schur_n,
                       "flops_per_launch": flops, "traffic_source": traffic_src}
    log("schur: %d launches, %.1f us avg, %.2f TFLOP/s (%.1f%% of FP64 peak)" % (
        schur_n, avg_s * 1e6, achieved, 100 * achieved / FP64_PEAK_TFLOPS))
    # mixed precision (FP32 Jacobian / W / V / Schur products, FP64 factor and
    # cost) on the same scene, and both paths on the same geometry without
    # gross outliers, where every solve converges to one minimum and the final
    # RMS reprojection errors can be compared
    ms = max(1, a.steps // 2)
    mctx = P.Context.from_scene(sc, precision=P.PREC_MIXED)
    t_mix, r_mix = timed(mctx, ms, 1)
    rms_mix = float(np.sqrt(r_mix["final_err"] / (2 * sc.n_obs)))
    sc0 = P.Scene(2, a.seed, outlier_frac=0.0)
    t_p0, r_p0 = timed(P.Context.from_scene(sc0), ms, 1)
    t_m0, r_m0 = timed(P.Context.from_scene(sc0, precision=P.PREC_MIXED), ms, 1)
    rms_p0 = float(np.sqrt(r_p0["final_err"] / (2 * sc0.n_obs)))
    rms_m0 = float(np.sqrt(r_m0["final_err"] / (2 * sc0.n_obs)))
    out["mixed"] = {
        "solves_per_s": 1.0 / t_mix, "ms_per_solve": 1e3 * t_mix, "speedup_vs_plain": (dt / a.steps) / t_mix,
        "lm_iterations": r_mix["itno"], "final_rms_px": rms_mix,
        "rel_rms_vs_plain": abs(rms_mix - rms) / rms,
        "outlier_free": {"plain_ms": 1e3 * t_p0, "mixed_ms": 1e3 * t_m0, "speedup": t_p0 / t_m0,
                         "plain_iterations": r_p0["itno"], "mixed_iterations": r_m0["itno"],
                         "plain_rms_px": rms_p0, "mixed_rms_px": rms_m0,
                         "rel_rms": abs(rms_m0 - rms_p0) / rms_p0},
    }
    log("mixed: %.3f ms/solve (%.2fx), rms %.6f (rel %.2e); outlier-free: plain %.3f ms, mixed %.3f ms (%.2fx), "
        "rms rel %.2e" % (1e3 * t_mix, (dt / a.steps) / t_mix, rms_mix, abs(rms_mix - rms) / rms, 1e3 * t_p0,
                          1e3 * t_m0, t_p0 / t_m0, abs(rms_m0 - rms_p0) / rms_p0))
    if not a.no_cpu:
        out["cpu_baseline"] = cpu_baseline(sc, kw)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
