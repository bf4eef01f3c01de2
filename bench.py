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
HBM_PEAK_GBS = 8000.0     # MI355X HBM3E


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


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(P, scene, gpu_trials, reps=5):
    """The oracle (the CPU FP64 restatement of the reference's LM, with the
    reference's explicit S^-1 solve, PSBA/CL_files/SPD_inv.cl:248-411 +
    matVec_mul.cl:7-18) on this host's cores.  Bounded sample: one LM
    iteration (Jacobian, normal equations, explicit-inverse step, trial cost)
    from the starting point of each scene, median of `reps` runs, on one
    thread and on all cores used here; configs 0, 1 (cnp 11), 2 and the
    three request classes of config 4.  For configs 0, 1 and 4 also the
    time to converge, median of `reps` full solves, beside the GPU's
    (`gpu_ms_to_converge`, same stop rule).  The headline value scales the
    all-core iteration time of config 2 by the GPU solve's trial count."""
    from oracle import binding as O
    nproc = os.cpu_count() or 1
    allc = int(os.environ.get("OMP_NUM_THREADS") or min(16, nproc))
    cases = [("config2", scene), ("config0", P.Scene(0, 1000)), ("config1", P.Scene(1, 1001)),
             ("config4_small", P.Scene(4, 1004, problem=0, size_class=0)),
             ("config4_medium", P.Scene(4, 1004, problem=10000, size_class=1)),
             ("config4_large", P.Scene(4, 1004, problem=20000, size_class=2))]
    # full solves to convergence beside the GPU's, where the CPU finishes in
    # seconds: the configs' own stop rules (config 0 and 1: at most 50
    # iterations; config 4: relative decrease < 1e-8 or 100 iterations), LM
    # without the dogleg hand-off on both sides; 1-thread runs where cheap
    converge = {"config0": (dict(max_iter=50), True), "config1": (dict(max_iter=50), False),
                "config4_small": (dict(max_iter=100, rel_tol=1e-8), True),
                "config4_medium": (dict(max_iter=100, rel_tol=1e-8), False),
                "config4_large": (dict(max_iter=100, rel_tol=1e-8), False)}
    per = {}
    for name, sc in cases:
        row = {"n_cams": sc.n_cams, "n_pts": sc.n_pts, "n_obs": sc.n_obs}
        for th in sorted({1, allc}):
            O.set_threads(th)
            ts = []
            for _ in range(reps):
                o = O.Oracle.from_scene(sc)
                t0 = time.perf_counter()
                o.levmar(max_iter=1, solve_mode=O.SOLVE_INVERSE, flags=O.FLAG_NO_HANDOFF)
                ts.append(time.perf_counter() - t0)
            row["ms_per_iteration_threads_%d" % th] = 1e3 * float(np.median(ts))
        if name in converge:
            kw, one_thread = converge[name]
            g = P.Context.from_scene(sc)
            g.snapshot()
            gt = []
            for _ in range(reps + 1):
                g.restore()
                g.itno = 0
                t0 = time.perf_counter()
                r = g.levmar(flags=P.FLAG_NO_HANDOFF, **kw)
                gt.append(time.perf_counter() - t0)
            g.close()
            row["gpu_ms_to_converge"] = 1e3 * float(np.median(gt[1:]))
            row["gpu_iterations"] = r["itno"]
            for th in sorted({1, allc} if one_thread else {allc}):
                O.set_threads(th)
                ts = []
                for _ in range(reps):
                    o = O.Oracle.from_scene(sc)
                    t0 = time.perf_counter()
                    ro = o.levmar(solve_mode=O.SOLVE_INVERSE, flags=O.FLAG_NO_HANDOFF, **kw)
                    ts.append(time.perf_counter() - t0)
                row["ms_to_converge_threads_%d" % th] = 1e3 * float(np.median(ts))
            row["cpu_iterations"] = ro["itno"]
        per[name] = row
        log("cpu baseline %s: %s" % (name, ", ".join(("%s %.1f" % kv) if isinstance(kv[1], float) else "%s %s" % kv
                                                     for kv in row.items() if kv[0].startswith(("ms", "gpu", "cpu")))))
    it_s = per["config2"]["ms_per_iteration_threads_%d" % allc] * 1e-3
    return {"value": 1.0 / (it_s * gpu_trials), "unit": "solves/s", "cores": allc, "kind": "port",
            "sample": "median of %d runs of one LM iteration (explicit-inverse step) of the config-2 scene on %d "
                      "threads, %.3f s, scaled by the %d trials of the GPU solve; 1-thread, other configs and "
                      "times to converge (CPU and GPU) in per_config" % (reps, allc, it_s, gpu_trials),
            "cpu_model": cpu_model(), "nproc": nproc, "per_config": per}


def mixed_section(P, sc, a, t_plain, rms):
    """Mixed precision (FP32 Jacobian / W / V / Schur products, FP64 factor and
    cost) on the same scene, and both paths on the same geometry without gross
    outliers, where every solve converges to one minimum and the final RMS
    reprojection errors can be compared."""
    kw = dict(max_iter=100, flags=P.FLAG_NO_HANDOFF, rel_tol=1e-8)

    def solve_on(c):
        c.restore()
        c.itno = 0
        return c.levmar(**kw)

    def timed(c, steps, warmup):
        c.snapshot()
        for _ in range(warmup):
            solve_on(c)
        t0 = time.perf_counter()
        rr = [solve_on(c) for _ in range(steps)]
        return (time.perf_counter() - t0) / steps, rr[-1]

    ms = max(1, a.steps // 2)
    mctx = P.Context.from_scene(sc, precision=P.PREC_MIXED)
    t_mix, r_mix = timed(mctx, ms, 1)
    rms_mix = float(np.sqrt(r_mix["final_err"] / (2 * sc.n_obs)))
    sc0 = P.Scene(2, a.seed, outlier_frac=0.0)
    t_p0, r_p0 = timed(P.Context.from_scene(sc0), ms, 1)
    t_m0, r_m0 = timed(P.Context.from_scene(sc0, precision=P.PREC_MIXED), ms, 1)
    rms_p0 = float(np.sqrt(r_p0["final_err"] / (2 * sc0.n_obs)))
    rms_m0 = float(np.sqrt(r_m0["final_err"] / (2 * sc0.n_obs)))
    # mixed with one FP64 refinement sweep per step (follows the FP64 trials)
    kw_r = dict(kw, refine_steps=1)
    cref = P.Context.from_scene(sc0, precision=P.PREC_MIXED)
    cref.snapshot()
    cref.levmar(**kw_r)
    t0 = time.perf_counter()
    for _ in range(ms):
        cref.restore()
        cref.itno = 0
        r_r0 = cref.levmar(**kw_r)
    t_r0 = (time.perf_counter() - t0) / ms
    rms_r0 = float(np.sqrt(r_r0["final_err"] / (2 * sc0.n_obs)))
    step_err = step_errors(P, sc)
    res = {
        "solves_per_s": 1.0 / t_mix, "ms_per_solve": 1e3 * t_mix, "speedup_vs_plain": t_plain / t_mix,
        "lm_iterations": r_mix["itno"], "final_rms_px": rms_mix,
        "rel_rms_vs_plain": abs(rms_mix - rms) / rms,
        "outlier_free": {"plain_ms": 1e3 * t_p0, "mixed_ms": 1e3 * t_m0, "speedup": t_p0 / t_m0,
                         "plain_iterations": r_p0["itno"], "mixed_iterations": r_m0["itno"],
                         "plain_rms_px": rms_p0, "mixed_rms_px": rms_m0,
                         "rel_rms": abs(rms_m0 - rms_p0) / rms_p0,
                         "mixed_refined_ms": 1e3 * t_r0, "mixed_refined_iterations": r_r0["itno"],
                         "mixed_refined_rel_rms": abs(rms_r0 - rms_p0) / rms_p0},
        "first_step_rel_error_vs_fp64": step_err,
    }
    log("mixed: %.3f ms/solve (%.2fx), rms %.6f (rel %.2e); outlier-free: plain %.3f ms, mixed %.3f ms (%.2fx), "
        "rms rel %.2e" % (1e3 * t_mix, t_plain / t_mix, rms_mix, abs(rms_mix - rms) / rms, 1e3 * t_p0,
                          1e3 * t_m0, t_p0 / t_m0, abs(rms_m0 - rms_p0) / rms_p0))
    log("mixed refined: %.3f ms/solve, %d iterations, rms rel %.2e; first-step error vs FP64 %s" % (
        1e3 * t_r0, r_r0["itno"], abs(rms_r0 - rms_p0) / rms_p0,
        ", ".join("%s %.2e" % kv for kv in step_err.items())))
    return res


def step_errors(P, sc):
    """Relative error of the mixed path's first step (at mu0) against the FP64
    step, unrefined and after one and two FP64 refinement sweeps."""
    c64 = P.Context.from_scene(sc)
    c64.jacobian_normal(1.0, 1.0)
    C = sc.cnp
    diag = np.concatenate([c64.buffer("U").reshape(-1, C, C).diagonal(axis1=1, axis2=2).ravel(),
                           c64.buffer("V").reshape(-1, 3, 3).diagonal(axis1=1, axis2=2).ravel()])
    mu = 1e-3 * float(diag.max())
    c64.schur(mu)
    c64.factor_solve()
    c64.backsub()
    ref = c64.buffer("DP")
    cm = P.Context.from_scene(sc, precision=P.PREC_MIXED)
    cm.jacobian_normal(1.0, 1.0)
    cm.schur(mu)
    cm.factor_solve()
    cm.backsub()
    out = {}
    den = float(np.max(np.abs(ref)))
    for k in range(3):
        out["sweeps_%d" % k] = float(np.max(np.abs(cm.buffer("DP") - ref)) / den)
        if k < 2:
            cm.refine(mu, 1)
    return out


def batch_section(P, seed=1003, n=1024, reps=3):
    """1,024 independent windows (config 3) per launch: every LM stage is one
    launch over all problems, rounds are replayed HIP graphs, launch shapes
    autotuned; mixed precision, with plain FP64 beside it."""
    scenes = [P.Scene(3, seed, problem=k) for k in range(n)]
    n_obs = sum(s.n_obs for s in scenes)
    out = {"problems": n, "n_obs": n_obs}
    finals = {}
    for name, prec in (("mixed", P.PREC_MIXED), ("fp64", P.PREC_FP64)):
        b = P.Batch(6, precision=prec)
        for sc in scenes:
            b.add_scene(sc)
        b.commit()
        shapes = b.autotune()
        b.solve()  # warm-up
        t = []
        for _ in range(reps):
            t0 = time.perf_counter()
            rounds = b.solve()
            t.append(time.perf_counter() - t0)
        dev_ms, _ = b.elapsed()
        res = [b.result(k, params=False) for k in range(n)]
        its = np.array([r["itno"] for r in res])
        finals[name] = np.array([r["final_err"] for r in res])
        wall = float(np.median(t))
        out[name] = {"solves_per_s": n / wall, "ms_per_batch": 1e3 * wall, "device_ms": dev_ms,
                     "lm_rounds": rounds, "mean_iterations": float(its.mean()), "max_iterations": int(its.max()),
                     "jacobian_obs_per_s": float(n_obs * rounds / wall),
                     "autotuned_blocks": {"obs": shapes[0], "point": shapes[1]}, "bytes": b.nbytes()}
        b.close()
        log("batch %s: %d problems, %.1f ms/batch (%.0f solves/s), %d rounds" % (
            name, n, 1e3 * wall, n / wall, rounds))
    rms_m = np.sqrt(finals["mixed"] / np.array([2 * s.n_obs for s in scenes]))
    rms_p = np.sqrt(finals["fp64"] / np.array([2 * s.n_obs for s in scenes]))
    out["max_rel_rms_mixed_vs_fp64"] = float(np.max(np.abs(rms_m - rms_p) / rms_p))
    out["speedup_mixed_vs_fp64"] = out["mixed"]["solves_per_s"] / out["fp64"]["solves_per_s"]
    return out


def index_section(P, sc, reps=5):
    """Index construction at the headline scene: the host loops against the
    device build (stable radix sorts and scans), both bit-identical; and
    generate_idxs from the visibility mask, host and device."""
    vm = np.zeros((sc.n_pts, sc.n_cams), np.int8)
    vm[sc.iidx, sc.jidx] = 1
    c = P.Context(sc.cnp, sc.n_cams, sc.n_pts, sc.n_obs)
    c.set_index_device(sc.iidx, sc.jidx)  # warm-up (first launches)
    P.generate_idxs_device(sc.n_cams, sc.n_pts, sc.n_obs, vm)

    def med(fn):
        t = []
        for _ in range(reps):
            t0 = time.perf_counter()
            fn()
            t.append(time.perf_counter() - t0)
        return 1e3 * float(np.median(t))

    out = {"set_index_host_ms": med(lambda: c.set_index(sc.iidx, sc.jidx)),
           "set_index_device_ms": med(lambda: c.set_index_device(sc.iidx, sc.jidx)),
           "generate_idxs_host_ms": med(lambda: P.generate_idxs(sc.n_cams, sc.n_pts, sc.n_obs, vm)),
           "generate_idxs_device_ms": med(lambda: P.generate_idxs_device(sc.n_cams, sc.n_pts, sc.n_obs, vm)),
           "pair_items": c.info()["n_pair_items"]}
    c.close()
    log("index: set_index host %.2f ms, device %.2f ms; generate_idxs host %.2f ms, device %.2f ms" % (
        out["set_index_host_ms"], out["set_index_device_ms"], out["generate_idxs_host_ms"],
        out["generate_idxs_device_ms"]))
    return out


def stream_section(P, seed=1004, rates=(200, 1000, 2000, 4000, 8000), duration=2.0, serial_n=100, capacity_n=2000):
    """Open-loop Poisson request stream (config 4 mix: 90% small, 9% medium,
    1% large) through the request loop, against one-at-a-time plain FP64
    solves of the same mix; and the loop's capacity: capacity_n requests of
    the mix all submitted at once, solves/s until the last completes."""
    rng = np.random.default_rng(seed)
    pool = {0: [], 1: [], 2: []}
    for c, count in ((0, 300), (1, 30), (2, 6)):
        for k in range(count):
            pool[c].append(P.Scene(4, seed, problem=10000 * c + k, size_class=c))

    def draw(m):
        u = rng.random(m)
        cls = np.where(u < 0.90, 0, np.where(u < 0.99, 1, 2))
        return [pool[c][rng.integers(len(pool[c]))] for c in cls], cls

    kw = dict(max_iter=100, flags=P.FLAG_NO_HANDOFF, rel_tol=1e-8)
    sample, _ = draw(serial_n)
    serial = []
    t0 = time.perf_counter()
    for sc in sample:
        c = P.Context.from_scene(sc)
        r = c.levmar(**kw)
        c.read_params()
        c.close()
        serial.append(r["final_err"])
    serial_rate = serial_n / (time.perf_counter() - t0)
    out = {"serial_plain_fp64_solves_per_s": serial_rate, "serial_sample": serial_n, "precision": "mixed",
           "lanes": 4, "hbm_budget_bytes": 200e9, "rates": []}
    log("stream: serial plain FP64 %.1f solves/s" % serial_rate)

    def run(reqs, arr):
        srv = P.Server(6, precision=P.PREC_MIXED, lanes=4)
        ids, done = {}, {}
        T0 = P.server_clock() + 0.01
        i, m = 0, len(arr)
        while len(done) < m:
            now = P.server_clock() - T0
            while i < m and arr[i] <= now:
                ids[srv.submit_scene(reqs[i])] = i
                i += 1
            srv.step()
            for rid in srv.completed():
                done[ids[rid]] = srv.result(rid, params=False)
        st = srv.stats()
        srv.close()
        lat = np.array([done[k]["t_done"] - (T0 + arr[k]) for k in range(m)])
        span = max(done[k]["t_done"] for k in range(m)) - T0
        return done, lat, span, st

    for rate in rates:
        gaps = rng.exponential(1.0 / rate, int(rate * duration * 1.5) + 10)
        arr = np.cumsum(gaps)
        arr = arr[arr < duration]
        reqs, cls = draw(len(arr))
        done, lat, span, st = run(reqs, arr)
        rec = {"offered_per_s": rate, "requests": len(arr), "solves_per_s": len(arr) / span,
               "p50_ms": 1e3 * float(np.percentile(lat, 50)), "p99_ms": 1e3 * float(np.percentile(lat, 99)),
               "p99_ms_by_class": [1e3 * float(np.percentile(lat[cls == c], 99)) if np.any(cls == c) else None
                                   for c in (0, 1, 2)],
               "batches": st["batches"], "peak_bytes": st["bytes_peak"]}
        rec["vs_serial"] = rec["solves_per_s"] / serial_rate
        out["rates"].append(rec)
        log("stream %d req/s: %.1f solves/s (%.1fx serial), p50 %.1f ms, p99 %.1f ms" % (
            rate, rec["solves_per_s"], rec["vs_serial"], rec["p50_ms"], rec["p99_ms"]))
    # accuracy: the serial sample as one burst through the loop
    done, lat, span, st = run(sample, np.zeros(len(sample)))
    rel = [abs(done[k]["final_err"] - serial[k]) / serial[k] for k in range(len(sample))]
    out["max_rel_cost_vs_serial"] = float(max(rel))
    out["burst_solves_per_s"] = len(sample) / span
    reqs, cls = draw(capacity_n)
    done, lat, span, st = run(reqs, np.zeros(capacity_n))
    out["capacity"] = {"requests": capacity_n, "solves_per_s": capacity_n / span,
                       "vs_serial": capacity_n / span / serial_rate, "p50_ms": 1e3 * float(np.percentile(lat, 50)),
                       "p99_ms": 1e3 * float(np.percentile(lat, 99)), "batches": st["batches"]}
    log("stream capacity: %d requests at once, %.1f solves/s (%.1fx serial)" % (
        capacity_n, capacity_n / span, capacity_n / span / serial_rate))
    return out


def stage_rates(P, ctx, solve_on, sc, info, r, solve_s):
    """Absolute rates of the headline solve: LM iterations and trials per
    second, and the Jacobian, Schur-assembly and back-substitution stages,
    each timed by HIP events on the context's stream over one further solve,
    with their roofline fractions.  Jacobian bytes per observation: the reads
    (camera, point, measurement, indices: 40 B) and the writes (residual 2,
    J_A 2C, J_B 6 and W 3C doubles).  Schur flops: per (point, camera pair
    k >= l of its track) one C x 3 by 3 x C product, 6 C^2 flops (the launch
    also runs the e_a chunks, whose few flops are not counted).  Back
    substitution bytes: W rows (3C doubles) and camera indices per
    observation, V^-1, g_b and the outputs per point."""
    out = {"lm_iterations_per_s": r["itno"] / solve_s, "lm_trials_per_s": len(r["trace"]) / solve_s}
    timed = {}
    for name, which in (("schur", P.PROBE_SCHUR), ("backsub", P.PROBE_BSOLVE)):
        ctx.probe(which)
        ctx.probe_read()
        solve_on(ctx)
        timed[name] = ctx.probe_read()
        ctx.probe(P.PROBE_OFF)
    # the Jacobian stage runs gated in a device-controlled solve (its launches
    # in trials that open no iteration exit early), so it is timed over direct
    # stage calls, each a full pass
    ctx.probe(P.PROBE_JACOBIAN)
    ctx.probe_read()
    for _ in range(20):
        ctx.jacobian_normal()
    timed["jacobian"] = ctx.probe_read()
    ctx.probe(P.PROBE_OFF)
    C = sc.cnp
    ms, n = timed["jacobian"]
    avg = ms * 1e-3 / max(n, 1)
    nbytes = ((2 + 2 * C + 6 + 3 * C) * 8 + 40) * sc.n_obs
    out["jacobian"] = {"launches": n, "avg_launch_us": avg * 1e6, "obs_per_s": sc.n_obs / avg,
                       "bytes_per_launch": nbytes, "achieved_GBps": nbytes / avg / 1e9,
                       "frac_of_hbm_peak": nbytes / avg / 1e9 / HBM_PEAK_GBS}
    ms, n = timed["schur"]
    avg = ms * 1e-3 / max(n, 1)
    flops = 6.0 * C * C * info["n_pair_items"]
    traffic, src = profiled_traffic("k_schur_ea_chunks")
    out["schur"] = {"kernel": "k_schur_ea_chunks", "launches": n, "avg_launch_us": avg * 1e6, "flops_per_launch": flops,
                    "achieved_TFLOPs": flops / avg / 1e12, "frac_of_fp64_peak": flops / avg / 1e12 / FP64_PEAK_TFLOPS,
                    "traffic": traffic, "traffic_source": src}
    ms, n = timed["backsub"]
    avg = ms * 1e-3 / max(n, 1)
    nbytes = (3 * C * 8 + 4) * sc.n_obs + (6 + 3 + 3 + 3) * 8 * sc.n_pts
    out["backsub"] = {"launches": n, "avg_launch_us": avg * 1e6, "bytes_per_launch": nbytes,
                      "achieved_GBps": nbytes / avg / 1e9, "frac_of_hbm_peak": nbytes / avg / 1e9 / HBM_PEAK_GBS}
    log("stages: %.0f LM iterations/s; Jacobian %.0f us (%.2e obs/s, %.0f GB/s); Schur %.0f us (%.2f TFLOP/s); "
        "back substitution %.0f us (%.0f GB/s)" % (
            out["lm_iterations_per_s"], out["jacobian"]["avg_launch_us"], out["jacobian"]["obs_per_s"],
            out["jacobian"]["achieved_GBps"], out["schur"]["avg_launch_us"], out["schur"]["achieved_TFLOPs"],
            out["backsub"]["avg_launch_us"], out["backsub"]["achieved_GBps"]))
    return out


def path_roofline(sc, info, r, solve_s):
    """T_roof / T_measured for the whole solve: every stage at its own roof
    (HBM bytes at 8 TB/s or FP64 flops at 78.6 TFLOP/s, whichever is longer),
    summed over the solve's Jacobian passes (accepted iterations + 1) and
    trials.  Bytes and flops per stage are the algorithmic counts of the
    reference's stage graph with the Jacobian fused into U/V/W/g (FP64,
    s = 8 B): residual 40 B/obs; Jacobian + normal equations
    n_o (8 + s (2(C+3) + 3C + 2)) + s (n_c C^2 + 9 n_p); V^-1 18 s n_p;
    Schur 6 C^2 sum_i t_i (t_i + 1) / 2 flops; e_a, e_b, dp_b 2 s 3C n_o;
    factor and solves n^3 / 3 + 2 n^2 flops; step statistics and commit
    3 s n_tot."""
    C, no, nc, np_ = sc.cnp, sc.n_obs, sc.n_cams, sc.n_pts
    s = 8.0
    n = float(C * nc)
    ntot = n + 3 * np_
    bw, fl = HBM_PEAK_GBS * 1e9, FP64_PEAK_TFLOPS * 1e12
    t_res = 40.0 * no / bw
    t_jac = (no * (8 + s * (2 * (C + 3) + 3 * C + 2)) + s * (nc * C * C + 9 * np_)) / bw
    t_trial = (18 * s * np_ / bw + 6.0 * C * C * info["n_pair_items"] / fl + 2 * s * 3 * C * no / bw
               + (n ** 3 / 3 + 2 * n * n) / fl + 3 * s * ntot / bw + t_res)
    trials = len(r["trace"])
    jac = r["itno"] + 1
    t_roof = t_res + jac * t_jac + trials * t_trial
    return {"t_roof_ms": 1e3 * t_roof, "t_measured_ms": 1e3 * solve_s, "frac": t_roof / solve_s,
            "trials": trials, "jacobian_passes": jac,
            "per_trial_roof_us": 1e6 * t_trial, "per_jacobian_roof_us": 1e6 * t_jac}


def copy_calibration(nbytes=512 << 20):
    """A plain device-to-device copy of known size: the HBM bytes a
    profiler's FETCH_SIZE and WRITE_SIZE counters report for it give the
    factors that turn their readings for the solver's kernels into bytes."""
    import ctypes as ct
    H = None
    for name in ("libamdhip64.so", "/opt/rocm/lib/libamdhip64.so"):
        try:
            H = ct.CDLL(name)
            break
        except OSError:
            continue
    if H is None:
        raise RuntimeError("HIP runtime not found")
    a, b = ct.c_void_p(), ct.c_void_p()
    assert H.hipMalloc(ct.byref(a), ct.c_size_t(nbytes)) == 0 and H.hipMalloc(ct.byref(b), ct.c_size_t(nbytes)) == 0
    assert H.hipMemset(a, 1, ct.c_size_t(nbytes)) == 0
    assert H.hipDeviceSynchronize() == 0
    assert H.hipMemcpy(b, a, ct.c_size_t(nbytes), 3) == 0  # hipMemcpyDeviceToDevice
    assert H.hipDeviceSynchronize() == 0
    H.hipFree(a)
    H.hipFree(b)
    log("copied %d bytes device to device" % nbytes)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--seed", type=int, default=1002)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--headline-only", action="store_true",
                    help="only the timed solves and the roofline line (for profiler runs)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="one 512 MiB device-to-device copy and exit (scales the profiler's FETCH_SIZE/WRITE_SIZE)")
    ap.add_argument("--sections", default="mixed,index,batch,stream",
                    help="comma-separated sections after the headline solves: mixed, index, batch, stream")
    ap.add_argument("--stream-rates", default="200,1000,2000,4000,8000",
                    help="offered request rates (req/s) of the stream section")
    a = ap.parse_args()
    sections = set(x for x in a.sections.split(",") if x)
    if a.gpus != 1:
        print("bench.py measures one MI355X only: multi-GPU runs are not part of this build "
              "(run with --gpus 1)", file=sys.stderr)
        sys.exit(2)
    if a.copy_calibration:
        copy_calibration()
        return
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

    for _ in range(a.warmup):
        solve()
    # HIP events around every launch of the dense factor + solve (the kernel
    # with the largest share of the solve's device time) on the context's stream
    ctx.probe(P.PROBE_FACTOR)
    ctx.probe_read()
    res = []
    t0 = time.perf_counter()
    for _ in range(a.steps):
        res.append(solve())  # returns after its final device read-back (stream drained)
    dt = time.perf_counter() - t0
    fac_ms, fac_n = ctx.probe_read()
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
        "device": P.device_info(0),
        "config": {"workload": "config 2: 88 cams (f,k1,k2+pose), %d pts, %d obs, LM to rel. decrease < 1e-8, "
                               "plain FP64" % (sc.n_pts, sc.n_obs),
                   "n_obs": sc.n_obs, "sum_t2": sc.sum_t2, "lm_iterations": r["itno"], "lm_trials": trials,
                   "final_rms_px": rms},
    }
    # Dense Cholesky of the reduced camera matrix (order n = C n_c) and the two
    # triangular solves: n^3 / 3 + 2 n^2 flops per launch (n, not the padded
    # storage order)
    n = float(info["n"])
    flops = n ** 3 / 3.0 + 2.0 * n * n
    # the device-controlled solve queues one trial behind the one that stops
    # it, and that launch ends before factoring: the average is over the
    # launches that factored (one per trial), their few microseconds included
    factored = min(fac_n, trials * a.steps)
    avg_s = fac_ms * 1e-3 / max(factored, 1)
    achieved = flops / avg_s / 1e12
    traffic, traffic_src = profiled_traffic("k_chol_solve")
    out["roofline"] = {"bound": "mfma", "achieved": achieved, "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                       "frac": achieved / FP64_PEAK_TFLOPS, "traffic": traffic,
                       "kernel": "k_chol_solve", "avg_launch_us": avg_s * 1e6, "launches": fac_n,
                       "launches_factored": factored, "flops_per_launch": flops, "n": int(n),
                       "traffic_source": traffic_src}
    log("dense factor+solve: %d launches (%d factored), %.1f us avg, %.3f TFLOP/s (%.2f%% of FP64 peak)" % (
        fac_n, factored, avg_s * 1e6, achieved, 100 * achieved / FP64_PEAK_TFLOPS))
    if a.headline_only:
        print(json.dumps(out), flush=True)
        return
    out["stages"] = stage_rates(P, ctx, solve_on, sc, info, r, dt / a.steps)
    out["path_roofline"] = path_roofline(sc, info, r, dt / a.steps)
    log("path roofline: T_roof %.2f ms vs measured %.2f ms per solve (%.1f%%)" % (
        out["path_roofline"]["t_roof_ms"], out["path_roofline"]["t_measured_ms"], 100 * out["path_roofline"]["frac"]))
    if "mixed" in sections:
        out["mixed"] = mixed_section(P, sc, a, dt / a.steps, rms)
    if "index" in sections:
        out["index"] = index_section(P, sc)
    if "batch" in sections:
        out["batch"] = batch_section(P)
    if "stream" in sections:
        out["stream"] = stream_section(P, rates=tuple(int(x) for x in a.stream_rates.split(",") if x))
    if not a.no_cpu:
        out["cpu_baseline"] = cpu_baseline(P, sc, trials)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
