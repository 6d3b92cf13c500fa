"""Benchmark of the per-timestep foam path on one MI355X.

    python bench.py [--steps K] [--warmup W] [--legs batch,plain,foam,serve]

A "step" is one accepted integrate() call -- predictor, pairwise forces and
gas exchange, corrector, adaptive dt, removal of dissolved bubbles and the
neighbour search when triggered -- of every foam of the workload, with the
foams resident in HBM.  Throughput is live bubbles summed over accepted steps
per second of wall time.

The headline (`value`) is the batched mixed-precision form: 256 independent
foams (half 2D, half 3D, 1e3..1e5 bubbles each) advanced together by replayed
HIP graphs with autotuned launch shapes, K steps each.  The same line carries:
  plain_fp64_same_foams  the same 256 foams, FP64, each run alone, one after
                         another (the plain path), and the speedup over it;
  foam3d_1e6_fp64        one 3D foam of 1e6 bubbles, FP64, 2000 steps, with
                         the roofline of its dominant kernel;
  serving                Poisson requests (50/s, 3D, 1e3..1e6 bubbles, 500
                         steps or half dissolved) through the continuous-
                         batching request loop: p50 / p99 latency, throughput;
  mixed_precision_error  the 3D 1e7-bubble foam: per-quantity error of a
                         mixed-precision attempt against FP64 from the same
                         state, and the drift of a 500-step mixed run; the
                         same per-step error on the batch's largest 2D foam;
  cpu_baseline           the OpenMP FP64 restatement on the host cores, on a
                         bounded sample of the headline workload.
Prints one JSON line.
"""
import argparse
import dataclasses
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "bubble-steps/s mixed+batched vs plain FP64 on GPU; p50/p99 request latency"
HBM_PEAK_GBS = 8000.0


def log(msg):
    print("bench: " + msg, file=sys.stderr, flush=True)


def kernel_bytes(dim, p, mixed=False):
    """Algorithmic (compulsory) HBM bytes per bubble of each kernel of an
    attempt; p = half-list pairs per bubble (each bubble reads 2p list
    entries).  Positions and radii are FP64 in both precision modes; mixed
    precision stores velocities in FP32, and its predictor writes the 16-byte
    fixed-point neighbour record instead of the 32-byte FP64 one (the pair
    kernel reads only that, the corrector recomputes the prediction)."""
    c = dim + 1
    rec = 16 if mixed else 8 * c
    # bytes per velocity component, saved position, error and overlap value
    # (FP32 in mixed precision)
    vb = 4 if mixed else 8
    return {
        # read x, r, v, v_old; write the predicted {x, y, z, r} record (in
        # mixed precision only the 16-byte fixed-point record)
        "predict": 8 * c + 2 * vb * c + rec,
        # read own record, neighbour count, 2p neighbour indices (the gathered
        # neighbour records are other bubbles' own records: L2 hits when the
        # lists are cell ordered); write v_p incl. the pair dR/dt, overlap
        "pair": rec + 4 + 8 * p + vb * c + vb,
        # read predicted record (mixed precision: v_old, to recompute it), x, v,
        # v_p, saved (c each), overlap, path, error; write corrected x, dR/dt,
        # path, error (+1 byte deletion mark)
        "correct": (vb if mixed else 8) * c + 8 * c + 2 * vb * c + vb * c + 2 * vb + 8
                   + 8 * c + vb + vb + 9,
    }


def step_bytes(dim, p, mixed=False):
    """Compulsory bytes per bubble-step of a whole accepted step: read x,r, v,
    v_old, saved, path, error; write x,r, v_new, path, error; the 2p list
    entries.  Velocities, saved positions and error are FP32 in mixed
    precision."""
    c = dim + 1
    vb = 4 if mixed else 8
    return 8 * 2 * c + vb * 4 * c + 8 * 2 + vb * 2 + 8 * p


KERNEL_NAMES = {"predict": "k_predict", "pair": "k_pair", "correct": "k_correct",
                "search": "k_s_search"}


def kernel_name(part, mixed):
    return "k_pair_mixed" if part == "pair" and mixed else KERNEL_NAMES[part]


def search_bytes(dim, p):
    """Compulsory HBM bytes per bubble of the list kernel k_s_search: read
    the bubble's {x, y, z, r} record once (the candidates' records are
    neighbours' own records, L2 hits in cell order), write its upper list (p
    entries on average) and one lower-count increment per entry."""
    return 8 * (dim + 1) + 4 * p + 4 * p


def roofline(kt, st, per_foam, steps, attempts, searches, groups, mixed=False):
    """Roofline of the kernel with the largest share of a step.  kt: attempt
    kernel times (ms) of one launch over all foams; st: search timings of one
    search of every foam (cubble_time_search); per_foam: [(dim, count,
    pairs)]; steps, attempts, searches: counts over the timed run (searches
    summed over foams), groups: attempt groups replayed.  Shares: the attempt
    kernels run once per attempt, the list kernel once per search (its time
    for all foams scaled by the searched bubbles' share), plus the idle
    launches of the phase kernels once per group."""
    bubbles = sum(n for _, n, _ in per_foam)
    nf = max(len(per_foam), 1)
    times = {"predict": kt["predict_ms"], "pair": kt["pair_ms"], "correct": kt["correct_ms"],
             "search": st["list_kernel_ms"]}
    share = {k: times[k] * attempts / nf / max(steps, 1) for k in ("predict", "pair", "correct")}
    share["search"] = st["list_kernel_ms"] * searches / nf / max(steps, 1)
    dom = max(share, key=share.get)
    nbytes = 0.0
    for dim, n, pairs in per_foam:
        p = pairs / max(n, 1)
        nbytes += (search_bytes(dim, p) if dom == "search" else kernel_bytes(dim, p, mixed)[dom]) * n
    achieved = nbytes / (times[dom] * 1e-3) / 1e9
    return {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
            "kernel": kernel_name(dom, mixed), "bytes_per_launch": round(nbytes),
            "bubbles_per_launch": bubbles,
            "launch_ms": round(times[dom], 4),
            "kernels_ms": {k: round(v, 4) for k, v in times.items()},
            "ms_per_step_share": {k: round(v, 4) for k, v in share.items()},
            "search": {k: (round(v, 4) if isinstance(v, float) else v) for k, v in st.items()},
            "idle_phase_ms_per_group": round(st["idle_ms"], 4),
            "idle_phase_ms_per_step": round(st["idle_ms"] * groups / max(steps, 1), 4)}


def attempt_rooflines(kt, per_foam, mixed=False):
    """HBM fraction of each attempt kernel (algorithmic bytes / launch time)."""
    out = {}
    for part in ("predict", "pair", "correct"):
        nbytes = sum(kernel_bytes(d, pr / max(n, 1), mixed)[part] * n for d, n, pr in per_foam)
        gbs = nbytes / (kt[part + "_ms"] * 1e-3) / 1e9
        out[kernel_name(part, mixed)] = {"achieved": round(gbs, 1),
                                         "frac": round(gbs / HBM_PEAK_GBS, 4),
                                         "bytes_per_launch": round(nbytes),
                                         "launch_ms": round(kt[part + "_ms"], 4)}
    return out


def counters(before, after):
    keys = ("attempts", "searches", "removals")
    return {k: int(sum(a[k] for a in after) - sum(b[k] for b in before)) for k in keys}


def leg_batch(ca, args, ip):
    cfgs = ca.batch_workload(256, precision=ca.PREC_MIXED)
    t = time.perf_counter()
    b = ca.Batch(cfgs)
    log("batch: created 256 foams (%d bubbles) in %.1f s" %
        (sum(c.num_start for c in cfgs), time.perf_counter() - t))
    # warm up first: the shapes are then tuned on steps that include searches
    # and removals, as in the timed run
    if args.warmup > 0:
        b.run_steps(args.warmup, ip)
    t = time.perf_counter()
    if args.tuning:
        # fixed shapes (no timing trials): bubbles_per_thread,search_tile,graph_steps
        shapes = dict(zip(ca.TUNING_KEYS, [int(v) for v in args.tuning.split(",")]))
        b.retune(shapes)
        tune = {"chosen": shapes, "from_cache": False, "candidates_timed": 0, "fixed": True}
    else:
        # each trial runs as many steps as the timed run (at least 16, at most
        # --tune-steps): a short run starts while searches are frequent and
        # replays end more often, and favours other shapes than a long one
        trial = max(16, min(args.tune_steps, args.steps))
        tune = b.autotune(ip, trial_steps=trial, repeats=2 if trial >= 64 else 3,
                          cache_path=args.tuning_cache)
        tune["trial_steps"] = trial
    # the device loop of the chosen shapes is captured before the clock starts
    b.run_steps(0, ip)
    log("batch: launch shapes in %.1f s: %s" % (time.perf_counter() - t, json.dumps(tune)))
    s0 = [b.scalars(k) for k in range(len(cfgs))]
    t0 = time.perf_counter()
    res = b.run_steps(args.steps, ip)
    elapsed = time.perf_counter() - t0
    s1 = [b.scalars(k) for k in range(len(cfgs))]
    work = sum(x["bubble_steps"] for x in s1) - sum(x["bubble_steps"] for x in s0)
    assert all(r["steps"] >= args.steps for r in res)
    kt = b.time_kernels(args.kernel_reps, ip)
    st = b.time_search(args.kernel_reps)
    per_foam = [(x["dim"], x["count"], x["num_pairs"]) for x in s1]
    cn = counters(s0, s1)
    groups = max(x["attempts"] for x in s1) - max(x["attempts"] for x in s0)
    rf = roofline(kt, st, per_foam, args.steps, cn["attempts"], cn["searches"], groups, mixed=True)
    rf["attempt_kernels"] = attempt_rooflines(kt, per_foam, mixed=True)
    live = sum(x["count"] for x in s1)
    pairs = sum(x["num_pairs"] for x in s1)
    rf["step_bytes_per_bubble"] = round(
        sum(step_bytes(x["dim"], x["num_pairs"] / max(x["count"], 1), True) * x["count"]
            for x in s1) /
        max(live, 1), 2)
    b.close()
    log("batch: %d steps in %.3f s -> %.4g bubble-steps/s" % (args.steps, elapsed, work / elapsed))
    return {"value": work / elapsed, "elapsed_s": elapsed, "bubbles_start": sum(c.num_start for c in cfgs),
            "bubbles_end": live, "pairs_per_bubble": round(pairs / max(live, 1), 3),
            "tuning": tune, "counters": counters(s0, s1), "roofline": rf}, cfgs


def leg_plain(ca, args, ip, cfgs):
    """The same foams in FP64, each alone, one after another (creation and
    the first search untimed)."""
    work = 0.0
    elapsed = 0.0
    for k, c in enumerate(cfgs):
        c = dataclasses.replace(c, precision=ca.PREC_FP64)
        sim = ca.Simulator(c)
        if args.warmup > 0:
            sim.run_steps(args.warmup, ip)
        b0 = sim.scalars()["bubble_steps"]
        t0 = time.perf_counter()
        sim.run_steps(args.steps, ip)
        elapsed += time.perf_counter() - t0
        work += sim.scalars()["bubble_steps"] - b0
        sim.close()
    log("plain: 256 foams alone in %.2f s -> %.4g bubble-steps/s" % (elapsed, work / elapsed))
    return {"value": work / elapsed, "elapsed_s": elapsed, "precision": "fp64",
            "how": "each foam alone (cubble_create + cubble_run_steps), one after another"}


def leg_foam(ca, args, ip):
    cfg = ca.Config.workload("foam3d_1e6")
    t = time.perf_counter()
    sim = ca.Simulator(cfg)
    log("foam3d_1e6: created in %.1f s" % (time.perf_counter() - t))
    if args.warmup > 0:
        sim.run_steps(args.warmup, ip)
    r0 = sim.scalars()
    t0 = time.perf_counter()
    sim.run_steps(args.foam_steps, ip)
    elapsed = time.perf_counter() - t0
    r1 = sim.scalars()
    work = r1["bubble_steps"] - r0["bubble_steps"]
    kt = sim.time_kernels(args.kernel_reps, ip)
    st = sim.time_search(args.kernel_reps)
    cn = counters([r0], [r1])
    per_foam = [(3, r1["count"], r1["num_pairs"])]
    rf = roofline(kt, st, per_foam, args.foam_steps, cn["attempts"], cn["searches"],
                  cn["attempts"])
    rf["attempt_kernels"] = attempt_rooflines(kt, per_foam)
    p = r1["num_pairs"] / max(r1["count"], 1)
    rf["step_bytes_per_bubble"] = round(step_bytes(3, p), 2)
    rf["step_gbs"] = round(step_bytes(3, p) * work / elapsed / 1e9, 1)
    sim.close()
    log("foam3d_1e6: %d steps in %.3f s -> %.4g bubble-steps/s" %
        (args.foam_steps, elapsed, work / elapsed))
    return {"value": work / elapsed, "steps": args.foam_steps,
            "ms_per_step": elapsed / args.foam_steps * 1e3, "precision": "fp64",
            "pairs_per_bubble": round(p, 3), "final_count": r1["count"],
            "counters": counters([r0], [r1]), "roofline": rf}


def cpu_info():
    """The host's CPU model, logical CPUs, CPUs this process may run on, and
    the OpenMP thread count in use."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        avail = len(os.sched_getaffinity(0))
    except AttributeError:
        avail = os.cpu_count()
    return {"model": model, "nproc": os.cpu_count(), "available": avail,
            "omp_threads": int(os.environ.get("OMP_NUM_THREADS") or avail or 1)}


def median(v):
    v = sorted(v)
    return v[len(v) // 2]


def leg_foam2d(ca, args, ip):
    """The 2D foam of 1e4 bubbles (normal radii 1 +- 0.1 clipped to [0.5, 1.5],
    phi 0.9, FP64), 1000 accepted steps from creation, on the GPU and in the
    FP64 restatement on the host cores (all threads, and one), each the
    median of 3 runs; the same seed on both, with the end states compared."""
    import oracle as ora

    cfg = ca.Config.workload("foam2d_1e4")
    steps = args.foam2d_steps
    gpu, end = [], None
    for rep in range(3):
        sim = ca.Simulator(cfg)
        b0 = sim.scalars()["bubble_steps"]
        t0 = time.perf_counter()
        sim.run_steps(steps, ip)
        el = time.perf_counter() - t0
        sc = sim.scalars()
        gpu.append((sc["bubble_steps"] - b0) / el)
        if end is None:
            st, _ = sim.get_state()
            end = {"count": sc["count"], "mean_r": float(st["x"][3].mean()), "time":
                   sc["time_int"] + sc["time_frac"], "searches": sc["searches"],
                   "attempts": sc["attempts"]}
        sim.close()
    info = cpu_info()
    cpu = {}
    oend = None
    for threads in sorted({info["omp_threads"], 1}, reverse=True):
        ora.set_threads(threads)
        rates = []
        for rep in range(3):
            f = ora.Foam.synthetic(dim=cfg.dim, count=cfg.num_start, radius_dist=cfg.radius_dist,
                                   r_mean=cfg.r_mean, r_std=cfg.r_std, r_lo=cfg.r_lo,
                                   r_hi=cfg.r_hi, phi=cfg.phi, kappa=cfg.kappa,
                                   k_param=cfg.k_param, mu=cfg.mu, sigma=cfg.sigma,
                                   tol=cfg.error_tolerance, dt0=cfg.dt0, avg_rad=cfg.r_mean,
                                   seed=cfg.seed)
            b0 = f.scalars()["bubble_steps"]
            el = f.run(steps)
            sc = f.scalars()
            rates.append((sc["bubble_steps"] - b0) / el)
            if oend is None:
                ost, _ = f.get_state()
                oend = {"count": sc["count"], "mean_r": float(ost["x"][3].mean()),
                        "time": sc["time_int"] + sc["time_frac"], "searches": sc["searches"],
                        "attempts": sc["attempts"]}
        cpu["threads_%d" % threads] = {"value": median(rates), "runs": rates, "threads": threads}
    g = median(gpu)
    log("foam2d_1e4: GPU %.4g, CPU %s bubble-steps/s" %
        (g, {k: "%.4g" % v["value"] for k, v in cpu.items()}))
    return {"workload": "2D periodic foam, 10,000 bubbles, normal radii (1, 0.1) clipped to "
                        "[0.5, 1.5], phi 0.9, kappa 0.01, K 0.1, FP64, %d steps" % steps,
            "gpu": {"value": g, "runs": gpu, "unit": "bubble-steps/s"},
            "cpu_oracle": cpu, "cpu": info,
            "gpu_over_cpu_all_threads": round(g / cpu["threads_%d" % info["omp_threads"]]["value"], 2),
            "end_state": {"gpu": end, "oracle": oend,
                          "mean_r_rel_diff": abs(end["mean_r"] / oend["mean_r"] - 1.0)}}


def scaled_err(gpu, ref, scale):
    """max_i |gpu_i - ref_i| / max(|ref_i|, scale): relative error, judged
    against the quantity's scale where its value cancels to about zero."""
    import numpy as np

    gpu = np.asarray(gpu, float)
    ref = np.asarray(ref, float)
    if gpu.size == 0:
        return 0.0
    return float(np.max(np.abs(gpu - ref) / np.maximum(np.abs(ref), scale)))


def leg_mixed_error(ca, args, ip):
    """Mixed precision against FP64 on the 3D 1e7-bubble foam (phi 0.9).
    Per step: the FP64 state after `mixed_steps` steps is loaded into a
    mixed-precision simulator and one attempt of each is compared, quantity
    by quantity.  Over the run: the mixed and FP64 runs of the same seed,
    `mixed_steps` steps each, compared on count, mean radius and gas
    volume."""
    import numpy as np

    c64 = ca.Config.workload("foam3d_1e7")
    cmx = dataclasses.replace(c64, precision=ca.PREC_MIXED)
    a = ca.Simulator(c64)
    b = ca.Simulator(cmx)
    a.run_steps(args.mixed_steps, ip)
    b.run_steps(args.mixed_steps, ip)
    sa, sca = a.get_state()
    sb, scb = b.get_state()

    def vol(r):
        return float(np.sum(4.0 / 3.0 * np.pi * np.asarray(r) ** 3))

    traj = {"count_fp64": int(sca["count"]), "count_mixed": int(scb["count"]),
            "mean_radius_rel_diff": abs(float(np.mean(sb["x"][3])) / float(np.mean(sa["x"][3])) - 1),
            "gas_volume_rel_diff": abs(vol(sb["x"][3]) / vol(sa["x"][3]) - 1)}
    b.set_state(sa, sca)
    per_step = attempt_error(a, b, 3, sca["f0"], ip)
    a.close()
    b.close()
    # the batch's largest 2D foam, where 2D overlap terms near tangency and
    # nearly coincident centres of freshly seeded foams test the mixed path
    c2 = max((c for c in ca.batch_workload(256, precision=ca.PREC_FP64) if c.dim == 2),
             key=lambda c: c.num_start)
    per_step_2d = {}
    for steps in (3, 50):
        a = ca.Simulator(c2)
        b = ca.Simulator(dataclasses.replace(c2, precision=ca.PREC_MIXED))
        a.run_steps(steps, ip)
        sa2, sc2 = a.get_state()
        b.set_state(sa2, sc2)
        per_step_2d["after_%d_steps" % steps] = attempt_error(a, b, 2, sc2["f0"], ip)
        a.close()
        b.close()
    log("mixed error: %s %s 2D %s" % (json.dumps(per_step), json.dumps(traj),
                                      json.dumps(per_step_2d)))
    return {"workload": "3D periodic foam, 10,000,000 bubbles, log-normal radii sigma 0.15, "
                        "phi 0.9", "steps": args.mixed_steps,
            "per_step_vs_fp64": per_step, "metric": "max |mixed - fp64| / max(|fp64|, scale); "
            "scale = f0 for velocities, max|q| for dR/dt and overlap, 1 for positions and radii",
            "run_vs_fp64": traj,
            "per_step_vs_fp64_2d": {
                "workload": "the batch's largest 2D foam: %d bubbles, normal radii, phi %.3f, "
                            "FP64 state after 3 and after 50 steps" % (c2.num_start, c2.phi),
                **per_step_2d}}


def attempt_error(a, b, dim, f0, ip):
    """One attempt of FP64 simulator `a` and mixed-precision simulator `b`
    from the same state, compared quantity by quantity (scaled metric)."""
    import numpy as np

    a.step(ip)
    b.step(ip)
    ga, gb = a.get_attempt(), b.get_attempt()
    s_r = max(float(np.max(np.abs(ga["vp"][3]))), 1e-300)
    s_ov = max(float(np.max(np.abs(ga["ov"]))), 1e-300)
    return {
        "velocity": max(scaled_err(gb["vp"][k], ga["vp"][k], f0) for k in range(dim)),
        "drdt": scaled_err(gb["vp"][3], ga["vp"][3], s_r),
        "overlap_area": scaled_err(gb["ov"], ga["ov"], s_ov),
        "position": max(scaled_err(gb["xc"][k], ga["xc"][k], 1.0) for k in range(dim)),
        "radius": scaled_err(gb["xc"][3], ga["xc"][3], 1.0),
    }


def leg_serve(ca, args, ip):
    """Poisson requests through the serving loop.  Offered load: the bubble-
    steps the requests ask for (count x 500 steps, at most; a request ends
    early once half its bubbles dissolved) per second of arrivals.  Queue
    stability: the latency of the first and the last quarter of the requests
    (a growing queue shows as a rising latency), and the time to drain after
    the last arrival."""
    import numpy as np

    reqs = ca.serving_requests(args.serve_requests, rate_per_s=50.0, precision=ca.PREC_MIXED)
    srv = ca.Server(hbm_budget_bytes=args.serve_budget_gib << 30, round_steps=args.serve_round,
                    precision=ca.PREC_MIXED)
    ids = [srv.submit(cfg, max_steps=500, dissolve_fraction=0.5, arrival_s=at) for cfg, at in reqs]
    st = srv.run()
    res = [srv.result(i) for i in ids]
    srv.close()
    log("serve: %s" % json.dumps(st))
    last = reqs[-1][1]
    lat = np.array([r["finish_s"] - r["arrival_s"] for r in res])
    q = max(1, len(res) // 4)
    offered = sum(c.num_start * 500 for c, _ in reqs) / last
    return {"requests": st["requests"], "arrival_rate_per_s": 50.0, "arrivals_s": round(last, 2),
            "p50_latency_s": st["p50_latency_s"], "p99_latency_s": st["p99_latency_s"],
            "throughput": st["throughput"], "unit": "bubble-steps/s", "wall_s": st["wall_s"],
            "offered_max_bubble_steps_per_s": offered,
            "served_bubble_steps_per_s_of_arrivals": st["bubble_steps"] / last,
            "latency_first_quarter_p50_s": float(np.median(lat[:q])),
            "latency_last_quarter_p50_s": float(np.median(lat[-q:])),
            "drain_after_last_arrival_s": round(st["wall_s"] - last, 3),
            "max_resident": st["max_resident"], "peak_hbm_bytes": st["peak_bytes"],
            "hbm_budget_bytes": args.serve_budget_gib << 30, "round_steps": args.serve_round,
            "precision": "mixed"}


COPY_MIB = 256


def device_copy(mib, reps):
    """A plain device-to-device copy of known size (HIP's copy kernel): the
    reference reading that scales the FETCH_SIZE / WRITE_SIZE counters."""
    import ctypes

    hip = ctypes.CDLL("libamdhip64.so")
    n = ctypes.c_size_t(mib << 20)
    a, b = ctypes.c_void_p(), ctypes.c_void_p()
    for p in (a, b):
        if hip.hipMalloc(ctypes.byref(p), n) != 0:
            raise RuntimeError("hipMalloc failed")
    hip.hipMemset(a, 1, n)
    hip.hipDeviceSynchronize()
    for _ in range(reps):
        if hip.hipMemcpy(b, a, n, 3) != 0:  # hipMemcpyDeviceToDevice
            raise RuntimeError("hipMemcpy failed")
    hip.hipDeviceSynchronize()
    hip.hipFree(a)
    hip.hipFree(b)


def leg_probe(ca, args, ip, foam=False):
    """Child of the counter passes: a few launches of the attempt kernels of
    the headline workload (or, with foam=True, of the 1e6-bubble FP64 foam),
    then the reference copy."""
    if foam:
        h = ca.Simulator(ca.Config.workload("foam3d_1e6"))
    else:
        h = ca.Batch(ca.batch_workload(256, precision=ca.PREC_MIXED))
    h.time_kernels(3, ip)
    h.close()
    device_copy(COPY_MIB, 3)


def read_counters(folder, counter):
    """{kernel name: [value per dispatch]} from rocprofv3's counter CSV."""
    import csv
    import glob

    out = {}
    for path in glob.glob(os.path.join(folder, "**", "*counter_collection.csv"), recursive=True):
        with open(path) as f:
            for row in csv.DictReader(f):
                if row.get("Counter_Name") != counter:
                    continue
                # "ns::kernel(args)" -> "kernel"
                # "void ns::k_pair_mixed<false>(...)" -> "k_pair_mixed"
                name = row["Kernel_Name"].split("(")[0].split("<")[0].split("::")[-1]
                out.setdefault(name, []).append(float(row["Counter_Value"]))
    return out


ATTEMPT_KERNELS = ("k_predict", "k_pair_mixed", "k_correct")


def measure_traffic(kernels=ATTEMPT_KERNELS):
    """HBM bytes per launch of each of `kernels` over the headline workload,
    from two rocprofv3 --pmc passes over a child run of this script
    (FETCH_SIZE, then WRITE_SIZE; KiB per dispatch), each scaled by what the
    same counter reads for the copy of COPY_MIB MiB.  Returns ({kernel: bytes},
    {kernel: counter details}) or (None, reason).  Run before this process
    touches the GPU."""
    import shutil
    import subprocess
    import tempfile

    exe = shutil.which("rocprofv3") or "/opt/rocm/bin/rocprofv3"
    if not os.path.exists(exe):
        return None, "rocprofv3 not found"
    env = dict(os.environ, TMPDIR="/tmp")
    res = {k: {} for k in kernels}
    for counter in ("FETCH_SIZE", "WRITE_SIZE"):
        folder = tempfile.mkdtemp(prefix="cubble_pmc_", dir="/tmp")
        cmd = [exe, "--pmc", counter, "--output-format", "csv", "-d", folder, "-o", "pmc", "--",
               sys.executable, os.path.abspath(__file__), "--legs", "probe",
               "--no-cpu-baseline", "--no-traffic"]
        log("traffic: " + " ".join(cmd))
        try:
            r = subprocess.run(cmd, cwd="/tmp", env=env, timeout=90, capture_output=True)
        except subprocess.TimeoutExpired:
            return None, "rocprofv3 %s pass timed out" % counter
        vals = read_counters(folder, counter)
        shutil.rmtree(folder, ignore_errors=True)
        copy = [v for k, v in vals.items() if "copy" in k.lower()]
        missing = [k for k in kernels if not vals.get(k)]
        if r.returncode != 0:
            return None, "rocprofv3 %s pass failed (rc %d)" % (counter, r.returncode)
        if not copy or missing:
            return None, "%s pass: no reading for %s or the copy kernel" % (counter, missing)
        copy_kib = max(max(c) for c in copy)
        scale = (COPY_MIB << 20) / (copy_kib * 1024.0)
        for k in kernels:
            mine = vals[k]
            res[k][counter] = {"kib_per_dispatch": sum(mine) / len(mine), "copy_kib": copy_kib,
                               "scale": scale}
    total = {k: sum(v["kib_per_dispatch"] * 1024.0 * v["scale"] for v in res[k].values())
             for k in kernels}
    return total, res


def cpu_baseline(cfgs, steps, budget_s):
    """The OpenMP FP64 restatement on the host cores over the headline
    workload's foams in order, `steps` steps each, until the time budget."""
    import oracle as ora

    threads = int(os.environ.get("OMP_NUM_THREADS") or os.cpu_count() or 1)
    ora.set_threads(threads)
    elapsed, work, done, partial = 0.0, 0.0, 0, 0
    for c in cfgs:
        f = ora.Foam.synthetic(dim=c.dim, count=c.num_start, radius_dist=c.radius_dist,
                               r_mean=c.r_mean, r_std=c.r_std, r_lo=c.r_lo, r_hi=c.r_hi,
                               phi=c.phi, kappa=c.kappa, k_param=c.k_param, mu=c.mu,
                               sigma=c.sigma, tol=c.error_tolerance, dt0=c.dt0,
                               avg_rad=c.r_mean, seed=c.seed)
        b0 = f.scalars()["bubble_steps"]
        n = 0
        while n < steps and elapsed < budget_s:
            elapsed += f.run(10)
            n += 10
        work += f.scalars()["bubble_steps"] - b0
        if n >= steps:
            done += 1
        else:
            partial = n
            break
    return {"value": work / elapsed, "unit": "bubble-steps/s", "cores": threads, "kind": "port",
            "cpu": cpu_info(),
            "sample": "first %d foams of the 256-foam workload, %d steps each%s, FP64 OpenMP "
                      "restatement, %.1f s" % (done, steps, " (+%d steps of the next)" % partial
                                               if partial else "", elapsed)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000, help="timed steps of the batched workload")
    ap.add_argument("--warmup", type=int, default=50)
    ap.add_argument("--legs", default="batch,plain,foam,foam2d,serve,mixed-error")
    ap.add_argument("--mixed-steps", type=int, default=500)
    ap.add_argument("--foam2d-steps", type=int, default=1000)
    ap.add_argument("--foam-steps", type=int, default=2000)
    # trials long enough to see the search costs of a coarsening foam: at 64
    # steps the tuner preferred 2 bubbles per thread, which is about 1%
    # slower than 1 over the 1000 timed steps
    ap.add_argument("--tune-steps", type=int, default=192)
    ap.add_argument("--tuning-cache", default=None)
    ap.add_argument("--tuning", default=None,
                    help="fixed launch shapes bubbles_per_thread,search_tile,graph_steps "
                         "instead of autotuning")
    ap.add_argument("--serve-requests", type=int, default=500)
    ap.add_argument("--serve-budget-gib", type=int, default=64)
    ap.add_argument("--serve-round", type=int, default=50)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-traffic", action="store_true",
                    help="skip the two rocprofv3 counter passes behind roofline.traffic")
    ap.add_argument("--cpu-budget", type=float, default=15.0)
    ap.add_argument("--kernel-reps", type=int, default=20)
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this build measures one GPU only (the path has no multi-GPU sharding); "
              "run with --gpus 1", file=sys.stderr)
        sys.exit(2)
    legs = set(args.legs.split(","))

    import cubble_amd as ca

    ip = ca.IntegrationParams(True, False, True)
    if "probe" in legs or "probe-foam" in legs:
        leg_probe(ca, args, ip, foam="probe-foam" in legs)
        return
    traffic = (None, "not measured")
    if "batch" in legs and not args.no_traffic:
        # counter passes first: this process has not touched the GPU yet
        traffic = measure_traffic()
        log("traffic: %s" % json.dumps(traffic))
    out = {"metric": METRIC, "unit": "bubble-steps/s", "n_gpus": 1, "steps": args.steps,
           "warmup": args.warmup, "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None, "data": "synthetic (seeded; uniform positions, box sized for "
                                        "exact phi; random radii)"}
    cfgs = ca.batch_workload(256, precision=ca.PREC_MIXED)
    if "batch" in legs:
        b, cfgs = leg_batch(ca, args, ip)
        out.update({"value": b["value"], "ms_per_step": b["elapsed_s"] / args.steps * 1e3,
                    "dtype": "f64 state, f32 pair math",
                    "config": {"workload": "batch256_mixed",
                               "description": "256 independent foams, half 2D (normal radii) and "
                                              "half 3D (log-normal radii), bubble counts "
                                              "log-uniform in [1e3, 1e5], phi U[0.8, 0.95], "
                                              "kappa logU[1e-3, 1e-1], K 0.1, mixed precision, "
                                              "autotuned launch shapes, graph-replayed",
                               "foams": 256, "bubbles_start": b["bubbles_start"],
                               "bubbles_end": b["bubbles_end"],
                               "pairs_per_bubble": b["pairs_per_bubble"],
                               "tuning": b["tuning"]["chosen"],
                               "autotune": {k: v for k, v in b["tuning"].items() if k != "chosen"},
                               "counters": b["counters"]},
                    "roofline": b["roofline"]})
        rf = out["roofline"]
        if traffic[0] is not None and rf["kernel"] in traffic[0]:
            rf["traffic"] = round(traffic[0][rf["kernel"]])
            rf["traffic_counters"] = traffic[1][rf["kernel"]]
            rf["traffic_all_attempt_kernels"] = {k: round(v) for k, v in traffic[0].items()}
        else:
            rf["traffic_note"] = traffic[1] if traffic[0] is None else \
                "no counters for %s" % rf["kernel"]
    if "plain" in legs:
        pl = leg_plain(ca, args, ip, cfgs)
        out["plain_fp64_same_foams"] = pl
        if "value" in out:
            out["speedup_vs_plain_fp64"] = round(out["value"] / pl["value"], 3)
    if "foam" in legs:
        out["foam3d_1e6_fp64"] = leg_foam(ca, args, ip)
    if "foam2d" in legs:
        out["foam2d_1e4_fp64_gpu_vs_cpu"] = leg_foam2d(ca, args, ip)
    if "serve" in legs:
        out["serving"] = leg_serve(ca, args, ip)
    if "mixed-error" in legs:
        out["mixed_precision_error"] = leg_mixed_error(ca, args, ip)
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(cfgs, args.steps, args.cpu_budget)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
