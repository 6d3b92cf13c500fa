"""Benchmark: bubble-steps/s of the per-timestep foam path on one MI355X.

    python bench.py [--steps K] [--warmup W] [--workload foam3d_1e6] [--precision fp64]

A "step" is one accepted integrate() call -- predictor, pairwise forces and
gas exchange, corrector, adaptive dt, removal of dissolved bubbles and the
neighbour search when triggered -- over one synthetic foam resident in HBM.
value = live bubbles summed over the K timed accepted steps / wall time.
Prints one JSON line (rank 0, N = 1).
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "bubble-steps/s mixed+batched vs plain FP64 on GPU; p50/p99 request latency"
HBM_PEAK_GBS = 8000.0

WORKLOADS = {
    # name: (description, default steps)
    "foam2d_1e4": ("2D periodic foam, 10,000 bubbles, clipped-normal radii, phi 0.9", 1000),
    "foam3d_1e6": ("3D periodic foam, 1,000,000 bubbles, log-normal radii sigma 0.15, phi 0.84",
                   2000),
    "foam3d_1e7": ("3D periodic foam, 10,000,000 bubbles, log-normal radii sigma 0.15, phi 0.9",
                   500),
}


def kernel_bytes(dim, p):
    """Algorithmic (compulsory) HBM bytes per bubble of each kernel of an
    attempt, FP64 storage; p = half-list pairs per bubble."""
    c = dim + 1
    return {
        # read x, r, v, v_old; write x_p, r_p
        "predict": 8 * 4 * c,
        # read x_p, r_p, neighbour count, 2p neighbour indices; write v_p (incl. dR/dt part), overlap
        "pair": 8 * c + 4 + 8 * p + 8 * c + 8,
        # read x_p, x, v, v_p, saved (c each), overlap, path, error; write x_c, dR/dt, path_new,
        # error (+1 byte deletion mark)
        "correct": 8 * (5 * c + 3) + 8 * (c + 3) + 1,
    }


def step_bytes(dim, p):
    """Compulsory bytes per bubble-step of the whole step (read once, write once):
    read x,r, v, v_old, saved, path, error; write x,r, v_new, path, error, plus
    the 2p neighbour indices."""
    c = dim + 1
    return 8 * (6 * c + 4) + 8 * p


def cpu_baseline(cfg, budget_s):
    import oracle as ora

    threads = int(os.environ.get("OMP_NUM_THREADS") or os.cpu_count() or 1)
    ora.set_threads(threads)
    f = ora.Foam.synthetic(dim=cfg.dim, count=cfg.num_start, radius_dist=cfg.radius_dist,
                           r_mean=cfg.r_mean, r_std=cfg.r_std, r_lo=cfg.r_lo, r_hi=cfg.r_hi,
                           phi=cfg.phi, kappa=cfg.kappa, k_param=cfg.k_param, mu=cfg.mu,
                           sigma=cfg.sigma, tol=cfg.error_tolerance, dt0=cfg.dt0,
                           avg_rad=cfg.r_mean, seed=cfg.seed)
    elapsed, steps = 0.0, 0
    b0 = f.scalars()["bubble_steps"]
    while elapsed < budget_s and steps < 500:
        elapsed += f.run(2)
        steps += 2
    bs = f.scalars()["bubble_steps"] - b0
    return {"value": bs / elapsed, "unit": "bubble-steps/s", "cores": threads, "kind": "port",
            "sample": "%d accepted steps of the same workload (%d bubbles) on the OpenMP FP64 "
                      "oracle, %.1f s" % (steps, cfg.num_start, elapsed)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--workload", default="foam3d_1e6", choices=sorted(WORKLOADS))
    ap.add_argument("--precision", default="fp64", choices=["fp64", "mixed"])
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-budget", type=float, default=15.0)
    ap.add_argument("--kernel-reps", type=int, default=20)
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this build measures one GPU only (the path has no multi-GPU sharding); "
              "run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    import cubble_amd as ca

    desc, default_steps = WORKLOADS[args.workload]
    steps = args.steps if args.steps is not None else default_steps
    prec = ca.PREC_MIXED if args.precision == "mixed" else ca.PREC_FP64
    cfg = ca.Config.workload(args.workload, precision=prec)
    ip = ca.IntegrationParams(True, False, True)

    t_create = time.perf_counter()
    sim = ca.Simulator(cfg)
    t_create = time.perf_counter() - t_create
    print("bench: created %s in %.2f s" % (args.workload, t_create), file=sys.stderr, flush=True)
    if args.warmup > 0:
        sim.run_steps(args.warmup, ip)
    r0 = sim.scalars()
    t0 = time.perf_counter()
    r1 = sim.run_steps(steps, ip)
    t1 = time.perf_counter()
    elapsed = t1 - t0
    bubble_steps = r1["bubble_steps"] - r0["bubble_steps"]
    value = bubble_steps / elapsed
    print("bench: %d steps in %.3f s, %d attempts, %d searches, %d removals" %
          (steps, elapsed, r1["attempts"] - r0["attempts"], r1["searches"] - r0["searches"],
           r1["removals"] - r0["removals"]), file=sys.stderr, flush=True)

    sc = sim.scalars()
    n = sc["count"]
    p = sc["num_pairs"] / max(n, 1)
    kt = sim.time_kernels(args.kernel_reps, ip)
    kb = kernel_bytes(cfg.dim, p)
    times = {"predict": kt["predict_ms"], "pair": kt["pair_ms"], "correct": kt["correct_ms"]}
    dom = max(times, key=times.get)
    achieved = kb[dom] * n / (times[dom] * 1e-3) / 1e9
    roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                "kernel": "k_" + dom, "bytes_per_bubble": round(kb[dom], 2),
                "launch_ms": round(times[dom], 4),
                "kernels_ms": {k: round(v, 4) for k, v in times.items()},
                "step_bytes_per_bubble": round(step_bytes(cfg.dim, p), 2),
                "step_gbs": round(step_bytes(cfg.dim, p) * value / 1e9, 1)}

    out = {
        "metric": METRIC, "value": value, "unit": "bubble-steps/s", "n_gpus": 1, "steps": steps,
        "warmup": args.warmup, "ms_per_step": elapsed / steps * 1e3, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None,
        "dtype": "f64" if prec == ca.PREC_FP64 else "f64+f32",
        "data": "synthetic (seeded; uniform positions, box sized for exact phi)",
        "config": {"workload": args.workload, "description": desc, "bubbles": cfg.num_start,
                   "dim": cfg.dim, "phi": cfg.phi, "precision": args.precision,
                   "kappa": cfg.kappa, "K": cfg.k_param, "tol": cfg.error_tolerance,
                   "pairs_per_bubble": round(p, 3), "final_count": n,
                   "attempts": r1["attempts"] - r0["attempts"],
                   "searches": r1["searches"] - r0["searches"],
                   "removals": r1["removals"] - r0["removals"]},
        "roofline": roofline,
    }
    sim.close()
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(cfg, args.cpu_budget)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
