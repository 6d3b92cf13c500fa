"""The benchmark's workloads, each on its own terms.

* 3D foam, log-normal radii (sigma 0.15), phi 0.9, mixed precision (the
  workload of the 1e7-bubble error measurement, at 1.5e5 bubbles so that the
  FP64 oracle follows it step by step): 20 steps, each started from the same
  state on both sides, per-quantity error <= 1e-5 (scaled metric).
* The 256-foam mixed-precision batch (half 2D, half 3D, 1e3..1e5 bubbles):
  60 graph-replayed steps; every foam bit-equal to the same foam run alone.
* Serving: Poisson arrivals at 50/s, 3D, 1e3..1e6 bubbles, phi U[0.8, 0.95],
  mixed precision, each request 500 steps or half dissolved, capped here at
  40 steps: every result bit-equal to the same request run alone."""
import numpy as np
import pytest

import cubble_amd as ca
from parity import MIXED_TOL, attempt_errors, config, oracle_from_sim

pytestmark = pytest.mark.gpu

GAS = ca.IntegrationParams(True, False, True)


def test_mixed_3d_phi09_lognormal_stepwise_vs_oracle(ora):
    ora.set_threads(min(16, ora.max_threads()))
    n = 150_000
    cfg = config(3, n, dist=1, phi=0.9, seed=2, precision=ca.PREC_MIXED)
    print(f"workload mixed 3D phi 0.9 n={n}", flush=True)
    sim = ca.Simulator(cfg)
    worst = {}
    for step in range(20):
        f = oracle_from_sim(ora, sim)
        sim.step(GAS)
        f.attempt(True, True)
        errs = attempt_errors(3, sim.get_attempt(), f.attempt_out(), sim.scalars()["f0"])
        for q, e in errs.items():
            worst[q] = max(worst.get(q, 0.0), e)
        # advance the GPU foam to its next step (commit, removal, search)
        sim.integrate(GAS)
    print("mixed 3D phi 0.9 worst errors over 20 steps", worst, flush=True)
    for q, e in worst.items():
        assert e <= MIXED_TOL, (q, e)
    assert sim.scalars()["steps"] == 20
    sim.close()


def _same(a, b, what):
    sa, sca = a
    sb, scb = b
    for key in sa:
        assert np.array_equal(sa[key], sb[key]), (what, key)
    for key in ("count", "dt", "steps", "attempts", "searches", "removals", "bubble_steps"):
        assert sca[key] == scb[key], (what, key)


def test_batch256_mixed_each_foam_equals_alone():
    cfgs = ca.batch_workload(256, precision=ca.PREC_MIXED)
    print("workload batch256 mixed: %d bubbles" % sum(c.num_start for c in cfgs), flush=True)
    batch = ca.Batch(cfgs, tuning={"bubbles_per_thread": 1, "search_tile": 64, "graph_steps": 8})
    res = batch.run_steps(60, GAS)
    assert all(r["steps"] == 60 for r in res)
    assert sum(r["searches"] for r in res) > 256  # searches happened beyond the first
    for k, cfg in enumerate(cfgs):
        sim = ca.Simulator(cfg, tuning={"bubbles_per_thread": 1, "search_tile": 64,
                                        "graph_steps": 8})
        sim.run_steps(60, GAS)
        _same(batch.get_state(k), sim.get_state(), k)
        sim.close()
    batch.close()


def test_serving_poisson_50_per_s_each_result_equals_alone():
    reqs = ca.serving_requests(32, rate_per_s=50.0, precision=ca.PREC_MIXED)
    sizes = [c.num_start for c, _ in reqs]
    print("workload serving: 32 requests, %d..%d bubbles" % (min(sizes), max(sizes)), flush=True)
    assert max(sizes) > 1e5 and min(sizes) < 1e4
    steps = 40
    srv = ca.Server(hbm_budget_bytes=32 << 30, round_steps=10, keep_states=True,
                    precision=ca.PREC_MIXED)
    ids = [srv.submit(c, max_steps=steps, dissolve_fraction=0.5, arrival_s=t) for c, t in reqs]
    stats = srv.run()
    assert stats["requests"] == len(reqs)
    assert stats["max_resident"] >= 2
    for rid, (cfg, t) in zip(ids, reqs):
        r = srv.result(rid)
        assert r["status"] == 2 and r["steps"] == steps
        assert r["admit_s"] >= r["arrival_s"] >= t - 1e-9
        sim = ca.Simulator(cfg)
        sim.run_steps(steps, GAS)
        _same(srv.get_state(rid), sim.get_state(), rid)
        sim.close()
    srv.close()
