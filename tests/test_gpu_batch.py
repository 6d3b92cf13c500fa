"""Batched foams: each foam's result in a batch equals the same foam run alone
(bit for bit: the per-foam work, launch shape and reduction order do not
depend on the other foams), for the host-driven integrate() and the
graph-replayed device loop, in FP64 and mixed precision."""
import numpy as np
import pytest

import cubble_amd as ca
from parity import config

pytestmark = pytest.mark.gpu

GAS = ca.IntegrationParams(True, False, True)


def mixed_configs(precision=ca.PREC_FP64):
    specs = [(2, 3000, 0.85, 0.01), (3, 4000, 0.84, 0.003), (2, 1500, 0.92, 0.05),
             (3, 9000, 0.9, 0.02), (2, 20000, 0.8, 0.001), (3, 1200, 0.95, 0.1)]
    return [config(d, n, phi=phi, kappa=kappa, seed=100 + k, precision=precision)
            for k, (d, n, phi, kappa) in enumerate(specs)]


def assert_same(a, b):
    sa, sca = a
    sb, scb = b
    for key in sa:
        assert np.array_equal(sa[key], sb[key]), key
    for key in ("count", "dt", "steps", "attempts", "searches", "removals", "bubble_steps",
                "time_int", "time_frac", "num_pairs"):
        assert sca[key] == scb[key], key


@pytest.mark.parametrize("precision", [ca.PREC_FP64, ca.PREC_MIXED])
def test_batch_run_steps_equals_alone(precision):
    cfgs = mixed_configs(precision)
    batch = ca.Batch(cfgs)
    assert len(batch) == len(cfgs)
    res = batch.run_steps(40, GAS)
    assert all(r["steps"] == 40 for r in res)
    for k, cfg in enumerate(cfgs):
        sim = ca.Simulator(cfg)
        sim.run_steps(40, GAS)
        assert_same(batch.get_state(k), sim.get_state())
        sim.close()
    # keep going: a second replay batch stays identical
    batch.run_steps(17, GAS)
    sim = ca.Simulator(cfgs[3])
    sim.run_steps(57, GAS)
    assert_same(batch.get_state(3), sim.get_state())
    batch.close()


@pytest.mark.parametrize("precision", [ca.PREC_FP64, ca.PREC_MIXED])
def test_batch_with_wall_foams_equals_alone(precision):
    """Foams with walls and periodic ones in one batch: the launch takes the
    corrector variant with the wall code for all of them, and every foam
    still equals its run alone (where a periodic foam takes the variant
    without it), bit for bit."""
    print(f"batch with walls precision={precision}", flush=True)
    specs = [(2, 3000, (0, 0, 0)), (2, 2500, (1, 1, 0)), (3, 4000, (1, 0, 1)),
             (3, 3500, (0, 0, 0))]
    cfgs = [config(d, n, seed=200 + k, precision=precision, wall=w)
            for k, (d, n, w) in enumerate(specs)]
    batch = ca.Batch(cfgs)
    batch.run_steps(30, GAS)
    for k, cfg in enumerate(cfgs):
        sim = ca.Simulator(cfg)
        sim.run_steps(30, GAS)
        assert_same(batch.get_state(k), sim.get_state())
        sim.close()
    batch.close()


@pytest.mark.parametrize("precision", [ca.PREC_FP64, ca.PREC_MIXED])
def test_batch_with_flow_equals_alone(precision):
    """The imposed flow in a batch (each foam its own flow box and velocity,
    one with walls): every foam equals its run alone, bit for bit."""
    print(f"batch with flow precision={precision}", flush=True)
    flow = ca.IntegrationParams(True, True, True)
    specs = [(2, 3000, (0.5, -0.25, 0.0), (0, 0, 0)), (3, 4000, (0.2, 0.1, -0.3), (0, 0, 0)),
             (2, 2500, (-0.4, 0.3, 0.0), (1, 0, 0))]
    cfgs = [config(d, n, seed=300 + k, precision=precision, wall=w, flow=1, flow_vel=v,
                   flow_lbb=(0.3, 0.35, 0.3), flow_tfr=(0.7, 0.6, 0.65))
            for k, (d, n, v, w) in enumerate(specs)]
    batch = ca.Batch(cfgs)
    batch.run_steps(30, flow)
    for k, cfg in enumerate(cfgs):
        sim = ca.Simulator(cfg)
        sim.run_steps(30, flow)
        assert_same(batch.get_state(k), sim.get_state())
        sim.close()
    batch.close()


def test_batch_integrate_equals_alone():
    cfgs = mixed_configs()[:4]
    batch = ca.Batch(cfgs)
    sims = [ca.Simulator(c) for c in cfgs]
    for _ in range(12):
        res = batch.integrate(GAS)
        for s in sims:
            s.integrate(GAS)
    for k, s in enumerate(sims):
        assert res[k]["steps"] == 12
        assert_same(batch.get_state(k), s.get_state())
        s.close()
    batch.close()


def test_batch_many_small_foams_and_tuning():
    rng = np.random.default_rng(3)
    cfgs = []
    for k in range(48):
        dim = 2 + (k % 2)
        n = int(10 ** rng.uniform(3, 3.7))
        cfgs.append(config(dim, n, phi=float(rng.uniform(0.8, 0.95)),
                           kappa=float(10 ** rng.uniform(-3, -1)), seed=500 + k))
    a = ca.Batch(cfgs)
    b = ca.Batch(cfgs, tuning={"bubbles_per_thread": 2, "search_tile": 128})
    ra = a.run_steps(25, GAS)
    rb = b.run_steps(25, GAS)
    for k in range(len(cfgs)):
        assert ra[k]["steps"] == rb[k]["steps"] == 25
        sa, _ = a.get_state(k)
        sb, _ = b.get_state(k)
        # launch shapes change reduction trees only: integers exact, floats close
        assert np.array_equal(sa["idx"], sb["idx"]) and np.array_equal(sa["nbr"], sb["nbr"])
        assert np.allclose(sa["x"], sb["x"], rtol=1e-12, atol=1e-12)
    a.close()
    b.close()


def test_batch_rejects_mixed_precision_modes():
    cfgs = [config(2, 500, seed=1), config(2, 500, seed=2, precision=ca.PREC_MIXED)]
    with pytest.raises(ca.CubbleError):
        ca.Batch(cfgs)


def test_autotune_restores_state_and_chosen_shapes_match_alone(tmp_path):
    cfgs = mixed_configs()[:4]
    batch = ca.Batch(cfgs)
    batch.run_steps(5, GAS)
    before = [batch.get_state(k) for k in range(len(cfgs))]
    cache = str(tmp_path / "tuning.json")
    rep = batch.autotune(GAS, trial_steps=16, repeats=1, cache_path=cache)
    assert rep["candidates_timed"] >= 10 and not rep["from_cache"]
    assert rep["best_rate"] >= rep["start_rate"] > 0
    # tuning times real steps but hands the foams back unchanged
    for k in range(len(cfgs)):
        assert_same(batch.get_state(k), before[k])
    chosen = rep["chosen"]
    batch.run_steps(30, GAS)
    for k in (0, 3):
        # same shapes step for step: 5 steps with the defaults, then the chosen
        sim = ca.Simulator(cfgs[k])
        sim.run_steps(5, GAS)
        sim.retune(chosen)
        sim.run_steps(30, GAS)
        assert_same(batch.get_state(k), sim.get_state())
        sim.close()
    # same batch make-up: the cache answers without timing
    other = ca.Batch(cfgs)
    rep2 = other.autotune(GAS, cache_path=cache)
    assert rep2["from_cache"] and rep2["chosen"] == chosen and rep2["candidates_timed"] == 0
    other.close()
    batch.close()


def test_retune_live_simulation_equals_created_with_shapes():
    cfg = config(3, 30000, seed=21)
    shapes = {"bubbles_per_thread": 4, "search_tile": 64, "graph_steps": 4}
    a = ca.Simulator(cfg)
    a.run_steps(10, GAS)
    a.retune(shapes)
    b = ca.Simulator(cfg, tuning=shapes)
    b.retune({"bubbles_per_thread": 0, "search_tile": 256, "graph_steps": 16})
    b.run_steps(10, GAS)
    b.retune(shapes)
    a.run_steps(20, GAS)
    b.run_steps(20, GAS)
    assert_same(a.get_state(), b.get_state())
    with pytest.raises(ca.CubbleError):
        a.retune({"search_tile": 8})
    a.close()
    b.close()


def test_zero_step_run_prepares_without_stepping():
    """run_steps(0) captures the device loop and changes no state; a run after
    it equals a run without it, bit for bit."""
    cfgs = mixed_configs()[:3]
    a = ca.Batch(cfgs)
    b = ca.Batch(cfgs)
    before = [a.get_state(k) for k in range(len(cfgs))]
    res = a.run_steps(0, GAS)
    assert all(r["steps"] == 0 for r in res)
    for k in range(len(cfgs)):
        assert_same(a.get_state(k), before[k])
    a.run_steps(15, GAS)
    b.run_steps(15, GAS)
    for k in range(len(cfgs)):
        assert_same(a.get_state(k), b.get_state(k))
    a.close()
    b.close()
