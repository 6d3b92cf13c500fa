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
    b = ca.Batch(cfgs, tuning={"bubbles_per_thread": 2, "search_cells": 16})
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
