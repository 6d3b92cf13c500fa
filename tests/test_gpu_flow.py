"""GPU parity of the imposed flow (flow.impose; reference addFlowVelocity,
src/kernels.cu:497-544, with the neighbour-velocity sums of the pair kernel,
:191-208, :281-307) against the CPU FP64 oracle, through the C ABI.

Bars as for the rest of the path: FP64 velocities <= 1e-12 (scaled metric),
mixed precision <= 1e-5; lists, counts and the flow's divisor (the list length
at the last search, moved but not reduced by a removal) bit-exact."""
import numpy as np
import pytest

import cubble_amd as ca
from parity import (FP64_TOL, MIXED_TOL, assert_same_structure, attempt_errors, config,
                    oracle_from_sim, scaled_err)

pytestmark = pytest.mark.gpu

FLOW = ca.IntegrationParams(True, True, True)


def flow_config(dim, n, seed, **over):
    # a box around the middle third of the foam, flow at half of f0
    return config(dim, n, seed=seed, flow=1, flow_vel=(0.5, -0.25, 0.125),
                  flow_lbb=(0.3, 0.35, 0.3), flow_tfr=(0.7, 0.6, 0.65), **over)


def _inside_count(sim):
    st, sc = sim.get_state()
    dim = sc["dim"]
    x, r = st["x"], st["x"][3]
    inside = np.ones(sc["count"], bool)
    for a in range(dim):
        lo, hi = sc["flow_lbb"][a], sc["flow_tfr"][a]
        inside &= ((x[a] < hi) & (x[a] > lo)) | ((lo - x[a]) ** 2 <= r * r) | \
                  ((hi - x[a]) ** 2 <= r * r)
    return int(inside.sum())


@pytest.mark.parametrize("dim,n,seed,precision,wall", [
    (2, 4000, 41, ca.PREC_FP64, (0, 0, 0)),
    (3, 6000, 42, ca.PREC_FP64, (0, 0, 0)),
    (3, 5000, 43, ca.PREC_FP64, (1, 0, 0)),
    (2, 5000, 44, ca.PREC_MIXED, (0, 0, 0)),
    (3, 8000, 45, ca.PREC_MIXED, (0, 1, 0)),
    # a large 2D box with walls, a few steps after seeding: the flow's own
    # branch band next to nearly coincident and near-tangent pairs
    (2, 300000, 46, ca.PREC_MIXED, (1, 1, 0)),
])
def test_flow_attempt(ora, dim, n, seed, precision, wall):
    print(f"flow attempt dim={dim} n={n} seed={seed} precision={precision} wall={wall}",
          flush=True)
    sim = ca.Simulator(flow_config(dim, n, seed, precision=precision, wall=wall))
    sc = sim.scalars()
    f0 = sc["f0"]
    assert sc["flow_vel"][:dim] == pytest.approx([0.5 * f0, -0.25 * f0, 0.125 * f0][:dim])
    for _ in range(4):
        sim.integrate(FLOW)
    assert 0 < _inside_count(sim) < sim.scalars()["count"]
    f = oracle_from_sim(ora, sim)
    sim.step(FLOW)
    f.attempt(True, True, flow=True)
    g, o = sim.get_attempt(), f.attempt_out()
    errs = attempt_errors(dim, g, o, f0)
    print("flow errors", errs, flush=True)
    tol = FP64_TOL if precision == ca.PREC_FP64 else MIXED_TOL
    for q, e in errs.items():
        assert e <= tol, (q, e)
    # the flow changes the velocities: the same attempt without it differs
    h = oracle_from_sim(ora, sim)
    h.attempt(True, True, flow=False)
    assert max(scaled_err(h.attempt_out()["vp"][k], o["vp"][k], f0) for k in range(dim)) > 1e-3
    sim.close()


@pytest.mark.parametrize("dim,n,seed", [(2, 3000, 46), (3, 4000, 47)])
def test_flow_stepwise_integrate_from_shared_state(ora, dim, n, seed):
    """60 steps with the flow on, each started on both sides from the GPU's
    state (the divisor included): accept/reject, removals and searches."""
    print(f"flow stepwise dim={dim} n={n} seed={seed}", flush=True)
    sim = ca.Simulator(flow_config(dim, n, seed, kappa=0.05))
    stale = 0
    removals = 0
    for step in range(60):
        f = oracle_from_sim(ora, sim)
        sim.integrate(FLOW)
        f.integrate(True, True, flow=True)
        gst, gsc = sim.get_state()
        ost, osc = f.get_state()
        assert gsc["count"] == osc["count"]
        assert gsc["dt"] == pytest.approx(osc["dt"], rel=1e-12)
        assert gsc["removals"] == osc["removals"] and gsc["searches"] == osc["searches"]
        assert_same_structure(gst, ost)
        assert np.array_equal(gst["deg"], ost["deg"])
        assert np.all(gst["deg"] >= np.diff(gst["nbr_off"]))
        stale += int(np.any(gst["deg"] > np.diff(gst["nbr_off"])))
        f0 = gsc["f0"]
        assert max(scaled_err(gst["v"][k], ost["v"][k], f0) for k in range(dim)) <= FP64_TOL
        assert np.allclose(gst["x"], ost["x"], rtol=1e-13, atol=1e-13)
        removals = gsc["removals"]
    print("flow stepwise removals", removals, "steps with reduced lists", stale, flush=True)
    sim.close()


def test_flow_state_round_trip_keeps_divisor():
    """get_state / set_state carry the divisor; a state without it takes the
    current list lengths."""
    sim = ca.Simulator(flow_config(2, 3000, 48, kappa=0.05))
    for _ in range(30):
        sim.integrate(FLOW)
    st, sc = sim.get_state()
    st["deg"] = st["deg"] + 3
    sim.set_state(st, sc)
    assert np.array_equal(sim.get_state()[0]["deg"], st["deg"])
    del st["deg"]
    sim.set_state(st, sc)
    got, _ = sim.get_state()
    assert np.array_equal(got["deg"], np.diff(got["nbr_off"]))
    sim.close()
