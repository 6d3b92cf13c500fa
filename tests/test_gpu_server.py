"""Serving loop: requests arriving while others run are admitted at round
boundaries as memory frees up, their work overlaps on the device, and each
request's final state equals the same foam run alone (bit for bit)."""
import numpy as np
import pytest

import cubble_amd as ca
from parity import config

pytestmark = pytest.mark.gpu

GAS = ca.IntegrationParams(True, False, True)


def request_configs():
    specs = [(3, 3000, 0.85), (2, 4000, 0.9), (3, 1500, 0.92), (2, 2500, 0.8), (3, 5000, 0.88),
             (2, 1200, 0.95)]
    return [config(d, n, phi=phi, seed=700 + k) for k, (d, n, phi) in enumerate(specs)]


def alone(cfg, steps):
    sim = ca.Simulator(cfg)
    sim.run_steps(steps, GAS)
    out = sim.get_state()
    sim.close()
    return out


def test_server_results_equal_alone_and_overlap():
    cfgs = request_configs()
    # the first request runs long; the others arrive while it is on the GPU
    steps = [1500] + [45] * (len(cfgs) - 1)
    srv = ca.Server(hbm_budget_bytes=1 << 30, round_steps=20, keep_states=True)
    ids = [srv.submit(c, max_steps=m, arrival_s=0.002 * k)
           for k, (c, m) in enumerate(zip(cfgs, steps))]
    stats = srv.run()
    assert stats["requests"] == len(cfgs)
    assert stats["p99_latency_s"] >= stats["p50_latency_s"] > 0
    first = srv.result(ids[0])
    overlapped = [srv.result(i) for i in ids[1:] if srv.result(i)["admit_s"] < first["finish_s"]]
    assert overlapped and stats["max_resident"] >= 2   # admitted while another ran
    for rid, cfg, m in zip(ids, cfgs, steps):
        r = srv.result(rid)
        assert r["status"] == 2 and r["steps"] == m
        assert r["admit_s"] >= r["arrival_s"] and r["finish_s"] > r["admit_s"]
        gs, gsc = srv.get_state(rid)
        ws, wsc = alone(cfg, m)
        for key in gs:
            assert np.array_equal(gs[key], ws[key]), key
        assert gsc["dt"] == wsc["dt"] and gsc["bubble_steps"] == wsc["bubble_steps"]
    srv.close()


def test_server_admission_by_footprint():
    # four foams of one size: the budget holds two of them at a time
    cfgs = [config(3, 3000, phi=0.85, seed=40 + k) for k in range(4)]
    need = max(ca.foam_footprint(c) for c in cfgs)
    assert min(ca.foam_footprint(c) for c in cfgs) * 3 > 2.6 * need
    srv = ca.Server(hbm_budget_bytes=int(2.6 * need), round_steps=15)
    ids = [srv.submit(c, max_steps=30, arrival_s=0.0) for c in cfgs]
    stats = srv.run()
    assert stats["requests"] == 4
    assert stats["max_resident"] == 2
    assert stats["peak_bytes"] <= 2.6 * need
    res = [srv.result(i) for i in ids]
    assert all(r["status"] == 2 and r["steps"] == 30 for r in res)
    # the last two waited for memory: admitted only after an earlier one finished
    first_done = min(r["finish_s"] for r in res[:2])
    assert all(r["admit_s"] >= first_done for r in res[2:])
    srv.close()


def test_server_rejects_request_larger_than_budget():
    cfg = config(3, 20000, seed=5)
    srv = ca.Server(hbm_budget_bytes=ca.foam_footprint(cfg) // 2, round_steps=10)
    rid = srv.submit(cfg, max_steps=10)
    srv.run()
    assert srv.result(rid)["status"] == ca.E_NOMEM
    srv.close()


def test_server_refuses_flow_requests():
    srv = ca.Server(hbm_budget_bytes=1 << 28, round_steps=10)
    with pytest.raises(ca.CubbleError) as e:
        srv.submit(config(2, 2000, seed=3, flow=1), max_steps=10)
    assert e.value.code == ca.E_UNSUPPORTED
    srv.close()


def test_server_dissolution_stop():
    # a broad size distribution with strong gas exchange dissolves bubbles quickly
    cfg = config(2, 3000, dist=1, phi=0.9, seed=9, r_std=0.5, kappa=0.5, k_param=1.0)
    srv = ca.Server(hbm_budget_bytes=1 << 28, round_steps=25, keep_states=True)
    rid = srv.submit(cfg, max_steps=400, dissolve_fraction=0.01)
    srv.run()
    r = srv.result(rid)
    assert r["status"] == 2
    stop = int(np.floor(0.99 * 3000))
    assert r["final_count"] <= stop or r["steps"] == 400
    if r["steps"] < 400:
        gs, _ = srv.get_state(rid)
        ws, _ = alone(cfg, r["steps"])
        assert np.array_equal(gs["x"], ws["x"]) and np.array_equal(gs["idx"], ws["idx"])
    srv.close()


def test_server_fails_only_the_request_whose_lists_cannot_grow():
    """A request with a tiny neighbour capacity overflows at its first
    search; the budget holds both foams but not its grown lists, so it fails
    alone (CUBBLE_E_NOMEM) and the other request finishes as if alone."""
    a = config(3, 3000, phi=0.85, seed=61)
    b = config(3, 3000, phi=0.85, seed=62, nbr_cap=2)
    fa, fb = ca.foam_footprint(a), ca.foam_footprint(b)
    print(f"server nomem: footprints {fa} {fb}", flush=True)
    # B's lists grow from 2 to ~25 columns (+~90 KB per 1000 bubbles), more
    # than the 15% of B's footprint left free
    srv = ca.Server(hbm_budget_bytes=int(fa + 1.15 * fb), round_steps=10, keep_states=True)
    ra = srv.submit(a, max_steps=30, arrival_s=0.0)
    rb = srv.submit(b, max_steps=30, arrival_s=0.0)
    stats = srv.run()
    res_a, res_b = srv.result(ra), srv.result(rb)
    assert res_b["status"] == ca.E_NOMEM and res_b["finish_s"] >= res_b["admit_s"] >= 0
    assert res_a["status"] == 2 and res_a["steps"] == 30
    assert stats["requests"] == 1
    gs, gsc = srv.get_state(ra)
    ws, wsc = alone(a, 30)
    for key in gs:
        assert np.array_equal(gs[key], ws[key]), key
    srv.close()
