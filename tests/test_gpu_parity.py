"""GPU parity of the HIP path against the CPU FP64 oracle (through the C ABI).

Bars: neighbour lists, cell order, wrap counters, compaction: bit-exact.
Per-step velocities and dR/dt: <= 1e-12 (scaled metric, tests/parity.py) for
FP64, <= 1e-5 for the mixed-precision path.  Total gas volume after a volume
fix: phi * V_box to 1e-10, and GPU vs oracle to 1e-10."""
import numpy as np
import pytest

import cubble_amd as ca
from parity import (FP64_TOL, MIXED_TOL, VOLUME_TOL, assert_same_structure, attempt_errors,
                    config, oracle_for, oracle_from_sim, scaled_err, total_volume)

pytestmark = pytest.mark.gpu

GAS = ca.IntegrationParams(True, False, True)
CASES = [(2, 3000, 426), (3, 4000, 7), (2, 20000, 5), (3, 60000, 9)]


@pytest.mark.parametrize("dim,n,seed", CASES)
def test_initial_state_and_first_search_bit_exact(ora, dim, n, seed):
    cfg = config(dim, n, seed=seed)
    sim = ca.Simulator(cfg)
    f = oracle_for(ora, cfg)
    gst, gsc = sim.get_state()
    ost, osc = f.get_state()
    assert gsc["count"] == osc["count"] == n
    assert list(gsc["cell_dim"]) == list(osc["cell_dim"])
    assert_same_structure(gst, ost)
    for key in ("x", "saved"):
        assert np.array_equal(gst[key][:dim], ost[key][:dim]), key
        assert np.array_equal(gst[key][3], ost[key][3]), key
    assert gsc["num_pairs"] == osc["num_pairs"]
    sim.close()


@pytest.mark.parametrize("dim,n,seed", CASES)
def test_one_attempt_fp64(ora, dim, n, seed):
    cfg = config(dim, n, seed=seed)
    sim = ca.Simulator(cfg)
    # a few committed steps so velocity history and dt are non-trivial
    for _ in range(3):
        sim.integrate(GAS)
    f = oracle_from_sim(ora, sim)
    r = sim.step(GAS)
    f.attempt(True, True)
    g = sim.get_attempt()
    o = f.attempt_out()
    errs = attempt_errors(dim, g, o, sim.scalars()["f0"])
    print("fp64 errors", dim, n, errs, flush=True)
    for q, e in errs.items():
        assert e <= FP64_TOL, (q, e)
    assert np.array_equal(g["del"], o["del"])
    osc = f.scalars()
    for k in ("s_a", "s_ar", "s_oa", "s_oar", "s_v"):
        assert r is not None
        assert abs(sim.scalars()[k] - osc[k]) <= FP64_TOL * abs(osc[k]), k
    gsc = sim.scalars()
    assert gsc["max_err"] == pytest.approx(osc["max_err"], rel=1e-10)
    assert gsc["max_exp"] == pytest.approx(osc["max_exp"], rel=1e-10)
    assert gsc["dt"] == osc["dt"]
    sim.close()


@pytest.mark.parametrize("dim,n,seed", [(2, 3000, 1), (3, 4000, 2)])
def test_stepwise_integrate_from_shared_state(ora, dim, n, seed):
    """Each step starts both sides from the GPU's state; covers accept/reject,
    the commit rotation, searches and removals over 60 steps."""
    cfg = config(dim, n, seed=seed)
    sim = ca.Simulator(cfg)
    searches = 0
    for step in range(60):
        f = oracle_from_sim(ora, sim)
        sim.integrate(GAS)
        f.integrate(True, True)
        gst, gsc = sim.get_state()
        ost, osc = f.get_state()
        assert gsc["count"] == osc["count"]
        assert gsc["dt"] == pytest.approx(osc["dt"], rel=1e-12)
        assert gsc["steps"] == osc["steps"] and gsc["searches"] == osc["searches"]
        assert_same_structure(gst, ost)
        f0 = gsc["f0"]
        assert max(scaled_err(gst["v"][k], ost["v"][k], f0) for k in range(dim)) <= FP64_TOL
        s_r = max(float(np.max(np.abs(ost["v"][3]))), 1e-300)
        assert scaled_err(gst["v"][3], ost["v"][3], s_r) <= FP64_TOL
        assert np.allclose(gst["x"], ost["x"], rtol=1e-13, atol=1e-13)
        searches = gsc["searches"]
    assert searches >= 0
    sim.close()


@pytest.mark.parametrize("dim,n,seed", [(2, 5000, 3), (3, 8000, 4)])
def test_search_with_wrap_and_reorder_bit_exact(ora, dim, n, seed):
    cfg = config(dim, n, seed=seed)
    sim = ca.Simulator(cfg)
    st, sc = sim.get_state()
    rng = np.random.default_rng(seed)
    L = np.array(sc["interval"][:dim])
    # jitter every bubble and push some across the periodic faces
    st["x"][:dim] += rng.normal(0.0, 0.4, size=(dim, n))
    st["x"][0][: n // 50] -= L[0] * 0.999
    st["x"][1][n // 50: n // 25] += L[1] * 0.999
    sim.set_state(st, sc)
    f = oracle_from_sim(ora, sim)
    sim.search_neighbors()
    f.search()
    gst, gsc = sim.get_state()
    ost, osc = f.get_state()
    assert_same_structure(gst, ost)
    assert np.array_equal(gst["x"], ost["x"])
    assert np.array_equal(gst["saved"], ost["saved"])
    assert np.array_equal(gst["v"], ost["v"]) and np.array_equal(gst["vo"], ost["vo"])
    assert list(gsc["cell_dim"]) == list(osc["cell_dim"])
    assert gsc["num_pairs"] == osc["num_pairs"]
    # the half list equals the O(N^2) brute-force pair set
    if n <= 5000:
        assert np.array_equal(ca.half_pairs(gst["nbr_off"], gst["nbr"]), f.brute_pairs())
    sim.close()


@pytest.mark.parametrize("dim,n,frac,seed", [(2, 4000, 0.01, 5), (3, 6000, 0.03, 6),
                                             (3, 3000, 0.4, 8)])
def test_removal_compaction_and_volume_fix(ora, dim, n, frac, seed):
    cfg = config(dim, n, seed=seed)
    sim = ca.Simulator(cfg)
    st, sc = sim.get_state()
    rng = np.random.default_rng(seed)
    kill = rng.random(n) < frac
    st["x"][3][kill] = 0.05
    sim.set_state(st, sc)
    f = oracle_from_sim(ora, sim)
    sim.step(GAS)
    f.attempt(True, True)
    g = sim.get_attempt()
    o = f.attempt_out()
    assert np.array_equal(g["del"], o["del"]) and g["del"].sum() >= kill.sum()
    sim.remove_bubbles()
    f.remove()
    gst, gsc = sim.get_state()
    ost, osc = f.get_state()
    assert gsc["count"] == osc["count"] == n - int(g["del"].sum())
    assert_same_structure(gst, ost)
    assert np.allclose(gst["x"][3], ost["x"][3], rtol=1e-14, atol=0)
    for key in ("v", "vo", "saved"):
        assert np.array_equal(gst[key], ost[key]), key
    sim.close()


@pytest.mark.parametrize("dim", [2, 3])
def test_volume_conserved_after_removal(ora, dim):
    n = 5000
    cfg = config(dim, n, seed=12)
    sim = ca.Simulator(cfg)
    st, sc = sim.get_state()
    st["x"][3][::53] = 0.03
    sim.set_state(st, sc)
    f = oracle_from_sim(ora, sim)
    res = sim.integrate(GAS)
    f.integrate(True, True)
    assert res["removals"] == 1
    gst, gsc = sim.get_state()
    ost, _ = f.get_state()
    vbox = float(np.prod(gsc["interval"][:dim]))
    vg = total_volume(dim, gst["x"][3])
    vo = total_volume(dim, ost["x"][3])
    assert abs(vg - cfg.phi * vbox) / (cfg.phi * vbox) <= VOLUME_TOL
    assert abs(vg - vo) / vo <= VOLUME_TOL
    sim.close()


@pytest.mark.parametrize("dim,n", [(2, 6000), (3, 5000)])
def test_device_loop_equals_host_loop(dim, n):
    """cubble_run_steps (graph-captured, device-driven) gives bit-identical
    results to repeated cubble_integrate; the graph is replayed many times."""
    cfg = config(dim, n, seed=21)
    a = ca.Simulator(cfg)
    b = ca.Simulator(cfg)
    ra = a.run_steps(40, GAS)
    for _ in range(40):
        rb = b.integrate(GAS)
    assert ra["steps"] == rb["steps"] == 40
    assert ra["attempts"] == rb["attempts"]
    sa, sca = a.get_state()
    sb, scb = b.get_state()
    for key in sa:
        assert np.array_equal(sa[key], sb[key]), key
    assert sca["dt"] == scb["dt"] and sca["bubble_steps"] == scb["bubble_steps"]
    # a second replay batch continues identically
    ra = a.run_steps(25, GAS)
    for _ in range(25):
        b.integrate(GAS)
    sa, _ = a.get_state()
    sb, _ = b.get_state()
    assert np.array_equal(sa["x"], sb["x"]) and np.array_equal(sa["nbr"], sb["nbr"])
    a.close()
    b.close()


# walls on some axes: the minimum image applies only to the periodic ones
WALL_CASES = [(2, 4000, 61, (1, 0, 0)), (3, 6000, 62, (1, 0, 1)), (3, 5000, 63, (1, 1, 1))]


@pytest.mark.parametrize("dim,n,seed,wall", WALL_CASES)
def test_walls_first_search_and_attempt_fp64(ora, dim, n, seed, wall):
    cfg = config(dim, n, seed=seed, wall=wall)
    sim = ca.Simulator(cfg)
    f = oracle_for(ora, cfg)
    gst, gsc = sim.get_state()
    ost, osc = f.get_state()
    assert list(gsc["cell_dim"]) == list(osc["cell_dim"])
    assert_same_structure(gst, ost)
    assert gsc["num_pairs"] == osc["num_pairs"]
    for _ in range(3):
        sim.integrate(GAS)
    f = oracle_from_sim(ora, sim)
    sim.step(GAS)
    f.attempt(True, True)
    errs = attempt_errors(dim, sim.get_attempt(), f.attempt_out(), sim.scalars()["f0"])
    print("wall fp64 errors", dim, n, wall, errs, flush=True)
    for q, e in errs.items():
        assert e <= FP64_TOL, (q, e)
    sim.close()


# the 2D cases of 3e5 and 6e5 bubbles: boxes of ~1,030 and ~1,460 radii with
# walls (wall axes code positions in units of L / 2^30), three steps after
# seeding: they hold nearly coincident centres (where code quantisation turns
# the pair direction) and many near-tangent pairs (where it moves the 2D
# overlap term), whose values the mixed kernel takes from FP64; the last, a
# 3D box of 4e5 bubbles with walls on every axis
@pytest.mark.parametrize("dim,n,seed,wall", [(2, 5000, 31, (0, 0, 0)), (3, 8000, 32, (0, 0, 0)),
                                             (3, 60000, 33, (0, 0, 0)), (2, 4000, 34, (0, 1, 0)),
                                             (3, 6000, 35, (1, 0, 1)),
                                             (2, 300000, 36, (1, 1, 0)),
                                             (2, 600000, 37, (1, 0, 0)),
                                             (3, 400000, 38, (1, 1, 1))])
def test_mixed_precision_attempt_error(ora, dim, n, seed, wall):
    cfg = config(dim, n, seed=seed, precision=ca.PREC_MIXED, wall=wall)
    sim = ca.Simulator(cfg)
    for _ in range(3):
        sim.integrate(GAS)
    f = oracle_from_sim(ora, sim)
    sim.step(GAS)
    f.attempt(True, True)
    errs = attempt_errors(dim, sim.get_attempt(), f.attempt_out(), sim.scalars()["f0"])
    print("mixed errors", dim, n, errs, flush=True)
    for q, e in errs.items():
        assert e <= MIXED_TOL, (q, e)
    sim.close()


@pytest.mark.parametrize("dim,wall", [(2, (0, 0, 0)), (2, (1, 1, 0)), (3, (0, 0, 0)),
                                      (3, (1, 0, 1))])
def test_mixed_close_and_tangent_pairs(ora, dim, wall):
    """Hand-placed pairs that FP32 values from the fixed-point codes do not
    settle, 20 of each kind: centres 1e-6..1e-3 apart (quantisation turns the
    pair direction by up to ~s / |d|, here up to ~1e-1), centres within
    1e-9..1e-6 of tangency (in 2D the overlap term moves by up to ~sqrt of
    the error of R^2 - |d|^2), and centres within 1e-9..1e-6 of the inside
    point |d| = max(r1, r2) (the area jumps there).  Their bubbles are at
    rest, so the predictions keep the placed distances.  The attempt meets
    the 1e-5 bar, and the planted bubbles match the oracle to 1e-6."""
    print(f"mixed planted pairs dim={dim} wall={wall}", flush=True)
    n = 3000 if dim == 2 else 4000
    sim = ca.Simulator(config(dim, n, seed=90 + dim, precision=ca.PREC_MIXED, wall=wall))
    st, sc = sim.get_state()
    rng = np.random.default_rng(90 + dim)
    lo, hi = np.array(sc["lbb"][:dim]), np.array(sc["tfr"][:dim])
    x = st["x"]
    inner = np.flatnonzero(np.all((x[:dim].T > lo + 4.0) & (x[:dim].T < hi - 4.0), axis=1))
    pick = rng.choice(inner, size=120, replace=False)
    planted = []
    for k in range(60):
        a, b = pick[2 * k], pick[2 * k + 1]
        ra, rb = x[3][a], x[3][b]
        delta = 10.0 ** rng.uniform(-9, -6) * rng.choice([-1.0, 1.0])
        if k % 3 == 0:
            dist = 10.0 ** rng.uniform(-6, -3)
        elif k % 3 == 1:
            dist = ra + rb + delta
        else:
            dist = max(ra, rb) + delta
        u = rng.normal(size=dim)
        x[:dim, b] = x[:dim, a] + dist * u / np.linalg.norm(u)
        for key in ("v", "vo"):
            st[key][:, a] = 0.0
            st[key][:, b] = 0.0
        planted += [st["idx"][a], st["idx"][b]]
    sim.set_state(st, sc)
    sim.search_neighbors()
    gst, _ = sim.get_state()
    f = oracle_from_sim(ora, sim)
    sim.step(GAS)
    f.attempt(True, True)
    g, o = sim.get_attempt(), f.attempt_out()
    f0 = sc["f0"]
    errs = attempt_errors(dim, g, o, f0)
    mask = np.isin(gst["idx"], planted)
    assert mask.sum() == 120
    tight = {
        "velocity": max(scaled_err(g["vp"][k][mask], o["vp"][k][mask], f0) for k in range(dim)),
        "drdt": scaled_err(g["vp"][3][mask], o["vp"][3][mask], float(np.max(np.abs(o["vp"][3])))),
        "overlap": scaled_err(g["ov"][mask], o["ov"][mask], float(np.max(np.abs(o["ov"])))),
    }
    print("mixed planted errors", errs, tight, flush=True)
    for q, e in errs.items():
        assert e <= MIXED_TOL, (q, e)
    for q, e in tight.items():
        assert e <= 1e-6, (q, e)
    sim.close()


def test_mixed_second_pass_with_several_bubbles_per_thread(ora):
    """The mixed second pass gathers each bubble slot's flagged bubbles in
    turn when a thread holds several bubbles: with 1, 2 and 4 bubbles per
    thread the per-bubble terms (velocities, overlaps) are the same bits,
    dR/dt differs only through the block partials' summation order, and each
    run meets the oracle bar.  A large 2D box with walls, three steps after
    seeding, flags many bubbles."""
    cfg = config(2, 300000, seed=36, precision=ca.PREC_MIXED, wall=(1, 1, 0))
    base = ca.Simulator(cfg)
    for _ in range(3):
        base.integrate(GAS)
    st, sc = base.get_state()
    base.close()
    outs = []
    for ipt in (1, 2, 4):
        print(f"mixed second pass bubbles_per_thread={ipt}", flush=True)
        sim = ca.Simulator(cfg, tuning={"bubbles_per_thread": ipt, "search_tile": 64,
                                        "graph_steps": 8})
        sim.set_state(st, sc)
        f = oracle_from_sim(ora, sim)
        sim.step(GAS)
        f.attempt(True, True)
        g = sim.get_attempt()
        errs = attempt_errors(2, g, f.attempt_out(), sc["f0"])
        for q, e in errs.items():
            assert e <= MIXED_TOL, (ipt, q, e)
        outs.append(g)
        sim.close()
    for g in outs[1:]:
        for k in range(2):
            assert np.array_equal(g["vp"][k], outs[0]["vp"][k])
        assert np.array_equal(g["ov"], outs[0]["ov"])
        assert scaled_err(g["vp"][3], outs[0]["vp"][3],
                          float(np.max(np.abs(outs[0]["vp"][3])))) <= FP64_TOL


def test_large_foam_multiblock_partials(ora):
    """More than 256 * 4096 bubbles: several bubbles per thread, the full
    reduction tree and the XCD block remap."""
    n = 1_200_000
    cfg = config(3, n, seed=77)
    sim = ca.Simulator(cfg)
    f = oracle_for(ora, cfg)
    gst, gsc = sim.get_state()
    ost, osc = f.get_state()
    assert_same_structure(gst, ost)
    f2 = oracle_from_sim(ora, sim)
    sim.step(GAS)
    f2.attempt(True, True)
    errs = attempt_errors(3, sim.get_attempt(), f2.attempt_out(), gsc["f0"])
    print("large fp64 errors", errs, flush=True)
    for q, e in errs.items():
        assert e <= FP64_TOL, (q, e)
    sim.close()


def test_neighbour_capacity_growth(ora):
    """A tiny initial list capacity forces the overflow path: lists are
    rebuilt with a larger capacity and stay bit-exact."""
    cfg = config(3, 4000, seed=41, nbr_cap=2)
    sim = ca.Simulator(cfg)
    f = oracle_for(ora, cfg)
    gst, gsc = sim.get_state()
    ost, _ = f.get_state()
    assert gsc["nbr_cap"] > 2
    assert_same_structure(gst, ost)
    sim.close()


def test_lists_grow_ahead_of_overflow():
    """A list capacity just above the longest list: the host grows the lists,
    keeping their contents, as soon as a search comes within 1/8 of the
    capacity (here right after the initial search), before a later search can
    overflow them; the run equals one with ample capacity, bit for bit."""
    a = ca.Simulator(config(3, 6000, seed=43))
    st, _ = a.get_state()
    longest = int(np.max(np.diff(st["nbr_off"])))
    b = ca.Simulator(config(3, 6000, seed=43, nbr_cap=longest + 1))
    assert b.scalars()["nbr_cap"] > longest + 1
    sb0, _ = b.get_state()
    assert np.array_equal(st["nbr"], sb0["nbr"]) and np.array_equal(st["nbr_off"], sb0["nbr_off"])
    a.run_steps(60, GAS)
    b.run_steps(60, GAS)
    sa, sca = a.get_state()
    sb, scb = b.get_state()
    for key in sa:
        assert np.array_equal(sa[key], sb[key]), key
    assert sca["steps"] == scb["steps"] and sca["dt"] == scb["dt"]
    a.close()
    b.close()


def test_json_create_and_cli_run(tmp_path):
    import os
    path = os.path.join(os.path.dirname(__file__), "golden", "input_small_2d.json")
    sim = ca.Simulator(json_path=path)
    assert sim.scalars()["count"] == 2000
    sim.close()
    assert ca.run_cli(["cubble-cli", path]) == 0


@pytest.mark.parametrize("dim,n,seed", [(2, 2025, 61), (3, 4096, 62)])
def test_grid_placement_stepwise_from_shared_state(ora, dim, n, seed):
    """A foam started on the reference's lattice (init.placement "grid"):
    the first search and 30 steps, each started on both sides from the GPU's
    state."""
    print(f"grid placement dim={dim} n={n} seed={seed}", flush=True)
    sim = ca.Simulator(config(dim, n, seed=seed, placement=1))
    f = oracle_from_sim(ora, sim)
    f.search()
    assert_same_structure(sim.get_state()[0], f.get_state()[0])
    for step in range(30):
        f = oracle_from_sim(ora, sim)
        sim.integrate(GAS)
        f.integrate(True, True)
        gst, gsc = sim.get_state()
        ost, osc = f.get_state()
        assert gsc["count"] == osc["count"] and gsc["searches"] == osc["searches"]
        assert_same_structure(gst, ost)
        f0 = gsc["f0"]
        assert max(scaled_err(gst["v"][k], ost["v"][k], f0) for k in range(dim)) <= FP64_TOL
        assert np.allclose(gst["x"], ost["x"], rtol=1e-13, atol=1e-13)
    sim.close()


@pytest.mark.parametrize("dim,n,seed,precision", [(2, 3000, 71, ca.PREC_FP64),
                                                  (3, 4000, 72, ca.PREC_FP64),
                                                  (3, 6000, 73, ca.PREC_MIXED)])
def test_velocity_bootstrap_matches_oracle(ora, dim, n, seed, precision):
    """The AB2 bootstrap of init on a new foam, then (FP64; a mixed run may
    accept a different attempt than the oracle) 10 steps from the shared
    state; refused once a step has run."""
    print(f"bootstrap dim={dim} n={n} seed={seed} precision={precision}", flush=True)
    sim = ca.Simulator(config(dim, n, seed=seed, precision=precision))
    f = oracle_from_sim(ora, sim)
    sim.bootstrap_velocities()
    f.bootstrap()
    gst, gsc = sim.get_state()
    ost, _ = f.get_state()
    tol = FP64_TOL if precision == ca.PREC_FP64 else MIXED_TOL
    f0 = gsc["f0"]
    assert max(scaled_err(gst["vo"][k], ost["vo"][k], f0) for k in range(dim)) <= tol
    assert np.max(np.abs(ost["vo"][:dim])) > 1e-3 * f0
    assert not np.any(gst["v"]) and np.array_equal(gst["x"], ost["x"])
    for _ in range(10 if precision == ca.PREC_FP64 else 0):
        f = oracle_from_sim(ora, sim)
        sim.integrate(GAS)
        f.integrate(True, True)
        gst, gsc = sim.get_state()
        ost, _ = f.get_state()
        assert_same_structure(gst, ost)
        assert max(scaled_err(gst["v"][k], ost["v"][k], f0) for k in range(dim)) <= tol
    sim.integrate(GAS)
    with pytest.raises(ca.CubbleError) as e:
        sim.bootstrap_velocities()
    assert e.value.code == ca.E_STATE
    sim.close()


@pytest.mark.parametrize("dim,n,seed,rel", [(2, 2025, 81, (1.0, 1.0, 1.0)),
                                            (3, 4096, 82, (1.0, 1.0, 1.0)),
                                            (2, 2048, 83, (2.0, 1.0, 1.0))])
def test_rescale_box_matches_oracle(ora, dim, n, seed, rel):
    """A lattice in its own box (placement 2), scaled to phi: positions, box,
    flow box and volume multiplier equal the oracle's given the same total
    volume; the search that follows equals it bit for bit."""
    print(f"rescale dim={dim} n={n} seed={seed} rel={rel}", flush=True)
    cfg = config(dim, n, seed=seed, placement=2, rel_dim=rel, phi=0.85)
    sim = ca.Simulator(cfg)
    before = sim.scalars()
    f = oracle_from_sim(ora, sim)
    vol = sim.observe()["total_volume"]
    assert vol / np.prod(before["interval"][:dim]) < 0.8   # the lattice box is loose
    sim.rescale_box(0.85, rel)
    f.rescale(0.85, rel, vol)
    gst, gsc = sim.get_state()
    ost, _ = f.get_state()
    oc = f.consts()
    assert np.array_equal(gst["x"], ost["x"])
    for k in ("interval", "tfr", "flow_lbb", "flow_tfr"):
        assert list(gsc[k][:dim]) == list(oc[k][:dim]), k
    assert gsc["vol_mul"] == oc["vol_mul"]
    assert_same_structure(gst, ost)
    assert vol / np.prod(gsc["interval"][:dim]) == pytest.approx(0.85, rel=1e-12)
    ratio = np.array(gsc["interval"][:dim]) / np.array(rel[:dim])
    assert np.allclose(ratio, ratio[0], rtol=1e-14)
    sim.close()


# Boxes of a few cells per axis.  2D: periodic rows of 5 to 9 cells, where the
# search's row windows narrow and wrap at either row end), a wall in x or y,
# and boxes elongated along x (long rows): the lists equal the oracle's, and
# the brute-force pair set, after the first search and after a search with
# jittered positions pushed across the faces.
SMALL_2D = [(80, 91, (0, 0, 0), (1.0, 1.0, 1.0)), (110, 92, (0, 0, 0), (1.0, 1.0, 1.0)),
            (150, 93, (0, 0, 0), (1.0, 1.0, 1.0)), (200, 94, (0, 0, 0), (1.0, 1.0, 1.0)),
            (300, 95, (1, 0, 0), (1.0, 1.0, 1.0)), (400, 96, (0, 1, 0), (2.0, 1.0, 1.0)),
            (600, 97, (0, 0, 0), (3.0, 1.0, 1.0)), (2000, 98, (1, 0, 0), (4.0, 1.0, 1.0))]


# 3D: 5 to 7 cells per axis (rows of two runs at either end), a wall, a
# box elongated along x.
SMALL_3D = [(1400, 99, (0, 0, 0), (1.0, 1.0, 1.0)), (2400, 100, (0, 0, 0), (1.0, 1.0, 1.0)),
            (3000, 101, (0, 0, 1), (2.0, 1.0, 1.0))]


@pytest.mark.parametrize("dim,n,seed,wall,rel",
                         [(2,) + c for c in SMALL_2D] + [(3,) + c for c in SMALL_3D])
def test_small_and_long_boxes_search_bit_exact(ora, dim, n, seed, wall, rel):
    cfg = config(dim, n, seed=seed, wall=wall, rel_dim=rel)
    sim = ca.Simulator(cfg)
    f = oracle_for(ora, cfg)
    gst, gsc = sim.get_state()
    ost, osc = f.get_state()
    print("box", dim, n, wall, rel, "cells", list(gsc["cell_dim"][:dim]), flush=True)
    assert list(gsc["cell_dim"]) == list(osc["cell_dim"])
    assert_same_structure(gst, ost)
    assert np.array_equal(ca.half_pairs(gst["nbr_off"], gst["nbr"]), f.brute_pairs())
    # jitter, push some bubbles across the periodic faces, search again
    rng = np.random.default_rng(seed)
    L = np.array(gsc["interval"][:dim])
    gst["x"][:dim] += rng.normal(0.0, 0.3, size=(dim, n))
    for a in range(dim):
        if not wall[a]:
            gst["x"][a][a * (n // 10): (a + 1) * (n // 10)] += L[a] * 0.999
        else:
            lo, hi = gsc["lbb"][a], gsc["tfr"][a]
            gst["x"][a] = np.clip(gst["x"][a], lo + 0.5, hi - 0.5)
    sim.set_state(gst, gsc)
    f = oracle_from_sim(ora, sim)
    sim.search_neighbors()
    f.search()
    gst, gsc = sim.get_state()
    ost, osc = f.get_state()
    assert_same_structure(gst, ost)
    assert np.array_equal(gst["x"], ost["x"])
    assert gsc["num_pairs"] == osc["num_pairs"]
    assert np.array_equal(ca.half_pairs(gst["nbr_off"], gst["nbr"]), f.brute_pairs())
    sim.close()
