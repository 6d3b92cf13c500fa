"""GPU tests of the integration loop's outputs (through the C ABI): the
observables of a printed row (energy, volume, means) against the oracle, the
device stop rules and time-step window against a one-step-at-a-time replay,
and the command line's table, results.dat and snapshot files against a
replay of the reference's loop (src/simulator.cu:958-1157) driven from
Python one accepted step at a time.

Bars: per-bubble energies bit-exact (same lists, same order, FP64);
totals and means 1e-12 relative (block-order sums); stop steps, counts and
the time-step window exact; printed text identical."""
import ctypes
import glob
import json
import math
import os

import numpy as np
import pytest

import cubble_amd as ca
from parity import config, oracle_from_sim

pytestmark = pytest.mark.gpu

GAS = ca.IntegrationParams(True, False, True)


@pytest.mark.parametrize("dim,n,seed,precision,wall", [
    (2, 3000, 426, ca.PREC_FP64, (0, 0, 0)),
    (3, 4000, 7, ca.PREC_FP64, (0, 0, 0)),
    (3, 5000, 8, ca.PREC_FP64, (1, 0, 1)),
    (2, 20000, 5, ca.PREC_MIXED, (0, 0, 0)),
    (3, 60000, 9, ca.PREC_MIXED, (0, 1, 0)),
])
def test_observe_matches_oracle(ora, dim, n, seed, precision, wall):
    print(f"observe dim={dim} n={n} seed={seed} precision={precision} wall={wall}", flush=True)
    sim = ca.Simulator(config(dim, n, seed=seed, precision=precision, wall=wall))
    sim.run_steps(40, GAS)
    got = sim.observe(energy=True)
    f = oracle_from_sim(ora, sim)
    sums, e = f.observe()
    assert np.array_equal(got["energy"], e)
    cnt = sim.scalars()["count"]
    assert got["total_energy"] == pytest.approx(sums[0], rel=1e-12)
    assert got["total_volume"] == pytest.approx(sums[1], rel=1e-12)
    assert got["mean_r"] == pytest.approx(sums[2] / cnt, rel=1e-12)
    assert got["mean_path"] == pytest.approx(sums[3] / cnt, rel=1e-12)
    st, sc = sim.get_state()
    for q in range(4):
        scale = np.abs(st["v"][q]).sum() / cnt
        assert abs(got["mean_v"][q] - sums[4 + q] / cnt) <= 1e-12 * max(scale, 1e-300)
    iv = sc["interval"]
    assert got["box_volume"] == iv[0] * iv[1] * (iv[2] if dim == 3 else 1.0)
    sim.close()


def _split(t):
    i = int(t)
    return i, t - i


def _reached(sc, t):
    i, f = t
    return sc["time_int"] > i or (sc["time_int"] == i and sc["time_frac"] >= f)


@pytest.mark.parametrize("dim,n,seed,precision", [(2, 4000, 3, ca.PREC_FP64),
                                                  (3, 6000, 4, ca.PREC_MIXED)])
def test_time_stop_and_dt_window_match_stepwise(dim, n, seed, precision):
    print(f"stop rules dim={dim} n={n} seed={seed} precision={precision}", flush=True)
    cfg = config(dim, n, seed=seed, precision=precision)
    a = ca.Simulator(cfg)
    b = ca.Simulator(cfg)
    # stepwise: the time after each of 60 steps, and each step's dt
    times, dts = [], []
    for _ in range(60):
        r = a.integrate(GAS)
        sc = a.scalars()
        times.append((sc["time_int"], sc["time_frac"]))
        dts.append(r["dt"])
    # a stop between the times after steps 36 and 37 halts after step 37;
    # one equal to the time after step 45 halts after step 45
    b.reset_print_window()
    for k, stop in ((37, None), (45, times[44])):
        if stop is None:
            lo, hi = times[k - 2], times[k - 1]
            stop = _split(0.5 * ((lo[0] + lo[1]) + (hi[0] + hi[1])))
            assert not _reached({"time_int": lo[0], "time_frac": lo[1]}, stop)
        b.set_stop_rules(time=stop)
        b.run_steps(1000, GAS)
        sc = b.scalars()
        assert sc["steps"] == k
        assert (sc["time_int"], sc["time_frac"]) == times[k - 1]
    # the window holds steps 1..45 (never reset since the start)
    w = dts[:45]
    assert sc["ts_min"] == min(w) and sc["ts_max"] == max(w)
    acc = 0.0
    for x in w:
        acc += x
    assert sc["ts_sum"] == acc
    b.reset_print_window()
    sc = b.scalars()
    assert sc["ts_min"] == 9999999.9 and sc["ts_max"] == -1.0 and sc["ts_sum"] == 0.0
    st, _ = b.get_state()
    assert not np.any(st["err"])
    # without rules a run does all its steps
    b.set_stop_rules()
    b.run_steps(10, GAS)
    assert b.scalars()["steps"] == 55
    a.close()
    b.close()


def test_radius_and_count_stops():
    cfg = config(2, 3000, seed=21)
    sim = ca.Simulator(cfg)
    sim.run_steps(5, GAS)
    r0 = sim.scalars()["max_r"]
    # every step's largest radius exceeds half of it: halts after one step
    sim.set_stop_rules(max_radius=0.5 * r0)
    sim.run_steps(100, GAS)
    assert sim.scalars()["steps"] == 6
    # a count rule at the current count halts at the next removal, if any;
    # the run never goes past its step limit
    sim.set_stop_rules(min_count=sim.scalars()["count"])
    before = sim.scalars()
    sim.run_steps(300, GAS)
    after = sim.scalars()
    if after["removals"] > before["removals"]:
        assert after["steps"] < before["steps"] + 300
        assert after["count"] < before["count"]
    else:
        assert after["steps"] == before["steps"] + 300
    sim.close()


# ---------------------------------------------------------------- the CLI
def _g(x):
    """C++ ostream default formatting of a double (== printf %g)."""
    return "%g" % x


def _stabilize(sim, steps):
    e1 = sim.observe()["total_energy"]
    sim.run_steps(steps, ca.IntegrationParams(False, False, False))
    return e1, sim.observe()["total_energy"]


def _replay(json_path, rows):
    """The reference's init stabilisation and loop from Python, the loop one
    accepted step at a time: the stabilisation lines, then the first `rows`
    printed rows as (table text, results.dat line).  The stabilisation must
    stop after its first counted round."""
    with open(json_path) as fh:
        j = json.load(fh)
    r_mean = j["bubbles"]["radius"]["mean"]
    num_end = int(j["bubbles"]["numEnd"])
    tsf = j["constants"]["K"]["value"] / (r_mean * r_mean)
    steps = int(j["stabilization"]["steps"])
    sim = ca.Simulator(json_path=json_path)
    sim.bootstrap_velocities()
    for _ in range(5):
        _stabilize(sim, steps)
    searches0 = sim.scalars()["searches"]
    sim.reset_clock()
    e1, e2 = _stabilize(sim, steps)
    sc = sim.scalars()
    time = (sc["time_int"] + sc["time_frac"]) * tsf
    de = abs(e2 - e1)
    if de > 0:
        de *= 2.0 / ((e2 + e1) * time)
    assert de < j["stabilization"]["maxDeltaEnergy"]
    stab = ["Final energies:", "before: %9.5e" % e1, "after: %9.5e" % e2, "delta: %9.5e" % de,
            "time: %9.5g" % time]
    sim.reset_clock(clear_wraps=True)
    sim.reset_print_window()
    e1 = sim.observe()["total_energy"]
    printed, n_steps, ts_min, ts_max, ts_sum = 1, 0, 9999999.9, -1.0, 0.0
    table, results = [], []
    while printed <= rows:
        r = sim.integrate(GAS)
        sc = sim.scalars()
        ts = r["dt"]
        ts_min, ts_max = min(ts_min, ts), max(ts_max, ts)
        ts_sum += ts
        assert sc["count"] > num_end
        if _reached(sc, _split(printed / tsf)):
            o = sim.observe()
            e2 = o["total_energy"]
            de = abs(e2 - e1)
            if de > 0:
                de *= 2.0 / (e2 + e1)
            rel = o["mean_r"] / r_mean
            v = o["mean_v"]
            results.append(" ".join([str(printed), _g(rel), str(sc["count"]), _g(o["mean_path"]),
                                     _g(e2), _g(de), _g(v[0]), _g(v[1]), _g(v[2]),
                                     _g(math.sqrt(v[0] ** 2 + v[1] ** 2 + v[2] ** 2)),
                                     _g(v[3])]))
            table.append("%-5d %-#8.6g %-#6.4g %-9.5e %9d %10d %6ld %-9ld %-9.5e %-9.5e %-9.5e " % (
                printed, o["total_volume"] / o["box_volume"], rel, de, sc["count"],
                sc["num_pairs"], n_steps, sc["searches"] - searches0, ts_min, ts_max,
                ts_sum / n_steps if n_steps else math.inf))
            printed += 1
            e1 = e2
            n_steps, ts_min, ts_max, ts_sum = 0, 9999999.9, -1.0, 0.0
            searches0 = sc["searches"]
            sim.reset_print_window()
        n_steps += 1
    sim.close()
    return stab, table, results


def test_cli_table_results_and_snapshots(tmp_path, monkeypatch, capfd):
    src = os.path.join(os.path.dirname(__file__), "golden", "input_small_2d.json")
    with open(src) as fh:
        j = json.load(fh)
    # print every 0.01 time units (about 600 steps), a snapshot every 0.005,
    # stop at 2500 steps; stabilisation of 5 + 1 rounds of 30 steps
    j["constants"]["K"]["value"] = 100.0
    j["snapShot"] = {"filename": "bubbles", "frequency": 2.0}
    j["run"]["maxSteps"] = 2500
    j["stabilization"] = {"steps": 30, "maxDeltaEnergy": 1e9}
    path = str(tmp_path / "input.json")
    with open(path, "w") as fh:
        json.dump(j, fh)
    monkeypatch.chdir(tmp_path)
    print("cli run", flush=True)
    capfd.readouterr()
    assert ca.run_cli(["cubble-cli", path]) == 0
    ctypes.CDLL(None).fflush(None)  # the library prints through C stdio
    out = capfd.readouterr().out.splitlines()
    st = out.index("Stabilization") + 2
    assert out[st] == "%-7s %-11s %-11s %-11s %-9s" % ("#steps", "dE", "e1", "e2", "#searches")
    head = out.index("Integration") + 2
    assert out[head] == "%-5s %-8s %-6s %-11s %9s %10s %-6s %-9s %-11s %-11s %-11s " % (
        "T", "phi", "R", "dE", "#b   ", "#pairs", "#steps", "#searches", "min ts", "max ts",
        "avg ts")
    rows = []
    for line in out[head + 1:]:
        if line.startswith("Stopping simulation"):
            break
        rows.append(line)
    assert len(rows) >= 2, out[-5:]
    assert out[-2:] == ["Cleaning up...", "Done"]
    with open(tmp_path / "results.dat") as fh:
        res = fh.read().splitlines()
    assert len(res) == len(rows)
    for k, line in enumerate(res):
        f = line.split(" ")
        assert len(f) == 11 and int(f[0]) == k + 1
    # the first rows, replayed one step at a time, print the same text
    n = 2
    stab, table, results = _replay(path, n)
    assert out[st + 1:st + 6] == stab
    assert rows[:n] == table
    assert res[:n] == results
    # snapshots: 0 at the start, one per 0.005 time units, one at the end
    snaps = sorted(glob.glob(str(tmp_path / "bubbles.csv.*")), key=lambda p: int(p.rsplit(".", 1)[1]))
    assert len(snaps) >= 2 * len(rows) + 2
    hdr = "x,y,z,r,vx,vy,vz,vtot,vr,path,energy,displacement,error,index"
    first = np.genfromtxt(snaps[0], delimiter=",", names=True)
    last = np.genfromtxt(snaps[-1], delimiter=",", names=True)
    for p in (snaps[0], snaps[-1]):
        with open(p) as fh:
            assert fh.readline().strip() == hdr
    assert len(first) == j["bubbles"]["numStart"]
    assert not np.any(first["displacement"]) and not np.any(first["path"])
    assert np.all(first["energy"] >= 0) and first["energy"].sum() > 0
    assert int(res[-1].split(" ")[2]) >= len(last)
    assert np.all(np.diff(np.sort(last["index"])) > 0)
    assert np.all(last["displacement"] >= 0)


def test_reset_clock():
    sim = ca.Simulator(config(2, 3000, seed=22))
    sim.run_steps(200, GAS)
    st, sc = sim.get_state()
    assert sc["time_int"] + sc["time_frac"] > 0
    sim.reset_clock()
    st2, sc2 = sim.get_state()
    assert sc2["time_int"] == 0 and sc2["time_frac"] == 0.0
    assert np.array_equal(st2["wc"], st["wc"]) and np.array_equal(st2["x"], st["x"])
    assert sc2["steps"] == sc["steps"] and sc2["dt"] == sc["dt"]
    sim.reset_clock(clear_wraps=True)
    assert not np.any(sim.get_state()[0]["wc"])
    sim.close()


@pytest.mark.parametrize("precision", [ca.PREC_FP64, ca.PREC_MIXED])
def test_time_kernels_leaves_the_run_unchanged(precision):
    """The benchmark's kernel timer re-runs the attempt kernels back to back;
    the state, the counters and the rest of the run stay those of an untimed
    twin (only the per-bubble error, a diagnostic maximum, may grow)."""
    cfg = config(3, 20000, seed=23, precision=precision)
    sim, twin = ca.Simulator(cfg), ca.Simulator(cfg)
    sim.run_steps(20, GAS)
    twin.run_steps(20, GAS)
    st, sc = sim.get_state()
    t = sim.time_kernels(5, GAS)
    assert t["reps"] == 5 and min(t["predict_ms"], t["pair_ms"], t["correct_ms"]) > 0
    st2, sc2 = sim.get_state()
    for k in ("dt", "steps", "attempts", "time_int", "time_frac", "count", "searches"):
        assert sc2[k] == sc[k], k
    for k in ("x", "v", "vo", "saved", "nbr", "nbr_off", "idx"):
        assert np.array_equal(st2[k], st[k]), k
    sim.run_steps(30, GAS)
    twin.run_steps(30, GAS)
    a, b = sim.get_state()[0], twin.get_state()[0]
    for k in ("x", "v", "nbr", "idx"):
        assert np.array_equal(a[k], b[k]), k
    sim.close()
    twin.close()


def test_cli_runs_the_imposed_flow(tmp_path, monkeypatch, capfd):
    """flow.impose = 1 in the input runs the loop with the flow: the mean
    x velocity in results.dat follows the imposed flow's direction."""
    src = os.path.join(os.path.dirname(__file__), "golden", "input_small_2d.json")
    with open(src) as fh:
        j = json.load(fh)
    j["constants"]["K"]["value"] = 100.0
    j["run"]["maxSteps"] = 1500
    j["stabilization"] = {"steps": 0, "maxDeltaEnergy": 1.0}
    means = {}
    for impose in (0, 1):
        d = tmp_path / ("flow%d" % impose)
        d.mkdir()
        j["flow"] = {"impose": impose, "velocity": {"x": 0.5, "y": 0.0, "z": 0.0},
                     "lbb": {"x": 0.0, "y": 0.0, "z": 0.0}, "tfr": {"x": 1.0, "y": 0.5, "z": 0.0}}
        path = str(d / "input.json")
        with open(path, "w") as fh:
            json.dump(j, fh)
        monkeypatch.chdir(d)
        print("cli flow", impose, flush=True)
        assert ca.run_cli(["cubble-cli", path]) == 0
        with open(d / "results.dat") as fh:
            rows = [line.split(" ") for line in fh.read().splitlines()]
        assert rows
        means[impose] = float(rows[-1][6])   # mean dx/dt
    capfd.readouterr()
    assert means[1] > means[0] + 0.05


def test_cli_lattice_start_scales_the_box(tmp_path, monkeypatch, capfd):
    src = os.path.join(os.path.dirname(__file__), "golden", "input_small_2d.json")
    with open(src) as fh:
        j = json.load(fh)
    j["init"]["placement"] = "grid"
    j["stabilization"] = {"steps": 10, "maxDeltaEnergy": 1e9}
    j["run"]["maxSteps"] = 100
    path = str(tmp_path / "input.json")
    with open(path, "w") as fh:
        json.dump(j, fh)
    monkeypatch.chdir(tmp_path)
    capfd.readouterr()
    assert ca.run_cli(["cubble-cli", path]) == 0
    ctypes.CDLL(None).fflush(None)
    out = capfd.readouterr().out.splitlines()
    i = out.index("Scaling the simulation box")
    assert out.index("Stabilizing a few rounds after creation") < i
    first, second = out[i + 1], out[i + 2]
    assert first.startswith("Current phi: ") and second.startswith("Current phi: ")
    phi0 = float(first.split()[2].rstrip(","))
    phi1 = float(second.split()[2].rstrip(","))
    assert phi0 < 0.85 and phi1 == pytest.approx(0.9, abs=1e-9)
    assert out[i + 3] == "Neighbor search after scaling"
    assert out[i + 4] == "Stabilizing a few rounds after scaling"
    assert out[-1] == "Done"
