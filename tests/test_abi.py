"""CPU-only checks of the product library: it loads, exports every symbol the
C header declares, its host-side initial foam equals the oracle's bit for bit,
and it reads the reference's input_parameters.json schema.  No GPU compute."""
import ctypes
import os
import re

import numpy as np
import pytest

import cubble_amd as ca

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def header_functions():
    text = open(ca.HEADER).read()
    return sorted(set(re.findall(
        r"^\s*(?:int|int64_t|void|const char \*)\s*\*?\s*(cubble_\w+)\s*\(", text, re.M)))


def test_library_loads_and_exports_header_symbols():
    L = ca.lib()
    declared = header_functions()
    assert len(declared) >= 15
    for name in declared:
        assert hasattr(L, name), name
    assert sorted(declared) == sorted(ca.EXPORTS)
    assert "gfx950" in ca.version()


def test_no_torch_types_in_header():
    text = open(ca.HEADER).read()
    for token in ("torch::", "at::", "c10::", "Tensor", "#include <torch"):
        assert token not in text


@pytest.mark.parametrize("dim,dist,n,phi,seed", [(2, 0, 3000, 0.9, 426), (3, 1, 4000, 0.84, 11),
                                                 (3, 0, 1500, 0.95, 3)])
def test_host_initial_foam_matches_oracle(ora, dim, dist, n, phi, seed):
    r_std = 0.1 if dist == 0 else 0.15
    cfg = ca.Config(dim=dim, num_start=n, radius_dist=dist, r_std=r_std, r_lo=0.5, r_hi=1.5,
                    phi=phi, kappa=0.01, k_param=0.1, error_tolerance=1e-4, seed=seed)
    x4, sc = ca.generate_foam(cfg)
    f = ora.Foam.synthetic(dim=dim, count=n, phi=phi, seed=seed, radius_dist=dist, r_std=r_std,
                           kappa=0.01, k_param=0.1, tol=1e-4)
    st, osc = f.get_state()
    c = f.consts()
    order = np.argsort(st["idx"])
    for k in list(range(dim)) + [3]:
        assert np.array_equal(st["x"][k][order], x4[k]), k
    for key in ("interval", "tfr", "lbb"):
        assert list(sc[key]) == list(c[key])
    for key in ("skin", "min_rad", "f0", "k_param", "kappa", "avg_area_in", "vol_mul", "tol",
                "cell_cap"):
        assert sc[key] == c[key], key
    assert sc["max_r_cell"] == osc["max_r_cell"]


def test_json_schema_reader():
    cfg = ca.Config.from_json(os.path.join(GOLDEN, "input_small_2d.json"))
    assert cfg.dim == 2 and cfg.num_start == 2000 and cfg.num_end == 1000
    assert cfg.r_mean == 1.0 and cfg.r_std == 0.1 and cfg.phi == 0.9
    assert cfg.kappa == 0.01 and cfg.k_param == 0.1 and cfg.error_tolerance == 1e-4
    assert cfg.seed == 426 and cfg.max_steps == 200 and cfg.radius_dist == 0
    assert (cfg.r_lo, cfg.r_hi) == (0.5, 1.5) and tuple(cfg.wall) == (0, 0, 0)
    c3 = ca.Config.from_json(os.path.join(GOLDEN, "input_3d_lognormal.json"))
    assert c3.dim == 3 and c3.radius_dist == 1 and c3.precision == ca.PREC_MIXED
    assert c3.seed == 7 and c3.phi == 0.84


def test_json_flow_keys():
    # flow.impose only: the other flow keys keep input_parameters.json's values
    c = ca.Config.from_json(os.path.join(GOLDEN, "input_flow.json"))
    assert c.flow == 1
    assert c.flow_vel == (0.01, 0.0, 0.0)
    assert c.flow_lbb == (0.0, 0.475, 0.0) and c.flow_tfr == (0.4, 0.525, 0.0)
    c = ca.Config.from_json(os.path.join(GOLDEN, "input_small_2d.json"))
    assert c.flow == 0


def test_json_errors_are_reported():
    with pytest.raises(ca.CubbleError) as e:
        ca.Config.from_json(os.path.join(GOLDEN, "does_not_exist.json"))
    assert e.value.code == ca.E_IO


def test_invalid_config_rejected_on_host():
    with pytest.raises(ca.CubbleError) as e:
        ca.generate_foam(ca.Config(dim=4, num_start=10))
    assert e.value.code == ca.E_INVALID
    with pytest.raises(ca.CubbleError):
        ca.generate_foam(ca.Config(dim=2, num_start=10, phi=1.5))


def test_cli_usage_error_without_device_work():
    # wrong argument count: returns EXIT_FAILURE before touching the GPU
    assert ca.run_cli(["cubble-cli"]) == 1


# --------------------------------------------------------------------------
# the reference's lattice placement (init.placement "grid")
# --------------------------------------------------------------------------
@pytest.mark.parametrize("dim,n,rel,want", [
    (2, 2000, (1.0, 1.0, 1.0), 45 * 45), (2, 2025, (1.0, 1.0, 1.0), 45 * 45),
    (2, 2026, (1.0, 1.0, 1.0), 46 * 46), (2, 2000, (2.0, 1.0, 1.0), 64 * 32),
    (3, 1000, (1.0, 1.0, 1.0), 10 ** 3), (3, 1001, (1.0, 1.0, 1.0), 11 ** 3),
    (3, 5000, (1.0, 1.0, 1.0), 18 ** 3)])
def test_grid_placement_rounds_to_the_lattice(tmp_path, dim, n, rel, want):
    # numStart rounds up to a full lattice (src/simulator.cu:576-600); a
    # programmatic config must already hold the lattice count
    j = {"bubbles": {"numStart": n, "numEnd": 10, "radius": {"mean": 1.0, "std": 0.1}},
         "box": {"dimensionality": dim,
                 "relativeDimensions": {"x": rel[0], "y": rel[1], "z": rel[2]}},
         "init": {"placement": "grid"}}
    p = tmp_path / "in.json"
    p.write_text(__import__("json").dumps(j))
    c = ca.Config.from_json(str(p))
    assert c.placement == 1 and c.num_start == want
    if want != n:
        with pytest.raises(ca.CubbleError) as e:
            ca.generate_foam(ca.Config(dim=dim, num_start=n, placement=1, rel_dim=rel))
        assert e.value.code == ca.E_INVALID and str(want) in str(e.value)


@pytest.mark.parametrize("dim,bpd", [(2, (45, 45, 1)), (3, (12, 12, 12))])
def test_grid_placement_positions(dim, bpd):
    """Every bubble sits within R (scaled to the box) of its lattice site, in
    the positive orthant of it (modulo the box), and the box holds phi
    exactly."""
    n = bpd[0] * bpd[1] * bpd[2]
    cfg = ca.Config(dim=dim, num_start=n, placement=1, phi=0.9, r_mean=1.0, r_std=0.1)
    x4, sc = ca.generate_foam(cfg)
    L = np.array(sc["interval"][:dim])
    i = np.arange(n)
    site = np.stack([(i % bpd[0]) / bpd[0], ((i // bpd[0]) % bpd[1]) / bpd[1],
                     (i // (bpd[0] * bpd[1])) / bpd[2]])[:dim] * L[:, None]
    d = (x4[:dim] - site) % L[:, None]                   # offset, folded into the box
    scale = L / (2.0 * np.array(bpd[:dim]))              # R in box units, per axis
    assert np.all(d / scale[:, None] <= 1.0 + 1e-9)
    assert np.all(np.sqrt(((d / scale[:, None]) ** 2).sum(0)) <= 1.0 + 1e-9)
    assert np.all((x4[:dim] >= 0) & (x4[:dim] <= L[:, None]))
    r = x4[3]
    vol = np.pi * r ** 2 * (4.0 / 3.0 * r if dim == 3 else 1.0)
    assert vol.sum() == pytest.approx(0.9 * np.prod(L), rel=1e-12)
    # not all offsets vanish
    assert np.std(d) > 0.1 * scale.min()


def test_new_entry_points_reject_a_null_handle():
    """Argument checks come before any device work, so they run without a GPU."""
    L = ca.lib()
    obs = ca._Observables()
    rules = ca._StopRules()
    rel = np.ones(3)
    assert L.cubble_observe(None, ctypes.byref(obs), None) == ca.E_INVALID
    assert L.cubble_set_stop_rules(None, ctypes.byref(rules)) == ca.E_INVALID
    assert L.cubble_reset_print_window(None) == ca.E_INVALID
    assert L.cubble_reset_clock(None, 1) == ca.E_INVALID
    assert L.cubble_bootstrap_velocities(None) == ca.E_INVALID
    assert L.cubble_rescale_box(None, 0.9, rel.ctypes.data_as(ca._dp)) == ca.E_INVALID


# --------------------------------------------------------------------------
# the reference's own input: no "init" object
# --------------------------------------------------------------------------
def test_reference_input_defaults_to_lattice_and_reference_radii(tmp_path):
    # input_parameters.json has no "init" object: the reference's init places
    # numStart (rounded up to the lattice) on the lattice and draws
    # |N(mean, std)| radii (src/simulator.cu:576-606, src/kernels.cu:1034-1053)
    j = {"bubbles": {"numStart": 1e5, "numEnd": 1000, "radius": {"mean": 1.0, "std": 0.21}},
         "box": {"dimensionality": 2, "relativeDimensions": {"x": 1, "y": 1, "z": 1}},
         "constants": {"phi": {"value": 0.9}}}
    p = tmp_path / "input_parameters.json"
    p.write_text(__import__("json").dumps(j))
    c = ca.Config.from_json(str(p))
    assert c.placement == 1 and c.radius_dist == 2
    assert c.num_start == 317 * 317  # ceil(sqrt(1e5)) per axis
    # an "init" object, even an empty one, keeps uniform placement and the
    # clipped normal unless its keys say otherwise
    j["init"] = {}
    p.write_text(__import__("json").dumps(j))
    c = ca.Config.from_json(str(p))
    assert c.placement == 0 and c.radius_dist == 0 and c.num_start == 100000
    j["init"] = {"radiusDistribution": "reference", "placement": "grid"}
    p.write_text(__import__("json").dumps(j))
    c = ca.Config.from_json(str(p))
    assert c.placement == 1 and c.radius_dist == 2


@pytest.mark.parametrize("dim", [2, 3])
def test_reference_radius_rule_on_drawn_values(dim):
    """radius_dist 2 applies |v| and +minRad (v at or below minRad) to the
    very normals the unclipped normal draws from the same seed."""
    n, mean, std = 20000, 1.0, 0.6
    base = dict(dim=dim, num_start=n, r_mean=mean, r_std=std, phi=0.5, seed=77)
    raw, _ = ca.generate_foam(ca.Config(radius_dist=0, r_lo=-1e300, r_hi=1e300, **base))
    got, sc = ca.generate_foam(ca.Config(radius_dist=2, **base))
    v = raw[3]
    min_rad = 0.1 * mean
    assert sc["min_rad"] == min_rad
    a = np.abs(v)
    want = np.where(a > min_rad, a, a + min_rad)
    assert np.array_equal(got[3], want)
    # every branch was drawn: negative values, and |v| at or below minRad
    assert np.count_nonzero(v < 0) > 100 and np.count_nonzero(a <= min_rad) > 50
    assert np.all(got[3] > min_rad)


@pytest.mark.parametrize("dim", [2, 3])
def test_reference_radius_rule_matches_oracle_init(ora, dim):
    n = 3000
    cfg = ca.Config(dim=dim, num_start=n, radius_dist=2, r_std=0.6, phi=0.8, seed=5)
    x4, sc = ca.generate_foam(cfg)
    f = ora.Foam.synthetic(dim=dim, count=n, phi=0.8, seed=5, radius_dist=2, r_std=0.6)
    st, _ = f.get_state()
    order = np.argsort(st["idx"])
    assert np.array_equal(st["x"][3][order], x4[3])
    assert sc["avg_area_in"] == f.consts()["avg_area_in"]
