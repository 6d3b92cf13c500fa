"""Shared helpers for the GPU parity tests: paired product / oracle foams and
the error metric.

Tolerance metric (per quantity q, over bubbles i):
    err_q = max_i |q_gpu,i - q_oracle,i| / max(|q_oracle,i|, s_q)
with s_q = f0 for velocities and s_q = max_i |dR/dt_oracle,i| for dR/dt, so
that entries whose true value cancels to ~0 are judged against the scale of
the quantity instead of themselves.
"""
import numpy as np

import cubble_amd as ca

FP64_TOL = 1e-12      # FP64 GPU path vs the FP64 oracle
MIXED_TOL = 1e-5      # mixed-precision path vs the FP64 oracle
VOLUME_TOL = 1e-10    # total gas volume


def config(dim, n, dist=None, phi=None, seed=426, precision=ca.PREC_FP64, **over):
    dist = (0 if dim == 2 else 1) if dist is None else dist
    phi = (0.9 if dim == 2 else 0.84) if phi is None else phi
    kw = dict(dim=dim, num_start=n, radius_dist=dist, r_std=0.1 if dist == 0 else 0.15,
              r_lo=0.5, r_hi=1.5, phi=phi, kappa=0.01, k_param=0.1, error_tolerance=1e-4,
              dt0=1e-4, seed=seed, precision=precision)
    kw.update(over)
    return ca.Config(**kw)


def oracle_for(ora, cfg):
    return ora.Foam.synthetic(dim=cfg.dim, count=cfg.num_start, radius_dist=cfg.radius_dist,
                              r_mean=cfg.r_mean, r_std=cfg.r_std, r_lo=cfg.r_lo, r_hi=cfg.r_hi,
                              phi=cfg.phi, kappa=cfg.kappa, k_param=cfg.k_param, mu=cfg.mu,
                              sigma=cfg.sigma, tol=cfg.error_tolerance, dt0=cfg.dt0,
                              avg_rad=cfg.r_mean, seed=cfg.seed, rel_dim=cfg.rel_dim,
                              wall=cfg.wall, wall_drag=cfg.wall_drag,
                              cell_cap_factor=cfg.cell_cap_factor)


def oracle_from_sim(ora, sim):
    """An oracle foam holding exactly the simulator's current state."""
    st, sc = sim.get_state()
    consts = {k: sc[k] for k in ("dim", "wall", "lbb", "tfr", "interval", "wall_drag", "skin",
                                 "min_rad", "f0", "k_param", "kappa", "avg_area_in", "vol_mul",
                                 "tol", "cell_cap", "flow_lbb", "flow_tfr", "flow_vel")}
    osc = {k: sc[k] for k in ("count", "cell_dim", "dt", "time_int", "time_frac", "max_r_cell",
                              "steps", "attempts", "searches", "removals", "bubble_steps")}
    for k in ("s_a", "s_ar", "s_oa", "s_oar", "s_v", "max_err", "max_r", "max_exp"):
        osc[k] = 0.0
    for k in ("n_del", "accepted", "need_search", "num_pairs"):
        osc[k] = 0
    return ora.Foam.from_state(consts, st, osc)


def scaled_err(gpu, ref, scale):
    gpu = np.asarray(gpu, float)
    ref = np.asarray(ref, float)
    if gpu.size == 0:
        return 0.0
    den = np.maximum(np.abs(ref), scale)
    return float(np.max(np.abs(gpu - ref) / den))


def attempt_errors(dim, g, o, f0):
    """Per-quantity scaled errors of one attempt's outputs."""
    out = {}
    out["velocity"] = max(scaled_err(g["vp"][k], o["vp"][k], f0) for k in range(dim))
    s_r = max(float(np.max(np.abs(o["vp"][3]))), 1e-300)
    out["drdt"] = scaled_err(g["vp"][3], o["vp"][3], s_r)
    out["position"] = max(scaled_err(g["xc"][k], o["xc"][k], 1.0) for k in range(dim))
    out["radius"] = scaled_err(g["xc"][3], o["xc"][3], 1.0)
    s_ov = max(float(np.max(np.abs(o["ov"]))), 1e-300)
    out["overlap"] = scaled_err(g["ov"], o["ov"], s_ov)
    # the path grows by |x_c - x|, a difference of coordinates of size up to
    # the box, so it carries their absolute rounding: it is held to the
    # positions' bar, relative to the largest coordinate
    s_x = max(float(np.max(np.abs(o["xc"][k]))) for k in range(dim))
    out["path"] = scaled_err(g["path_new"], o["path_new"], max(s_x, 1e-3))
    return out


def assert_same_structure(gst, ost):
    """Integer state and neighbour lists must match bit for bit."""
    assert np.array_equal(gst["idx"], ost["idx"])
    assert np.array_equal(gst["wc"], ost["wc"])
    assert np.array_equal(gst["nbr_off"], ost["nbr_off"])
    assert np.array_equal(gst["nbr"], ost["nbr"])


def total_volume(dim, r):
    r = np.asarray(r)
    return float(np.sum(np.pi * r * r)) if dim == 2 else float(np.sum(4.0 / 3.0 * np.pi * r ** 3))
