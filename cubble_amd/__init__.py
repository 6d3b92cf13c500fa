"""cubble_amd: MI355X-native per-timestep path of the cubble wet-foam simulator.

Python mirror of the reference's simulator step interface (KJLankinen/cubble,
src/simulator.cu:60-337) over the C ABI in include/cubble_amd.h.  Everything
computational runs in the HIP library ``libcubble_amd.so`` built in-tree by
``python -m cubble_amd.build``; there is no CPU fallback -- if the library is
missing, every call raises.

    sim = Simulator(Config.workload("foam3d_1e6"))
    res = sim.integrate(IntegrationParams(use_gas_exchange=True, increment_path=True))
    sim.run_steps(1000)
    state, scalars = sim.get_state()
"""
import ctypes
import dataclasses
import math
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libcubble_amd.so")
HEADER = os.path.join(os.path.dirname(_HERE), "include", "cubble_amd.h")

OK = 0
E_INVALID, E_HIP, E_NOMEM, E_IO, E_UNSUPPORTED, E_CAPACITY, E_STATE = -1, -2, -3, -4, -5, -6, -7
FIND_PAIRS_NOALLOC_MAX = 800000  # CUBBLE_FIND_PAIRS_NOALLOC_MAX
PREC_FP64, PREC_MIXED = 0, 1

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int32)
_u8p = ctypes.POINTER(ctypes.c_uint8)


class CubbleError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("cubble_amd error %d: %s" % (code, msg))
        self.code = code


class _IntegrateParams(ctypes.Structure):
    _fields_ = [("use_gas_exchange", ctypes.c_int32), ("use_flow", ctypes.c_int32),
                ("increment_path", ctypes.c_int32)]


class _StepResult(ctypes.Structure):
    _fields_ = [
        ("error_too_large", ctypes.c_int32), ("num_to_be_deleted", ctypes.c_int32),
        ("count", ctypes.c_int32), ("num_pairs", ctypes.c_int32),
        ("need_search", ctypes.c_int32), ("halted", ctypes.c_int32),
        ("max_error", ctypes.c_double), ("max_radius", ctypes.c_double),
        ("max_expansion", ctypes.c_double), ("dt", ctypes.c_double), ("time", ctypes.c_double),
        ("steps", ctypes.c_int64), ("attempts", ctypes.c_int64), ("searches", ctypes.c_int64),
        ("removals", ctypes.c_int64), ("bubble_steps", ctypes.c_double),
    ]


class _Config(ctypes.Structure):
    _fields_ = [
        ("dim", ctypes.c_int32), ("num_start", ctypes.c_int32), ("num_end", ctypes.c_int32),
        ("radius_dist", ctypes.c_int32), ("r_mean", ctypes.c_double), ("r_std", ctypes.c_double),
        ("r_lo", ctypes.c_double), ("r_hi", ctypes.c_double), ("rel_dim", ctypes.c_double * 3),
        ("wall", ctypes.c_int32 * 3), ("wall_drag", ctypes.c_double), ("phi", ctypes.c_double),
        ("mu", ctypes.c_double), ("sigma", ctypes.c_double), ("kappa", ctypes.c_double),
        ("k_param", ctypes.c_double), ("error_tolerance", ctypes.c_double),
        ("dt0", ctypes.c_double), ("seed", ctypes.c_uint64), ("precision", ctypes.c_int32),
        ("nbr_cap", ctypes.c_int32), ("cell_cap_factor", ctypes.c_double),
        ("max_steps", ctypes.c_int64), ("flow", ctypes.c_int32),
        ("flow_vel", ctypes.c_double * 3), ("flow_lbb", ctypes.c_double * 3),
        ("flow_tfr", ctypes.c_double * 3), ("placement", ctypes.c_int32),
    ]


class _StateView(ctypes.Structure):
    _fields_ = [("x", _dp * 4), ("v", _dp * 4), ("vo", _dp * 4), ("saved", _dp * 4),
                ("path", _dp), ("error", _dp), ("wrap", _ip * 3), ("index", _ip),
                ("nbr_off", _ip), ("nbr", _ip), ("nbr_cap", ctypes.c_int64),
                ("degree", _ip)]


class _AttemptView(ctypes.Structure):
    _fields_ = [("xc", _dp * 4), ("vp", _dp * 4), ("overlap", _dp), ("path_new", _dp),
                ("deleted", _u8p)]


class _Scalars(ctypes.Structure):
    _fields_ = [
        ("dim", ctypes.c_int32), ("count", ctypes.c_int32), ("capacity", ctypes.c_int32),
        ("nbr_cap", ctypes.c_int32), ("cell_cap", ctypes.c_int32),
        ("cell_dim", ctypes.c_int32 * 3), ("wall", ctypes.c_int32 * 3),
        ("precision", ctypes.c_int32),
        ("lbb", ctypes.c_double * 3), ("tfr", ctypes.c_double * 3),
        ("interval", ctypes.c_double * 3),
        ("wall_drag", ctypes.c_double), ("skin", ctypes.c_double), ("min_rad", ctypes.c_double),
        ("f0", ctypes.c_double), ("k_param", ctypes.c_double), ("kappa", ctypes.c_double),
        ("avg_area_in", ctypes.c_double), ("vol_mul", ctypes.c_double), ("tol", ctypes.c_double),
        ("dt", ctypes.c_double), ("time_int", ctypes.c_uint64), ("time_frac", ctypes.c_double),
        ("max_r_cell", ctypes.c_double), ("steps", ctypes.c_int64),
        ("attempts", ctypes.c_int64), ("searches", ctypes.c_int64),
        ("removals", ctypes.c_int64), ("bubble_steps", ctypes.c_double),
        ("s_a", ctypes.c_double), ("s_ar", ctypes.c_double), ("s_oa", ctypes.c_double),
        ("s_oar", ctypes.c_double), ("s_v", ctypes.c_double), ("max_err", ctypes.c_double),
        ("max_r", ctypes.c_double), ("max_exp", ctypes.c_double), ("n_del", ctypes.c_int32),
        ("accepted", ctypes.c_int32), ("need_search", ctypes.c_int32),
        ("num_pairs", ctypes.c_int32), ("ts_min", ctypes.c_double), ("ts_max", ctypes.c_double),
        ("ts_sum", ctypes.c_double), ("flow_lbb", ctypes.c_double * 3),
        ("flow_tfr", ctypes.c_double * 3), ("flow_vel", ctypes.c_double * 3),
    ]


class _Observables(ctypes.Structure):
    _fields_ = [("total_energy", ctypes.c_double), ("total_volume", ctypes.c_double),
                ("box_volume", ctypes.c_double), ("mean_r", ctypes.c_double),
                ("mean_path", ctypes.c_double), ("mean_v", ctypes.c_double * 4)]


class _StopRules(ctypes.Structure):
    _fields_ = [("min_count", ctypes.c_int32), ("use_time", ctypes.c_int32),
                ("max_radius", ctypes.c_double), ("time_int", ctypes.c_uint64),
                ("time_frac", ctypes.c_double)]


class _KernelTimes(ctypes.Structure):
    _fields_ = [("predict_ms", ctypes.c_double), ("pair_ms", ctypes.c_double),
                ("correct_ms", ctypes.c_double), ("reps", ctypes.c_int32)]


class Box(ctypes.Structure):
    """cubble_box: dimensionality, walls (1 = wall), lower and upper corners."""
    _fields_ = [("dim", ctypes.c_int32), ("wall", ctypes.c_int32 * 3),
                ("lbb", ctypes.c_double * 3), ("tfr", ctypes.c_double * 3)]

    @classmethod
    def of(cls, dim, lbb, tfr, wall=(0, 0, 0)):
        b = cls()
        b.dim = dim
        for k in range(3):
            b.wall[k] = int(wall[k])
            b.lbb[k] = float(lbb[k]) if k < len(lbb) else 0.0
            b.tfr[k] = float(tfr[k]) if k < len(tfr) else 0.0
        return b


class _SearchTimes(ctypes.Structure):
    _fields_ = [("search_ms", ctypes.c_double), ("lists_ms", ctypes.c_double),
                ("list_kernel_ms", ctypes.c_double), ("reorder_ms", ctypes.c_double),
                ("idle_ms", ctypes.c_double), ("bubbles", ctypes.c_double),
                ("pairs", ctypes.c_double), ("reps", ctypes.c_int32)]


def _search_times(t):
    return {name: getattr(t, name) for name, _ in _SearchTimes._fields_}


class _Tuning(ctypes.Structure):
    _fields_ = [("bubbles_per_thread", ctypes.c_int32), ("search_tile", ctypes.c_int32),
                ("graph_steps", ctypes.c_int32)]


TUNING_KEYS = ("bubbles_per_thread", "search_tile", "graph_steps")


def _tuning(t):
    """None, a dict with TUNING_KEYS (missing = 0, the built-in default) or a
    sequence in that order."""
    if t is None:
        return None
    if isinstance(t, dict):
        return ctypes.byref(_Tuning(*[int(t.get(k, 0)) for k in TUNING_KEYS]))
    vals = [int(v) for v in t] + [0] * (3 - len(t))
    return ctypes.byref(_Tuning(*vals[:3]))


class _AutotuneOpts(ctypes.Structure):
    _fields_ = [("trial_steps", ctypes.c_int32), ("repeats", ctypes.c_int32),
                ("ignore_cache", ctypes.c_int32), ("cache_path", ctypes.c_char_p)]


class _AutotuneReport(ctypes.Structure):
    _fields_ = [("chosen", _Tuning), ("from_cache", ctypes.c_int32),
                ("candidates_timed", ctypes.c_int32), ("start_rate", ctypes.c_double),
                ("best_rate", ctypes.c_double)]


def _autotune(fn, h, ip, trial_steps, repeats, cache_path, ignore_cache, check):
    o = _AutotuneOpts(int(trial_steps), int(repeats), int(bool(ignore_cache)),
                      cache_path.encode() if cache_path else None)
    r = _AutotuneReport()
    check(fn(h, ctypes.byref((ip or IntegrationParams())._c()), ctypes.byref(o), ctypes.byref(r)))
    return {"chosen": {k: getattr(r.chosen, k) for k in TUNING_KEYS},
            "from_cache": bool(r.from_cache), "candidates_timed": r.candidates_timed,
            "start_rate": r.start_rate, "best_rate": r.best_rate}


# exported symbols of include/cubble_amd.h (checked by the ABI test)
EXPORTS = [
    "cubble_create", "cubble_create_from_json", "cubble_config_from_json",
    "cubble_config_default", "cubble_destroy", "cubble_last_error", "cubble_step",
    "cubble_integrate", "cubble_run_steps", "cubble_search_neighbors", "cubble_remove_bubbles",
    "cubble_get_scalars", "cubble_get_state", "cubble_get_attempt", "cubble_set_state",
    "cubble_generate_foam", "cubble_run", "cubble_version", "cubble_time_kernels",
    "cubble_create_tuned", "cubble_batch_create", "cubble_batch_size", "cubble_batch_integrate",
    "cubble_batch_run_steps", "cubble_batch_get_scalars", "cubble_batch_get_state",
    "cubble_batch_last_error", "cubble_batch_destroy", "cubble_batch_time_kernels",
    "cubble_time_search", "cubble_batch_time_search", "cubble_find_pairs",
    "cubble_find_pairs_noalloc",
    "cubble_server_create", "cubble_server_submit", "cubble_server_run", "cubble_server_result",
    "cubble_server_get_state", "cubble_server_last_error", "cubble_server_destroy",
    "cubble_foam_footprint", "cubble_autotune", "cubble_retune", "cubble_batch_autotune",
    "cubble_batch_retune", "cubble_observe", "cubble_set_stop_rules",
    "cubble_reset_print_window", "cubble_reset_clock", "cubble_bootstrap_velocities",
    "cubble_rescale_box",
]


class _ServerOpts(ctypes.Structure):
    _fields_ = [("hbm_budget_bytes", ctypes.c_uint64), ("max_active", ctypes.c_int32),
                ("round_steps", ctypes.c_int32), ("keep_states", ctypes.c_int32),
                ("precision", ctypes.c_int32)]


class _RequestResult(ctypes.Structure):
    _fields_ = [("id", ctypes.c_int64), ("status", ctypes.c_int32),
                ("start_count", ctypes.c_int32), ("final_count", ctypes.c_int32),
                ("pad", ctypes.c_int32), ("steps", ctypes.c_int64),
                ("bubble_steps", ctypes.c_double), ("arrival_s", ctypes.c_double),
                ("admit_s", ctypes.c_double), ("finish_s", ctypes.c_double)]


class _ServerStats(ctypes.Structure):
    _fields_ = [("requests", ctypes.c_int64), ("rounds", ctypes.c_int64),
                ("max_resident", ctypes.c_int32), ("pad", ctypes.c_int32),
                ("wall_s", ctypes.c_double), ("bubble_steps", ctypes.c_double),
                ("throughput", ctypes.c_double), ("p50_latency_s", ctypes.c_double),
                ("p99_latency_s", ctypes.c_double), ("peak_bytes", ctypes.c_double)]

_lib = None


def lib():
    """The loaded HIP library; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("libcubble_amd.so is not built (python -m cubble_amd.build); "
                               "there is no CPU fallback")
        L = ctypes.CDLL(LIB_PATH)
        vp = ctypes.c_void_p
        P = ctypes.POINTER
        L.cubble_create.argtypes = [P(_Config), P(vp)]
        L.cubble_create_from_json.argtypes = [ctypes.c_char_p, P(vp)]
        L.cubble_config_from_json.argtypes = [ctypes.c_char_p, P(_Config)]
        L.cubble_config_default.argtypes = [P(_Config)]
        L.cubble_destroy.argtypes = [vp]
        L.cubble_last_error.argtypes = [vp]
        L.cubble_last_error.restype = ctypes.c_char_p
        for name in ("cubble_step", "cubble_integrate"):
            getattr(L, name).argtypes = [vp, P(_IntegrateParams), P(_StepResult)]
        L.cubble_run_steps.argtypes = [vp, P(_IntegrateParams), ctypes.c_int64, P(_StepResult)]
        L.cubble_search_neighbors.argtypes = [vp]
        L.cubble_remove_bubbles.argtypes = [vp]
        L.cubble_get_scalars.argtypes = [vp, P(_Scalars)]
        L.cubble_get_state.argtypes = [vp, P(_StateView)]
        L.cubble_get_attempt.argtypes = [vp, P(_AttemptView)]
        L.cubble_set_state.argtypes = [vp, P(_StateView), P(_Scalars)]
        L.cubble_generate_foam.argtypes = [P(_Config), _dp, P(_Scalars)]
        L.cubble_run.argtypes = [ctypes.c_int, P(ctypes.c_char_p)]
        L.cubble_version.restype = ctypes.c_char_p
        L.cubble_time_kernels.argtypes = [vp, P(_IntegrateParams), ctypes.c_int32,
                                          P(_KernelTimes)]
        L.cubble_time_search.argtypes = [vp, ctypes.c_int32, P(_SearchTimes)]
        L.cubble_find_pairs.argtypes = [vp, vp, vp, vp, ctypes.c_int32, P(Box), ctypes.c_double,
                                        vp, vp, ctypes.c_int64, vp, vp]
        L.cubble_find_pairs_noalloc.argtypes = L.cubble_find_pairs.argtypes
        L.cubble_batch_time_search.argtypes = [vp, ctypes.c_int32, P(_SearchTimes)]
        L.cubble_observe.argtypes = [vp, P(_Observables), _dp]
        L.cubble_set_stop_rules.argtypes = [vp, P(_StopRules)]
        L.cubble_reset_print_window.argtypes = [vp]
        L.cubble_reset_clock.argtypes = [vp, ctypes.c_int32]
        L.cubble_bootstrap_velocities.argtypes = [vp]
        L.cubble_rescale_box.argtypes = [vp, ctypes.c_double, _dp]
        L.cubble_create_tuned.argtypes = [P(_Config), P(_Tuning), P(vp)]
        L.cubble_batch_create.argtypes = [P(_Config), ctypes.c_int32, P(_Tuning), P(vp)]
        L.cubble_batch_size.argtypes = [vp]
        L.cubble_batch_integrate.argtypes = [vp, P(_IntegrateParams), P(_StepResult)]
        L.cubble_batch_run_steps.argtypes = [vp, P(_IntegrateParams), ctypes.c_int64,
                                             P(_StepResult)]
        L.cubble_batch_get_scalars.argtypes = [vp, ctypes.c_int32, P(_Scalars)]
        L.cubble_batch_get_state.argtypes = [vp, ctypes.c_int32, P(_StateView)]
        L.cubble_batch_last_error.argtypes = [vp]
        L.cubble_batch_last_error.restype = ctypes.c_char_p
        L.cubble_batch_destroy.argtypes = [vp]
        L.cubble_batch_time_kernels.argtypes = [vp, P(_IntegrateParams), ctypes.c_int32,
                                                P(_KernelTimes)]
        L.cubble_server_create.argtypes = [P(_ServerOpts), P(vp)]
        L.cubble_server_submit.argtypes = [vp, P(_Config), ctypes.c_int64, ctypes.c_double,
                                           ctypes.c_double, P(ctypes.c_int64)]
        L.cubble_server_run.argtypes = [vp, P(_ServerStats)]
        L.cubble_server_result.argtypes = [vp, ctypes.c_int64, P(_RequestResult)]
        L.cubble_server_get_state.argtypes = [vp, ctypes.c_int64, P(_StateView), P(_Scalars)]
        L.cubble_server_last_error.argtypes = [vp]
        L.cubble_server_last_error.restype = ctypes.c_char_p
        L.cubble_server_destroy.argtypes = [vp]
        L.cubble_foam_footprint.argtypes = [P(_Config)]
        L.cubble_foam_footprint.restype = ctypes.c_int64
        for name in ("cubble_autotune", "cubble_batch_autotune"):
            getattr(L, name).argtypes = [vp, P(_IntegrateParams), P(_AutotuneOpts),
                                         P(_AutotuneReport)]
        for name in ("cubble_retune", "cubble_batch_retune"):
            getattr(L, name).argtypes = [vp, P(_Tuning)]
        _lib = L
    return _lib


def _check(rc, handle=None):
    if rc != OK:
        msg = lib().cubble_last_error(handle)
        raise CubbleError(rc, msg.decode() if msg else "")


@dataclasses.dataclass
class IntegrationParams:
    """Mirror of the reference's IntegrationParams flags (src/data_definitions.h:227-236)."""
    use_gas_exchange: bool = True
    use_flow: bool = False
    increment_path: bool = True

    def _c(self):
        return _IntegrateParams(int(self.use_gas_exchange), int(self.use_flow),
                                int(self.increment_path))


@dataclasses.dataclass
class Config:
    """Synthetic initial foam + constants; field meanings follow input_parameters.json."""
    dim: int = 2
    num_start: int = 100000
    num_end: int = 1000
    radius_dist: int = 0          # 0 clipped normal, 1 log-normal, 2 |normal| (+minRad)
    r_mean: float = 1.0
    r_std: float = 0.21
    r_lo: float = 0.5
    r_hi: float = 1.5
    rel_dim: tuple = (1.0, 1.0, 1.0)
    wall: tuple = (0, 0, 0)
    wall_drag: float = 0.1
    phi: float = 0.9
    mu: float = 1.0
    sigma: float = 1.0
    kappa: float = 0.1
    k_param: float = 0.001
    error_tolerance: float = 1e-5
    dt0: float = 1e-4
    seed: int = 426
    precision: int = PREC_FP64
    nbr_cap: int = 0
    cell_cap_factor: float = 2.0
    max_steps: int = 0
    flow: int = 0                 # flow.impose
    flow_vel: tuple = (0.01, 0.0, 0.0)
    flow_lbb: tuple = (0.0, 0.475, 0.0)
    flow_tfr: tuple = (0.4, 0.525, 0.0)
    placement: int = 0            # 0 uniform, 1 the reference's lattice

    _VEC = ("rel_dim", "wall", "flow_vel", "flow_lbb", "flow_tfr")

    def _c(self):
        c = _Config()
        for f in dataclasses.fields(self):
            v = getattr(self, f.name)
            if f.name in self._VEC:
                arr = getattr(c, f.name)
                for k in range(3):
                    arr[k] = v[k]
            else:
                setattr(c, f.name, v)
        return c

    @classmethod
    def _from_c(cls, c):
        kw = {}
        for f in dataclasses.fields(cls):
            v = getattr(c, f.name)
            kw[f.name] = tuple(v) if f.name in cls._VEC else v
        return cls(**kw)

    @classmethod
    def from_json(cls, path):
        c = _Config()
        _check(lib().cubble_config_from_json(path.encode(), ctypes.byref(c)))
        return cls._from_c(c)

    @classmethod
    def workload(cls, name, **over):
        """The named synthetic workloads of the benchmark configurations."""
        base = dict(mu=1.0, sigma=1.0, kappa=0.01, k_param=0.1, error_tolerance=1e-4, dt0=1e-4,
                    r_mean=1.0, seed=426)
        if name == "foam2d_1e4":       # 2D, normal(1, 0.1) clipped [0.5, 1.5], phi 0.9
            base.update(dim=2, num_start=10000, radius_dist=0, r_std=0.1, r_lo=0.5, r_hi=1.5,
                        phi=0.9)
        elif name == "foam3d_1e6":     # 3D, log-normal sigma 0.15, phi 0.84
            base.update(dim=3, num_start=1000000, radius_dist=1, r_std=0.15, phi=0.84)
        elif name == "foam3d_1e7":     # 3D, log-normal sigma 0.15, phi 0.9
            base.update(dim=3, num_start=10000000, radius_dist=1, r_std=0.15, phi=0.9)
        else:
            raise ValueError("unknown workload %r" % name)
        base.update(over)
        return cls(**base)


def batch_workload(num_foams=256, seed=2024, precision=PREC_MIXED, lo=1e3, hi=1e5):
    """The batched workload: independent foams, alternately 2D (normal radii,
    sd 0.1, clipped to [0.5, 1.5]) and 3D (log-normal radii, sigma 0.15),
    bubble counts log-uniform in [lo, hi], phi uniform in [0.8, 0.95], kappa
    log-uniform in [1e-3, 1e-1], K = 0.1, each with its own seed."""
    rng = np.random.default_rng(seed)
    out = []
    for k in range(num_foams):
        dim = 2 + (k % 2)
        n = int(round(10 ** rng.uniform(np.log10(lo), np.log10(hi))))
        phi = float(rng.uniform(0.8, 0.95))
        kappa = float(10 ** rng.uniform(-3, -1))
        dist = dict(radius_dist=0, r_std=0.1, r_lo=0.5, r_hi=1.5) if dim == 2 else \
            dict(radius_dist=1, r_std=0.15)
        out.append(Config(dim=dim, num_start=n, phi=phi, kappa=kappa, k_param=0.1, mu=1.0,
                          error_tolerance=1e-4, dt0=1e-4, seed=seed * 1000 + k,
                          precision=precision, **dist))
    return out


def serving_requests(num_requests, rate_per_s=50.0, seed=77, precision=PREC_MIXED, lo=1e3,
                     hi=1e6):
    """The serving workload: a Poisson stream of 3D foam requests, bubble counts
    log-uniform in [lo, hi], log-normal radii (sigma 0.15), phi uniform in
    [0.8, 0.95].  Returns [(config, arrival_s)]."""
    rng = np.random.default_rng(seed)
    t = 0.0
    out = []
    for k in range(num_requests):
        t += float(rng.exponential(1.0 / rate_per_s))
        n = int(round(10 ** rng.uniform(np.log10(lo), np.log10(hi))))
        phi = float(rng.uniform(0.8, 0.95))
        out.append((Config(dim=3, num_start=n, radius_dist=1, r_std=0.15, phi=phi, kappa=0.01,
                           k_param=0.1, mu=1.0, error_tolerance=1e-4, dt0=1e-4,
                           seed=seed * 1000 + k, precision=precision), t))
    return out


def _dptr(a):
    return a.ctypes.data_as(_dp)


def _iptr(a):
    return a.ctypes.data_as(_ip)


def _res(r):
    return {name: getattr(r, name) for name, _ in _StepResult._fields_}


def _scal(s):
    out = {}
    for name, _ in _Scalars._fields_:
        v = getattr(s, name)
        out[name] = list(v) if hasattr(v, "__len__") else v
    return out


class Simulator:
    """One foam on the GPU.  Mirrors the reference's step interface:
    step() (one attempt), integrate(), search_neighbors(), remove_bubbles()."""

    def __init__(self, config=None, json_path=None, tuning=None):
        h = ctypes.c_void_p()
        if json_path is not None:
            _check(lib().cubble_create_from_json(json_path.encode(), ctypes.byref(h)))
        else:
            cfg = config if config is not None else Config()
            self.config = cfg
            _check(lib().cubble_create_tuned(ctypes.byref(cfg._c()), _tuning(tuning),
                                             ctypes.byref(h)))
        self._h = h

    def close(self):
        if getattr(self, "_h", None) and self._h.value:
            lib().cubble_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def _call(self, rc):
        _check(rc, self._h)

    def step(self, ip=None):
        r = _StepResult()
        self._call(lib().cubble_step(self._h, ctypes.byref((ip or IntegrationParams())._c()),
                                     ctypes.byref(r)))
        return _res(r)

    def integrate(self, ip=None):
        r = _StepResult()
        self._call(lib().cubble_integrate(self._h, ctypes.byref((ip or IntegrationParams())._c()),
                                          ctypes.byref(r)))
        return _res(r)

    def run_steps(self, n, ip=None):
        r = _StepResult()
        self._call(lib().cubble_run_steps(self._h, ctypes.byref((ip or IntegrationParams())._c()),
                                          int(n), ctypes.byref(r)))
        return _res(r)

    def autotune(self, ip=None, trial_steps=64, repeats=2, cache_path=None, ignore_cache=False):
        """Pick launch shapes by timing them on this foam; the state is
        restored afterwards.  Returns the chosen shapes and the rates."""
        return _autotune(lib().cubble_autotune, self._h, ip, trial_steps, repeats, cache_path,
                         ignore_cache, self._call)

    def retune(self, tuning):
        self._call(lib().cubble_retune(self._h, _tuning(tuning)))

    def time_kernels(self, reps=20, ip=None):
        """Average device ms of predictor / pair / corrector, from HIP events."""
        t = _KernelTimes()
        self._call(lib().cubble_time_kernels(self._h, ctypes.byref((ip or IntegrationParams())._c()),
                                             int(reps), ctypes.byref(t)))
        return {"predict_ms": t.predict_ms, "pair_ms": t.pair_ms, "correct_ms": t.correct_ms,
                "reps": t.reps}

    def time_search(self, reps=10):
        """Average device ms of a search of the foam and of its parts, and of an
        attempt group's idle removal and search kernels (the state is put back)."""
        t = _SearchTimes()
        self._call(lib().cubble_time_search(self._h, int(reps), ctypes.byref(t)))
        return _search_times(t)

    def observe(self, energy=False):
        """Total energy and volume, box volume and the means of r, path and
        the velocities, reduced on the device; with energy=True also each
        bubble's energy (state order)."""
        o = _Observables()
        e = np.zeros(self.scalars()["count"]) if energy else None
        self._call(lib().cubble_observe(self._h, ctypes.byref(o),
                                        _dptr(e) if e is not None and len(e) else None))
        out = {name: (list(getattr(o, name)) if name == "mean_v" else getattr(o, name))
               for name, _ in _Observables._fields_}
        if energy:
            out["energy"] = e
        return out

    def set_stop_rules(self, min_count=0, max_radius=0.0, time=None):
        """Device-side stop rules checked after every accepted step; time is
        (integer part, fraction) or None."""
        r = _StopRules(int(min_count), 0 if time is None else 1, float(max_radius),
                       0 if time is None else int(time[0]), 0.0 if time is None else float(time[1]))
        self._call(lib().cubble_set_stop_rules(self._h, ctypes.byref(r)))

    def reset_print_window(self):
        """Zero the per-bubble error and restart the time-step window."""
        self._call(lib().cubble_reset_print_window(self._h))

    def rescale_box(self, phi, rel_dim=(1.0, 1.0, 1.0)):
        """The box rescale of the reference's init (then a search)."""
        rel = np.ascontiguousarray(rel_dim, dtype=np.float64)
        self._call(lib().cubble_rescale_box(self._h, float(phi), _dptr(rel)))

    def bootstrap_velocities(self):
        """The AB2 velocity bootstrap of the reference's init (new foam only)."""
        self._call(lib().cubble_bootstrap_velocities(self._h))

    def reset_clock(self, clear_wraps=False):
        """Time to zero; with clear_wraps also the wrap counters."""
        self._call(lib().cubble_reset_clock(self._h, int(bool(clear_wraps))))

    def search_neighbors(self):
        self._call(lib().cubble_search_neighbors(self._h))

    def remove_bubbles(self):
        self._call(lib().cubble_remove_bubbles(self._h))

    def scalars(self):
        s = _Scalars()
        self._call(lib().cubble_get_scalars(self._h, ctypes.byref(s)))
        return _scal(s)

    def get_state(self):
        sc = self.scalars()
        return _read_state(sc, lambda v: self._call(lib().cubble_get_state(self._h, v))), sc

    def get_attempt(self):
        n = self.scalars()["count"]
        o = {"xc": np.zeros((4, n)), "vp": np.zeros((4, n)), "ov": np.zeros(n),
             "path_new": np.zeros(n), "del": np.zeros(n, np.uint8)}
        a = _AttemptView()
        for k in range(4):
            a.xc[k], a.vp[k] = _dptr(o["xc"][k]), _dptr(o["vp"][k])
        a.overlap, a.path_new = _dptr(o["ov"]), _dptr(o["path_new"])
        a.deleted = o["del"].ctypes.data_as(_u8p)
        self._call(lib().cubble_get_attempt(self._h, ctypes.byref(a)))
        return o

    def set_state(self, st, scalars):
        n = int(scalars["count"])
        keep = []

        def d(a):
            a = np.ascontiguousarray(a, dtype=np.float64)
            keep.append(a)
            return a

        def i32(a):
            a = np.ascontiguousarray(a, dtype=np.int32)
            keep.append(a)
            return a

        v = _StateView()
        x, vv, vo, sv = d(st["x"]), d(st["v"]), d(st["vo"]), d(st["saved"])
        for k in range(4):
            v.x[k], v.v[k], v.vo[k], v.saved[k] = _dptr(x[k]), _dptr(vv[k]), _dptr(vo[k]), _dptr(sv[k])
        wc = i32(st["wc"])
        for k in range(3):
            v.wrap[k] = _iptr(wc[k])
        path, err, idx = d(st["path"]), d(st["err"]), i32(st["idx"])
        v.path, v.error, v.index = _dptr(path), _dptr(err), _iptr(idx)
        off = i32(st["nbr_off"])
        nbr = i32(st["nbr"] if len(st["nbr"]) else [0])
        v.nbr_off, v.nbr, v.nbr_cap = _iptr(off), _iptr(nbr), len(st["nbr"])
        if st.get("deg") is not None:
            v.degree = _iptr(i32(st["deg"]))
        s = _Scalars()
        for name, _ in _Scalars._fields_:
            if name in scalars:
                val = scalars[name]
                if hasattr(getattr(s, name), "__len__"):
                    arr = getattr(s, name)
                    for k in range(3):
                        arr[k] = val[k]
                else:
                    setattr(s, name, val)
        s.count = n
        self._call(lib().cubble_set_state(self._h, ctypes.byref(v), ctypes.byref(s)))

    def pairs(self):
        """Half-list pair set (i < j), the reference's `pairs` (src/kernels.cu:703-711)."""
        st, _ = self.get_state()
        return half_pairs(st["nbr_off"], st["nbr"])


def _read_state(sc, fill):
    n = sc["count"]
    st = {"x": np.zeros((4, n)), "v": np.zeros((4, n)), "vo": np.zeros((4, n)),
          "saved": np.zeros((4, n)), "path": np.zeros(n), "err": np.zeros(n),
          "wc": np.zeros((3, n), np.int32), "idx": np.zeros(n, np.int32),
          "nbr_off": np.zeros(n + 1, np.int32), "deg": np.zeros(n, np.int32)}
    nbr = np.zeros(max(1, sc["nbr_cap"] * n), np.int32)
    v = _StateView()
    for k in range(4):
        v.x[k], v.v[k] = _dptr(st["x"][k]), _dptr(st["v"][k])
        v.vo[k], v.saved[k] = _dptr(st["vo"][k]), _dptr(st["saved"][k])
    for k in range(3):
        v.wrap[k] = _iptr(st["wc"][k])
    v.path, v.error, v.index = _dptr(st["path"]), _dptr(st["err"]), _iptr(st["idx"])
    v.nbr_off, v.nbr, v.nbr_cap = _iptr(st["nbr_off"]), _iptr(nbr), nbr.size
    v.degree = _iptr(st["deg"])
    fill(ctypes.byref(v))
    st["nbr"] = nbr[: st["nbr_off"][n]].copy()
    return st


class Batch:
    """Many independent foams advanced together (2D and 3D mixed, uneven
    sizes, per-foam constants); every kernel launch covers the whole batch and
    a run of steps is a replayed HIP graph."""

    def __init__(self, configs, tuning=None):
        configs = list(configs)
        arr = (_Config * len(configs))(*[c._c() for c in configs])
        h = ctypes.c_void_p()
        rc = lib().cubble_batch_create(arr, len(configs), _tuning(tuning), ctypes.byref(h))
        if rc != OK:
            raise CubbleError(rc, lib().cubble_batch_last_error(None).decode())
        self._h = h
        self.configs = configs

    def close(self):
        if getattr(self, "_h", None) and self._h.value:
            lib().cubble_batch_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _call(self, rc):
        if rc != OK:
            raise CubbleError(rc, lib().cubble_batch_last_error(self._h).decode())

    def __len__(self):
        return lib().cubble_batch_size(self._h)

    def integrate(self, ip=None):
        r = (_StepResult * len(self))()
        self._call(lib().cubble_batch_integrate(self._h, ctypes.byref((ip or IntegrationParams())._c()), r))
        return [_res(x) for x in r]

    def run_steps(self, n, ip=None):
        r = (_StepResult * len(self))()
        self._call(lib().cubble_batch_run_steps(self._h, ctypes.byref((ip or IntegrationParams())._c()),
                                                int(n), r))
        return [_res(x) for x in r]

    def scalars(self, k):
        s = _Scalars()
        self._call(lib().cubble_batch_get_scalars(self._h, int(k), ctypes.byref(s)))
        return _scal(s)

    def get_state(self, k):
        sc = self.scalars(k)
        return _read_state(sc, lambda v: self._call(lib().cubble_batch_get_state(self._h, int(k), v))), sc

    def autotune(self, ip=None, trial_steps=64, repeats=2, cache_path=None, ignore_cache=False):
        """Pick launch shapes by timing them on these foams; the state is
        restored afterwards.  Returns the chosen shapes and the rates."""
        return _autotune(lib().cubble_batch_autotune, self._h, ip, trial_steps, repeats,
                         cache_path, ignore_cache, self._call)

    def retune(self, tuning):
        self._call(lib().cubble_batch_retune(self._h, _tuning(tuning)))

    def time_kernels(self, reps=20, ip=None):
        t = _KernelTimes()
        self._call(lib().cubble_batch_time_kernels(self._h, ctypes.byref((ip or IntegrationParams())._c()),
                                                   int(reps), ctypes.byref(t)))
        return {"predict_ms": t.predict_ms, "pair_ms": t.pair_ms, "correct_ms": t.correct_ms,
                "reps": t.reps}

    def time_search(self, reps=10):
        t = _SearchTimes()
        self._call(lib().cubble_batch_time_search(self._h, int(reps), ctypes.byref(t)))
        return _search_times(t)


class Server:
    """Request loop with continuous batching: requests are admitted in arrival
    order when their device footprint fits `hbm_budget_bytes`, stay resident on
    the GPU and advance together in rounds of `round_steps` steps."""

    def __init__(self, hbm_budget_bytes, max_active=0, round_steps=50, keep_states=False,
                 precision=PREC_FP64):
        o = _ServerOpts(int(hbm_budget_bytes), int(max_active), int(round_steps),
                        int(bool(keep_states)), int(precision))
        h = ctypes.c_void_p()
        rc = lib().cubble_server_create(ctypes.byref(o), ctypes.byref(h))
        if rc != OK:
            raise CubbleError(rc, "cubble_server_create failed")
        self._h = h

    def close(self):
        if getattr(self, "_h", None) and self._h.value:
            lib().cubble_server_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _call(self, rc):
        if rc != OK:
            raise CubbleError(rc, lib().cubble_server_last_error(self._h).decode())

    def submit(self, config, max_steps, dissolve_fraction=0.0, arrival_s=0.0):
        rid = ctypes.c_int64()
        self._call(lib().cubble_server_submit(self._h, ctypes.byref(config._c()), int(max_steps),
                                              float(dissolve_fraction), float(arrival_s),
                                              ctypes.byref(rid)))
        return rid.value

    def run(self):
        st = _ServerStats()
        self._call(lib().cubble_server_run(self._h, ctypes.byref(st)))
        return {name: getattr(st, name) for name, _ in _ServerStats._fields_ if name != "pad"}

    def result(self, rid):
        r = _RequestResult()
        self._call(lib().cubble_server_result(self._h, int(rid), ctypes.byref(r)))
        return {name: getattr(r, name) for name, _ in _RequestResult._fields_ if name != "pad"}

    def get_state(self, rid):
        s = _Scalars()
        self._call(lib().cubble_server_get_state(self._h, int(rid), None, ctypes.byref(s)))
        sc = _scal(s)
        return _read_state(sc, lambda v: self._call(
            lib().cubble_server_get_state(self._h, int(rid), v, None))), sc


def foam_footprint(config):
    """Device bytes one foam of this configuration occupies (admission size)."""
    return lib().cubble_foam_footprint(ctypes.byref(config._c()))


def half_pairs(off, nbr):
    off = np.asarray(off)
    nbr = np.asarray(nbr)
    n = len(off) - 1
    i = np.repeat(np.arange(n, dtype=np.int32), np.diff(off))
    m = nbr > i
    p = np.stack([i[m], nbr[m]], axis=1)
    order = np.lexsort((p[:, 1], p[:, 0]))
    return p[order]


def find_pairs(x, y, z, r, count, box, skin, pair_i, pair_j, pair_cap, num_pairs, stream=0,
               noalloc=False):
    """cubble_find_pairs (noalloc: cubble_find_pairs_noalloc) on device
    addresses (ints): positions / radii x, y, z, r (FP64), pair_i / pair_j
    (int32, pair_cap each), num_pairs (one int64); stream a hipStream_t
    address (0 = the null stream).  Enqueues only."""
    fn = lib().cubble_find_pairs_noalloc if noalloc else lib().cubble_find_pairs
    rc = fn(x, y, z, r, int(count), ctypes.byref(box), float(skin), pair_i, pair_j, int(pair_cap),
            num_pairs, stream or None)
    if rc != OK:
        raise CubbleError(rc, lib().cubble_last_error(None).decode())


def generate_foam(config):
    """Host-only: the initial arrays create() uploads (x, y, z, r rows) and the derived
    constants.  Needs no GPU."""
    n = config.num_start
    x4 = np.zeros(4 * n)
    s = _Scalars()
    _check(lib().cubble_generate_foam(ctypes.byref(config._c()), _dptr(x4), ctypes.byref(s)))
    return x4.reshape(4, n), _scal(s)


def version():
    return lib().cubble_version().decode()


def run_cli(argv):
    """cubble_run(argc, argv) -- the drop-in command line entry."""
    arr = (ctypes.c_char_p * len(argv))(*[a.encode() for a in argv])
    return lib().cubble_run(len(argv), arr)
