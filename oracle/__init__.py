"""ctypes wrapper of the CPU FP64 oracle.

TEST INFRASTRUCTURE ONLY: a checker for the HIP product library and the CPU
baseline timer.  Only tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg import this module; the product package never does.

State dictionaries use the same keys as ``cubble_amd``'s state views:
``x`` (4, n) holds x, y, z, r; ``v``/``vo``/``saved`` (4, n) the matching
velocities, old velocities and positions at the last search; ``path``,
``err`` (n,); ``wc`` (3, n) and ``idx`` (n,) int32; ``nbr_off`` (n+1,) and
``nbr`` the CSR full neighbour lists; ``deg`` (n,) the list lengths at the
last search (the divisor of the imposed flow's neighbour mean).
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libcubble_oracle.so")

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int32)
_u8p = ctypes.POINTER(ctypes.c_uint8)


class Consts(ctypes.Structure):
    _fields_ = [
        ("dim", ctypes.c_int32),
        ("wall", ctypes.c_int32 * 3),
        ("lbb", ctypes.c_double * 3),
        ("tfr", ctypes.c_double * 3),
        ("interval", ctypes.c_double * 3),
        ("wall_drag", ctypes.c_double),
        ("skin", ctypes.c_double),
        ("min_rad", ctypes.c_double),
        ("f0", ctypes.c_double),
        ("k_param", ctypes.c_double),
        ("kappa", ctypes.c_double),
        ("avg_area_in", ctypes.c_double),
        ("vol_mul", ctypes.c_double),
        ("tol", ctypes.c_double),
        ("cell_cap", ctypes.c_int32),
        ("flow_lbb", ctypes.c_double * 3),
        ("flow_tfr", ctypes.c_double * 3),
        ("flow_vel", ctypes.c_double * 3),
    ]


class Scalars(ctypes.Structure):
    _fields_ = [
        ("count", ctypes.c_int32),
        ("cell_dim", ctypes.c_int32 * 3),
        ("dt", ctypes.c_double),
        ("time_int", ctypes.c_uint64),
        ("time_frac", ctypes.c_double),
        ("max_r_cell", ctypes.c_double),
        ("steps", ctypes.c_int64),
        ("attempts", ctypes.c_int64),
        ("searches", ctypes.c_int64),
        ("removals", ctypes.c_int64),
        ("bubble_steps", ctypes.c_double),
        ("s_a", ctypes.c_double),
        ("s_ar", ctypes.c_double),
        ("s_oa", ctypes.c_double),
        ("s_oar", ctypes.c_double),
        ("s_v", ctypes.c_double),
        ("max_err", ctypes.c_double),
        ("max_r", ctypes.c_double),
        ("max_exp", ctypes.c_double),
        ("n_del", ctypes.c_int32),
        ("accepted", ctypes.c_int32),
        ("need_search", ctypes.c_int32),
        ("num_pairs", ctypes.c_int32),
    ]


class State(ctypes.Structure):
    _fields_ = [
        ("x", _dp * 4),
        ("v", _dp * 4),
        ("vo", _dp * 4),
        ("saved", _dp * 4),
        ("path", _dp),
        ("err", _dp),
        ("wc", _ip * 3),
        ("idx", _ip),
        ("nbr_off", _ip),
        ("nbr", _ip),
        ("nbr_cap", ctypes.c_int64),
        ("deg", _ip),
    ]


class AttemptOut(ctypes.Structure):
    _fields_ = [
        ("xc", _dp * 4),
        ("vp", _dp * 4),
        ("ov", _dp),
        ("path_new", _dp),
        ("dele", _u8p),
    ]


class InitParams(ctypes.Structure):
    _fields_ = [
        ("dim", ctypes.c_int32),
        ("count", ctypes.c_int32),
        ("radius_dist", ctypes.c_int32),
        ("r_mean", ctypes.c_double),
        ("r_std", ctypes.c_double),
        ("r_lo", ctypes.c_double),
        ("r_hi", ctypes.c_double),
        ("phi", ctypes.c_double),
        ("rel_dim", ctypes.c_double * 3),
        ("wall", ctypes.c_int32 * 3),
        ("wall_drag", ctypes.c_double),
        ("avg_rad", ctypes.c_double),
        ("mu", ctypes.c_double),
        ("sigma", ctypes.c_double),
        ("kappa", ctypes.c_double),
        ("k_param", ctypes.c_double),
        ("tol", ctypes.c_double),
        ("dt0", ctypes.c_double),
        ("cell_cap_factor", ctypes.c_double),
        ("seed", ctypes.c_uint64),
    ]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("oracle library not built: run `make -C oracle`")
        L = ctypes.CDLL(LIB_PATH)
        vp = ctypes.c_void_p
        L.ora_create.restype = vp
        L.ora_create.argtypes = [ctypes.POINTER(Consts), ctypes.c_int32]
        L.ora_destroy.argtypes = [vp]
        L.ora_get_consts.argtypes = [vp, ctypes.POINTER(Consts)]
        L.ora_set_state.argtypes = [vp, ctypes.POINTER(State), ctypes.POINTER(Scalars)]
        L.ora_get_state.argtypes = [vp, ctypes.POINTER(State), ctypes.POINTER(Scalars)]
        L.ora_get_state.restype = ctypes.c_int
        L.ora_get_attempt.argtypes = [vp, ctypes.POINTER(AttemptOut)]
        L.ora_attempt.argtypes = [vp, ctypes.c_int, ctypes.c_int, ctypes.c_int]
        L.ora_attempt.restype = ctypes.c_int
        L.ora_integrate.argtypes = [vp, ctypes.c_int, ctypes.c_int, ctypes.c_int]
        L.ora_integrate.restype = ctypes.c_int
        L.ora_run.argtypes = [vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int64]
        L.ora_run.restype = ctypes.c_double
        L.ora_search.argtypes = [vp]
        L.ora_remove.argtypes = [vp]
        L.ora_brute_pairs.argtypes = [vp, _ip, ctypes.c_int64]
        L.ora_brute_pairs.restype = ctypes.c_int64
        L.ora_init_synthetic.argtypes = [ctypes.POINTER(InitParams)]
        L.ora_init_synthetic.restype = vp
        L.ora_observe.argtypes = [vp, _dp, _dp]
        L.ora_bootstrap.argtypes = [vp]
        L.ora_rescale.argtypes = [vp, ctypes.c_double, _dp, ctypes.c_double]
        L.ora_wall_velocity.argtypes = [ctypes.POINTER(Consts), _dp, ctypes.c_double, _dp]
        L.ora_cbrt.argtypes = [ctypes.c_double]
        L.ora_cbrt.restype = ctypes.c_double
        u64p = ctypes.POINTER(ctypes.c_uint64)
        L.ora_splitmix64.argtypes = [ctypes.c_uint64, ctypes.c_int32, u64p]
        L.ora_xoshiro256ss.argtypes = [u64p, ctypes.c_int32, u64p]
        L.ora_set_threads.argtypes = [ctypes.c_int]
        L.ora_get_threads.restype = ctypes.c_int
        _lib = L
    return _lib


def _dptr(a):
    return a.ctypes.data_as(_dp)


def _iptr(a):
    return a.ctypes.data_as(_ip)


def scalars_to_dict(sc):
    out = {}
    for name, _ in Scalars._fields_:
        v = getattr(sc, name)
        out[name] = list(v) if name == "cell_dim" else v
    return out


def consts_to_dict(c):
    out = {}
    for name, _ in Consts._fields_:
        v = getattr(c, name)
        out[name] = list(v) if hasattr(v, "__len__") else v
    return out


def consts_from_dict(d):
    """Consts from a dict; the flow keys default to zero (no flow box)."""
    c = Consts()
    for name, _ in Consts._fields_:
        v = d.get(name, (0.0, 0.0, 0.0)) if name.startswith("flow_") else d[name]
        if hasattr(v, "__len__"):
            arr = getattr(c, name)
            for k in range(3):
                arr[k] = v[k]
        else:
            setattr(c, name, v)
    return c


def init_params(dim, count, radius_dist=0, r_mean=1.0, r_std=0.1, r_lo=0.5, r_hi=1.5,
                phi=0.9, kappa=0.01, k_param=0.1, mu=1.0, sigma=1.0, tol=1e-4,
                dt0=1e-4, avg_rad=1.0, seed=426, rel_dim=(1.0, 1.0, 1.0),
                wall=(0, 0, 0), wall_drag=0.1, cell_cap_factor=2.0):
    p = InitParams()
    p.dim, p.count, p.radius_dist = dim, count, radius_dist
    p.r_mean, p.r_std, p.r_lo, p.r_hi = r_mean, r_std, r_lo, r_hi
    p.phi = phi
    for k in range(3):
        p.rel_dim[k] = rel_dim[k]
        p.wall[k] = wall[k]
    p.wall_drag = wall_drag
    p.avg_rad, p.mu, p.sigma = avg_rad, mu, sigma
    p.kappa, p.k_param, p.tol, p.dt0 = kappa, k_param, tol, dt0
    p.cell_cap_factor = cell_cap_factor
    p.seed = seed
    return p


class Foam:
    """One oracle foam (owns the C object)."""

    def __init__(self, handle):
        self.h = handle

    @classmethod
    def synthetic(cls, **kw):
        p = init_params(**kw)
        return cls(lib().ora_init_synthetic(ctypes.byref(p)))

    @classmethod
    def from_state(cls, consts, state, scalars):
        c = consts if isinstance(consts, Consts) else consts_from_dict(consts)
        f = cls(lib().ora_create(ctypes.byref(c), int(scalars["count"])))
        f.set_state(state, scalars)
        return f

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.ora_destroy(self.h)
            self.h = None

    def consts(self):
        c = Consts()
        lib().ora_get_consts(self.h, ctypes.byref(c))
        return consts_to_dict(c)

    def scalars(self):
        sc = Scalars()
        lib().ora_get_state(self.h, None, ctypes.byref(sc))
        return scalars_to_dict(sc)

    def get_state(self):
        sc = self.scalars()
        n = sc["count"]
        st = {
            "x": np.zeros((4, n)), "v": np.zeros((4, n)), "vo": np.zeros((4, n)),
            "saved": np.zeros((4, n)), "path": np.zeros(n), "err": np.zeros(n),
            "wc": np.zeros((3, n), np.int32), "idx": np.zeros(n, np.int32),
            "nbr_off": np.zeros(n + 1, np.int32), "deg": np.zeros(n, np.int32),
        }
        nbr = np.zeros(max(1, 2 * sc["num_pairs"]), np.int32)
        s = State()
        for k in range(4):
            s.x[k] = _dptr(st["x"][k])
            s.v[k] = _dptr(st["v"][k])
            s.vo[k] = _dptr(st["vo"][k])
            s.saved[k] = _dptr(st["saved"][k])
        for k in range(3):
            s.wc[k] = _iptr(st["wc"][k])
        s.path, s.err = _dptr(st["path"]), _dptr(st["err"])
        s.idx, s.nbr_off, s.nbr = _iptr(st["idx"]), _iptr(st["nbr_off"]), _iptr(nbr)
        s.nbr_cap = nbr.size
        s.deg = _iptr(st["deg"])
        rc = lib().ora_get_state(self.h, ctypes.byref(s), None)
        if rc != 0:
            raise RuntimeError("oracle neighbour list larger than expected")
        st["nbr"] = nbr[: st["nbr_off"][n]].copy()
        return st, sc

    def observe(self):
        """(sums, energy): sums = total energy, total volume, then the sums of
        r, path, dxdt, dydt, dzdt, drdt; energy = per-bubble energies."""
        out = np.zeros(8)
        e = np.zeros(max(1, self.scalars()["count"]))
        lib().ora_observe(self.h, _dptr(out), _dptr(e))
        return out, e[: self.scalars()["count"]]

    def set_state(self, st, scalars):
        n = int(scalars["count"])
        keep = {}

        def d(a, shape=None):
            a = np.ascontiguousarray(a, dtype=np.float64)
            keep[id(a)] = a
            return a

        def i(a):
            a = np.ascontiguousarray(a, dtype=np.int32)
            keep[id(a)] = a
            return a

        s = State()
        x, v, vo, sv = d(st["x"]), d(st["v"]), d(st["vo"]), d(st["saved"])
        for k in range(4):
            s.x[k] = _dptr(x[k])
            s.v[k] = _dptr(v[k])
            s.vo[k] = _dptr(vo[k])
            s.saved[k] = _dptr(sv[k])
        wc = i(st["wc"])
        for k in range(3):
            s.wc[k] = _iptr(wc[k])
        path, err = d(st["path"]), d(st["err"])
        s.path, s.err = _dptr(path), _dptr(err)
        idx, off, nbr = i(st["idx"]), i(st["nbr_off"]), i(st["nbr"] if len(st["nbr"]) else [0])
        s.idx, s.nbr_off, s.nbr = _iptr(idx), _iptr(off), _iptr(nbr)
        s.nbr_cap = len(st["nbr"])
        if st.get("deg") is not None:
            s.deg = _iptr(i(st["deg"]))
        sc = Scalars()
        for name, _ in Scalars._fields_:
            if name == "cell_dim":
                for k in range(3):
                    sc.cell_dim[k] = int(scalars["cell_dim"][k])
            else:
                setattr(sc, name, scalars[name])
        sc.count = n
        lib().ora_set_state(self.h, ctypes.byref(s), ctypes.byref(sc))

    def attempt_out(self):
        n = self.scalars()["count"]
        o = {"xc": np.zeros((4, n)), "vp": np.zeros((4, n)), "ov": np.zeros(n),
             "path_new": np.zeros(n), "del": np.zeros(n, np.uint8)}
        a = AttemptOut()
        for k in range(4):
            a.xc[k] = _dptr(o["xc"][k])
            a.vp[k] = _dptr(o["vp"][k])
        a.ov, a.path_new = _dptr(o["ov"]), _dptr(o["path_new"])
        a.dele = o["del"].ctypes.data_as(_u8p)
        lib().ora_get_attempt(self.h, ctypes.byref(a))
        return o

    def attempt(self, gas=True, path=True, flow=False):
        return lib().ora_attempt(self.h, int(gas), int(path), int(flow))

    def integrate(self, gas=True, path=True, flow=False):
        return lib().ora_integrate(self.h, int(gas), int(path), int(flow))

    def run(self, steps, gas=True, path=True, flow=False):
        return lib().ora_run(self.h, int(gas), int(path), int(flow), int(steps))

    def bootstrap(self):
        lib().ora_bootstrap(self.h)

    def rescale(self, phi, rel_dim, volume):
        rel = np.ascontiguousarray(rel_dim, dtype=np.float64)
        lib().ora_rescale(self.h, float(phi), _dptr(rel), float(volume))

    def search(self):
        lib().ora_search(self.h)

    def remove(self):
        lib().ora_remove(self.h)

    def brute_pairs(self):
        n = self.scalars()["count"]
        cap = 64 * max(n, 1)
        out = np.zeros(2 * cap, np.int32)
        np_ = lib().ora_brute_pairs(self.h, _iptr(out), cap)
        if np_ > cap:
            raise RuntimeError("too many pairs")
        return out[: 2 * np_].reshape(-1, 2)

    def pairs(self):
        """Half-list pair set (i < j) from the full neighbour lists."""
        st, sc = self.get_state()
        return half_pairs(st["nbr_off"], st["nbr"])


def half_pairs(off, nbr):
    off = np.asarray(off)
    nbr = np.asarray(nbr)
    n = len(off) - 1
    i = np.repeat(np.arange(n, dtype=np.int32), np.diff(off))
    m = nbr > i
    p = np.stack([i[m], nbr[m]], axis=1)
    order = np.lexsort((p[:, 1], p[:, 0]))
    return p[order]


def cbrt(a):
    return lib().ora_cbrt(float(a))


def wall_velocity(consts, pos, rad, v):
    """The oracle's wall term alone: velocity v (3) of a bubble at pos (3)
    with radius rad, after the walls of `consts` (a dict)."""
    c = consts_from_dict(consts)
    p = np.array(list(pos) + [0.0] * (3 - len(pos)), float)
    out = np.array(list(v) + [0.0] * (3 - len(v)), float)
    lib().ora_wall_velocity(ctypes.byref(c), _dptr(p), float(rad), _dptr(out))
    return out


def splitmix64(seed, n):
    out = np.zeros(n, np.uint64)
    lib().ora_splitmix64(seed, n, out.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)))
    return out


def xoshiro256ss(state, n):
    st = np.asarray(state, np.uint64)
    out = np.zeros(n, np.uint64)
    u64p = ctypes.POINTER(ctypes.c_uint64)
    lib().ora_xoshiro256ss(st.ctypes.data_as(u64p), n, out.ctypes.data_as(u64p))
    return out


def set_threads(n):
    lib().ora_set_threads(int(n))


def max_threads():
    return lib().ora_get_threads()
