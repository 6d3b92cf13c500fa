"""Pins the CPU FP64 oracle against closed forms, brute force, invariants
and published known-answer vectors.

The reference (KJLankinen/cubble) ships no tests, golden vectors or
fixtures for this path, so every check here comes from the mathematics of
Durian's bubble model or from a textbook routine:

* RNG: SplitMix64 and xoshiro256** published test vectors;
* pair search: O(N^2) brute force, permutation / translation invariance;
* pair velocity: f0 (1/|d| - 1/R) d for two bubbles;
* overlap area: lens-circle radius computed from the *other* sphere;
* liquid-mediated gas flux: two-bubble closed form;
* predictor / corrector: AB2 / trapezoid on a free bubble;
* dt controller: scripted error (tolerance) sequences;
* volume fix: sum V(r) = phi * V_box afterwards;
* compaction: survivors and their pair set versus an explicit filter.
"""
import math

import numpy as np
import pytest

PI = 3.1415926535897932384626433832795028841971693993


def consts(dim, L=20.0, wall=(0, 0, 0), tol=1e-4, f0=1.0, kappa=0.01, k=0.1, ain=1.0,
           vol_mul=1.0, skin=0.3, min_rad=0.1, drag=0.1, cell_cap=1 << 20,
           flow_lbb=(0.0, 0.0, 0.0), flow_tfr=(0.0, 0.0, 0.0), flow_vel=(0.0, 0.0, 0.0)):
    iv = [L, L, L if dim == 3 else 0.0]
    return {
        "dim": dim, "wall": list(wall), "lbb": [0.0, 0.0, 0.0], "tfr": iv, "interval": iv,
        "wall_drag": drag, "skin": skin, "min_rad": min_rad, "f0": f0, "k_param": k,
        "kappa": kappa, "avg_area_in": ain, "vol_mul": vol_mul, "tol": tol, "cell_cap": cell_cap,
        "flow_lbb": list(flow_lbb), "flow_tfr": list(flow_tfr), "flow_vel": list(flow_vel),
    }


def make_state(pos, r, dim, v=None, vo=None):
    n = len(r)
    x = np.zeros((4, n))
    pos = np.asarray(pos, float)
    x[:dim] = pos.T[:dim]
    x[3] = r
    st = {
        "x": x, "v": np.zeros((4, n)) if v is None else np.asarray(v, float),
        "vo": np.zeros((4, n)) if vo is None else np.asarray(vo, float),
        "saved": x.copy(), "path": np.zeros(n), "err": np.zeros(n),
        "wc": np.zeros((3, n), np.int32), "idx": np.arange(n, dtype=np.int32),
        "nbr_off": np.zeros(n + 1, np.int32), "nbr": np.zeros(0, np.int32),
    }
    return st


def scalars(n, dt=1e-4, max_r=1.0):
    return {"count": n, "cell_dim": [1, 1, 1], "dt": dt, "time_int": 0, "time_frac": 0.0,
            "max_r_cell": max_r, "steps": 0, "attempts": 0, "searches": 0, "removals": 0,
            "bubble_steps": 0.0, "s_a": 0.0, "s_ar": 0.0, "s_oa": 0.0, "s_oar": 0.0,
            "s_v": 0.0, "max_err": 0.0, "max_r": 0.0, "max_exp": 0.0, "n_del": 0,
            "accepted": 0, "need_search": 0, "num_pairs": 0}


def foam(ora, pos, r, dim, **kw):
    c = consts(dim, **{k: v for k, v in kw.items() if k not in ("dt", "v", "vo")})
    st = make_state(pos, r, dim, kw.get("v"), kw.get("vo"))
    f = ora.Foam.from_state(c, st, scalars(len(r), dt=kw.get("dt", 1e-4), max_r=max(r)))
    f.search()
    return f


# --------------------------------------------------------------------------
# RNG known answers
# --------------------------------------------------------------------------
def test_splitmix64_published_vectors(ora):
    got = [int(v) for v in ora.splitmix64(0, 3)]
    assert got == [0xE220A8397B1DCDAF, 0x6E789E6AA1B965F4, 0x06C45D188009454F]


def test_xoshiro256ss_published_vectors(ora):
    got = [int(v) for v in ora.xoshiro256ss([1, 2, 3, 4], 4)]
    assert got == [11520, 0, 1509978240, 1215971899390074240]


# --------------------------------------------------------------------------
# deterministic cube root
# --------------------------------------------------------------------------
@pytest.mark.parametrize("a", [1e-300, 3.7e-9, 0.5, 1.0, 2.0, 8.0, 27.0, 1234.5678, 6.02e23, 1e300])
def test_cbrt_close_to_libm(ora, a):
    got = ora.cbrt(a)
    want = np.cbrt(a)
    assert abs(got - want) <= 2 * np.spacing(want)


def test_cbrt_exact_cubes(ora):
    for k in range(1, 200):
        assert ora.cbrt(float(k ** 3)) == float(k)


# --------------------------------------------------------------------------
# pair search
# --------------------------------------------------------------------------
@pytest.mark.parametrize("dim,n,phi,seed", [(2, 500, 0.9, 1), (2, 3000, 0.8, 2), (3, 800, 0.84, 3),
                                            (3, 4000, 0.95, 4)])
def test_cell_search_equals_brute_force(ora, dim, n, phi, seed):
    f = ora.Foam.synthetic(dim=dim, count=n, phi=phi, seed=seed, radius_dist=dim - 2, r_std=0.15)
    assert np.array_equal(f.pairs(), f.brute_pairs())
    st, sc = f.get_state()
    # full lists are symmetric and ascending
    off, nbr = st["nbr_off"], st["nbr"]
    for i in range(0, n, max(1, n // 50)):
        lst = nbr[off[i]:off[i + 1]]
        assert np.all(np.diff(lst) > 0)
        for j in lst:
            assert i in nbr[off[j]:off[j + 1]]


def test_cell_order_is_stable_counting_sort(ora):
    f = ora.Foam.synthetic(dim=3, count=2000, phi=0.84, seed=7, radius_dist=1, r_std=0.15)
    st, sc = f.get_state()
    c = f.consts()
    cd = sc["cell_dim"]
    ids = []
    for i in range(sc["count"]):
        q = [min(max(int(math.floor(cd[k] * (st["x"][k][i] - c["lbb"][k]) / c["interval"][k])), 0),
                 cd[k] - 1) for k in range(3)]
        ids.append(q[2] * cd[0] * cd[1] + q[1] * cd[0] + q[0])
    ids = np.array(ids)
    assert np.all(np.diff(ids) >= 0)
    # within a cell the original indices ascend (first search from identity order)
    for cell in np.unique(ids)[:200]:
        sel = st["idx"][ids == cell]
        assert np.all(np.diff(sel) > 0)


@pytest.mark.parametrize("dim", [2, 3])
def test_pair_set_invariant_to_permutation_and_translation(ora, dim):
    rng = np.random.default_rng(11)
    n = 600
    L = 30.0 if dim == 3 else 60.0
    pos = rng.uniform(0, L, size=(n, 3))
    r = rng.uniform(0.7, 1.3, size=n)

    def pair_set(pos, r, ids):
        f = foam(ora, pos, r, dim, L=L)
        st, _ = f.get_state()
        p = ora.half_pairs(st["nbr_off"], st["nbr"])
        orig = st["idx"]
        return {tuple(sorted((ids[orig[a]], ids[orig[b]]))) for a, b in p}

    base = pair_set(pos, r, np.arange(n))
    perm = rng.permutation(n)
    permuted = pair_set(pos[perm], r[perm], perm)
    assert base == permuted
    shift = np.array([7.25, -3.5, 11.0])
    moved = np.mod(pos + shift, L)
    assert pair_set(moved, r, np.arange(n)) == base


@pytest.mark.parametrize("dim,L", [(2, 2.0), (2, 4.0), (2, 6.0), (3, 2.0), (3, 4.0), (3, 5.0)])
def test_small_box_cell_dims_below_three(ora, dim, L):
    """cellDim < 3 maps several stencil offsets onto one cell; the search must
    still produce each minimum-image pair exactly once."""
    rng = np.random.default_rng(5)
    n = 40 if dim == 3 else 25
    pos = rng.uniform(0, L, size=(n, 3))
    r = rng.uniform(0.4, 0.8, size=n)
    f = foam(ora, pos, r, dim, L=L)
    assert min(f.scalars()["cell_dim"][:dim]) < 3
    assert np.array_equal(f.pairs(), f.brute_pairs())


def test_walls_do_not_wrap(ora):
    # two bubbles at opposite faces: neighbours through a periodic face only
    pos = [[0.2, 5.0, 5.0], [9.9, 5.0, 5.0]]
    r = [0.5, 0.5]
    per = foam(ora, pos, r, 3, L=10.0)
    wal = foam(ora, pos, r, 3, L=10.0, wall=(1, 0, 0))
    assert len(per.pairs()) == 1
    assert len(wal.pairs()) == 0


# --------------------------------------------------------------------------
# pair force and overlap area: two-bubble closed forms
# --------------------------------------------------------------------------
def lens_a_from_first(r1, r2, d):
    """Squared radius of the circle where two spheres meet, from sphere 1."""
    x1 = (d * d + r1 * r1 - r2 * r2) / (2 * d)
    return r1 * r1 - x1 * x1


@pytest.mark.parametrize("dim", [2, 3])
@pytest.mark.parametrize("r1,r2,d", [(1.0, 1.0, 1.7), (0.8, 1.3, 1.9), (1.2, 0.6, 1.5),
                                     (1.0, 0.3, 0.5), (0.4, 1.1, 0.6)])
def test_two_bubble_velocity_and_overlap(ora, dim, r1, r2, d):
    L = 20.0
    pos = [[5.0, 5.0, 5.0], [5.0 + d * 0.6, 5.0 + d * 0.8, 5.0]]
    f0 = 1.7
    f = foam(ora, pos, [r1, r2], dim, L=L, f0=f0, dt=1e-6, tol=1e9)
    f.attempt(gas=True, path=False)
    o = f.attempt_out()
    st, _ = f.get_state()
    # identify which slot holds which bubble after the cell sort
    k1 = int(np.where(st["idx"] == 0)[0][0])
    k2 = 1 - k1
    dvec = np.array([0.6 * d, 0.8 * d])
    u = f0 * (1.0 / d - 1.0 / (r1 + r2)) * dvec
    assert np.allclose(o["vp"][:2, k2], u, rtol=1e-13, atol=1e-15)
    assert np.allclose(o["vp"][:2, k1], -u, rtol=1e-13, atol=1e-15)
    if d * d < r1 * r1 or d * d < r2 * r2:
        a = min(r1, r2) ** 2
    else:
        a = abs(lens_a_from_first(r1, r2, d))
    oa = PI * a if dim == 3 else 2.0 * math.sqrt(a)
    assert o["ov"][k1] == pytest.approx(oa, rel=1e-12)
    assert o["ov"][k2] == pytest.approx(oa, rel=1e-12)


def test_non_overlapping_listed_pair_has_no_force(ora):
    # inside the skin but not overlapping: listed, no contribution
    f = foam(ora, [[5, 5, 5], [7.1, 5, 5]], [1.0, 1.0], 3, dt=1e-6, tol=1e9)
    assert len(f.pairs()) == 1
    f.attempt()
    o = f.attempt_out()
    assert np.all(o["vp"][:3] == 0.0) and np.all(o["ov"] == 0.0)


# --------------------------------------------------------------------------
# gas exchange
# --------------------------------------------------------------------------
@pytest.mark.parametrize("dim", [2, 3])
def test_liquid_mediated_flux_two_free_bubbles(ora, dim):
    r1, r2 = 0.8, 1.4
    kappa, K, ain = 0.05, 0.1, 9.0
    f = foam(ora, [[3, 3, 3], [12, 12, 12]], [r1, r2], dim, kappa=kappa, k=K, ain=ain, dt=1e-9,
             tol=1e9)
    f.attempt()
    o = f.attempt_out()
    st, _ = f.get_state()
    r = st["x"][3]

    def area(x):
        return 2 * PI * x if dim == 2 else 4 * PI * x * x

    s_a = area(r1) + area(r2)
    s_ar = area(r1) / r1 + area(r2) / r2
    inv_rho = s_ar / s_a
    for k in range(2):
        want = K * kappa * ain * 2 / s_a * (inv_rho - 1.0 / r[k])
        assert o["vp"][3][k] == pytest.approx(want, rel=1e-12)
    # volume is conserved by the RHS
    assert abs(area(r[0]) * o["vp"][3][0] + area(r[1]) * o["vp"][3][1]) < 1e-15


def test_single_bubble_radius_constant(ora):
    f = foam(ora, [[3, 3, 3]], [1.1], 3, dt=1e-3, tol=1e9)
    f.attempt()
    assert f.attempt_out()["vp"][3][0] == 0.0


@pytest.mark.parametrize("dim,n", [(2, 2000), (3, 3000)])
def test_gas_rhs_conserves_volume_and_force_antisymmetric(ora, dim, n):
    f = ora.Foam.synthetic(dim=dim, count=n, phi=0.9, seed=3, radius_dist=dim - 2, r_std=0.15)
    for _ in range(3):
        f.integrate()
    st, sc = f.get_state()
    dt = sc["dt"]
    f.attempt()
    o = f.attempt_out()
    # radii the attempt used: the AB2 predictor from the committed state
    rpred = st["x"][3] + 0.5 * dt * (3.0 * st["v"][3] - st["vo"][3])
    area = 2 * PI * rpred if dim == 2 else 4 * PI * rpred ** 2
    flux = area * o["vp"][3]
    assert abs(flux.sum()) <= 1e-10 * np.abs(flux).sum()
    vsum = o["vp"][:dim].sum(axis=1)
    vabs = np.abs(o["vp"][:dim]).sum(axis=1)
    assert np.all(np.abs(vsum) <= 1e-12 * vabs)


# --------------------------------------------------------------------------
# predictor / corrector and the dt controller
# --------------------------------------------------------------------------
def test_ab2_predictor_and_trapezoid_corrector_free_bubble(ora):
    dt = 0.01
    v = np.zeros((4, 1))
    vo = np.zeros((4, 1))
    v[:3, 0] = [0.3, -0.2, 0.1]
    vo[:3, 0] = [0.1, 0.4, -0.5]
    f = foam(ora, [[5, 5, 5]], [1.0], 3, v=v, vo=vo, dt=dt, tol=1e9)
    f.attempt(gas=False, path=True)
    o = f.attempt_out()
    x0 = np.array([5.0, 5.0, 5.0])
    pred = x0 + 0.5 * dt * (3 * v[:3, 0] - vo[:3, 0])
    corr = x0 + 0.5 * dt * (v[:3, 0] + 0.0)
    assert np.allclose(o["xc"][:3, 0], corr, rtol=0, atol=1e-15)
    assert f.scalars()["max_err"] == pytest.approx(np.max(np.abs(pred - corr)), rel=1e-12)
    assert o["path_new"][0] == pytest.approx(np.linalg.norm(corr - x0), rel=1e-12)


def test_dt_controller_reject_and_grow(ora):
    # rejection: any error exceeds a tiny tolerance -> dt *= 0.37, no commit
    v = np.zeros((4, 1))
    v[0, 0] = 1.0
    f = foam(ora, [[5, 5, 5]], [1.0], 3, v=v, dt=1e-3, tol=1e-300)
    assert f.attempt() == 0
    assert f.scalars()["dt"] == pytest.approx(1e-3 * 0.37, rel=1e-15)
    assert f.attempt() == 0
    assert f.scalars()["dt"] == pytest.approx(1e-3 * 0.37 * 0.37, rel=1e-15)
    # growth: error < 0.45 tol and dt < 10 -> dt *= 1.269 for the next step; time uses it
    g = foam(ora, [[5, 5, 5]], [1.0], 3, v=v, dt=1e-3, tol=1e30)
    dts = []
    t = 0.0
    dt = 1e-3
    for _ in range(5):
        g.integrate()
        dt *= 1.269
        t += dt
        dts.append(g.scalars()["dt"])
    assert np.allclose(dts, [1e-3 * 1.269 ** k for k in range(1, 6)], rtol=1e-14)
    sc = g.scalars()
    assert sc["time_int"] + sc["time_frac"] == pytest.approx(t, rel=1e-14)
    # growth stops once dt >= 10
    h = foam(ora, [[5, 5, 5]], [1.0], 3, v=v, dt=9.0, tol=1e30)
    h.integrate()
    assert h.scalars()["dt"] == pytest.approx(9.0 * 1.269)
    h.integrate()
    assert h.scalars()["dt"] == pytest.approx(9.0 * 1.269)


def test_error_between_tolerances_keeps_dt(ora):
    v = np.zeros((4, 1))
    v[0, 0] = 1.0
    # error = |AB2 - trapezoid| = dt * |v| = 1e-3 ; tolerance 1.5e-3 -> between 0.45 tol and tol
    f = foam(ora, [[5, 5, 5]], [1.0], 3, v=v, dt=1e-3, tol=1.5e-3)
    assert f.attempt() == 1
    assert f.scalars()["dt"] == 1e-3


# --------------------------------------------------------------------------
# wrap, removal, volume fix
# --------------------------------------------------------------------------
def test_wrap_counters_preserve_unwrapped_position(ora):
    L = 10.0
    pos = [[-0.5, 3.0, 10.25], [4.0, 10.75, 2.0], [9.0, -1.25, 5.0]]
    f = foam(ora, [[1, 1, 1]] * 3, [0.5] * 3, 3, L=L)
    st, sc = f.get_state()
    st["x"][:3] = np.asarray(pos).T
    f.set_state(st, sc)
    f.search()
    st2, _ = f.get_state()
    order = np.argsort(st2["idx"])
    x = st2["x"][:3, order]
    wc = st2["wc"][:, order]
    assert np.all((x >= 0) & (x <= L))
    # x += L * m and wc -= m, so the unwrapped position is x + wc * L
    assert np.allclose(x + wc * L, np.asarray(pos).T, atol=1e-12)


@pytest.mark.parametrize("dim,n,frac", [(2, 1500, 0.02), (3, 2500, 0.05), (3, 2000, 0.3)])
def test_removal_compaction_and_volume_fix(ora, dim, n, frac):
    f = ora.Foam.synthetic(dim=dim, count=n, phi=0.9, seed=9, radius_dist=dim - 2, r_std=0.15)
    st, sc = f.get_state()
    rng = np.random.default_rng(1)
    kill = rng.random(n) < frac
    st["x"][3][kill] = 0.05  # below minRad = 0.1: marked by the corrector
    f.set_state(st, sc)
    pairs_before = {tuple(p) for p in f.pairs()}
    idx_before = st["idx"].copy()
    # one attempt marks the bubbles below minRad; remove() then compacts the
    # (uncommitted) current state with those marks and S_V
    f.attempt()
    o = f.attempt_out()
    dele = o["del"].astype(bool)
    assert dele.sum() >= kill.sum()
    assert np.all(dele[kill])
    sv = f.scalars()["s_v"]
    nd = int(dele.sum())
    st_pre, _ = f.get_state()
    f.remove()
    st2, sc2 = f.get_state()
    # the imposed flow's divisor moves with its bubble and is not reduced by
    # the dropped pairs (the reference's numNeighbors, src/kernels.cu:820-827)
    deg_before = dict(zip(idx_before.tolist(), st_pre["deg"].tolist()))
    assert [deg_before[i] for i in st2["idx"].tolist()] == st2["deg"].tolist()
    assert np.all(st2["deg"] >= np.diff(st2["nbr_off"]))
    assert np.any(st2["deg"] > np.diff(st2["nbr_off"]))
    nnew = n - nd
    assert sc2["count"] == nnew
    survivors = idx_before[~dele]
    assert sorted(st2["idx"].tolist()) == sorted(survivors.tolist())
    # the first nnew slots keep survivors in place; holes get tail survivors, last first
    holes = [i for i in np.where(dele)[0] if i < nnew]
    tail = [k for k in range(n - 1, nnew - 1, -1) if not dele[k]]
    expect_idx = idx_before[:nnew].copy()
    for h, s in zip(holes, tail):
        expect_idx[h] = idx_before[s]
    assert np.array_equal(st2["idx"], expect_idx)
    # pair set: old pairs among survivors, expressed in original ids
    pairs_after = {tuple(sorted((st2["idx"][a], st2["idx"][b]))) for a, b in
                   ora.half_pairs(st2["nbr_off"], st2["nbr"])}
    want = {tuple(sorted((idx_before[a], idx_before[b]))) for a, b in pairs_before
            if not dele[a] and not dele[b]}
    assert pairs_after == want
    # volume fix scaling law r_new = r_old * M / S_V^(1/D) (the full
    # phi * V_box restoration is checked through integrate() below)
    c = f.consts()
    vm = c["vol_mul"] / (sv ** (1.0 / dim))
    order_old = {int(i): k for k, i in enumerate(idx_before)}
    r_old = np.array([st_pre["x"][3][order_old[int(i)]] for i in st2["idx"]])
    assert np.allclose(st2["x"][3], r_old * vm, rtol=1e-14)


@pytest.mark.parametrize("dim", [2, 3])
def test_volume_fix_restores_phi_in_integrate(ora, dim):
    n = 2000
    f = ora.Foam.synthetic(dim=dim, count=n, phi=0.85, seed=21, radius_dist=dim - 2, r_std=0.15)
    st, sc = f.get_state()
    st["x"][3][::97] = 0.02
    f.set_state(st, sc)
    f.integrate()
    sc2 = f.scalars()
    assert sc2["removals"] == 1 and sc2["count"] < n
    st2, _ = f.get_state()
    c = f.consts()
    vbox = np.prod(c["interval"][:dim])
    r = st2["x"][3]
    vol = (PI * r ** 2).sum() if dim == 2 else (4.0 / 3.0 * PI * r ** 3).sum()
    assert abs(vol - 0.85 * vbox) / (0.85 * vbox) < 1e-10


# --------------------------------------------------------------------------
# synthetic initial state
# --------------------------------------------------------------------------
@pytest.mark.parametrize("dim,dist,phi", [(2, 0, 0.9), (3, 1, 0.84)])
def test_synthetic_init_box_and_distributions(ora, dim, dist, phi):
    n = 20000
    f = ora.Foam.synthetic(dim=dim, count=n, phi=phi, seed=426, radius_dist=dist,
                           r_std=0.1 if dist == 0 else 0.15)
    st, sc = f.get_state()
    c = f.consts()
    r = st["x"][3]
    vbox = np.prod(c["interval"][:dim])
    vol = (PI * r ** 2).sum() if dim == 2 else (4.0 / 3.0 * PI * r ** 3).sum()
    assert vol / vbox == pytest.approx(phi, rel=1e-12)
    assert np.all((st["x"][:dim] >= 0) & (st["x"][:dim] < np.array(c["interval"][:dim])[:, None]))
    if dist == 0:
        assert r.min() >= 0.5 and r.max() <= 1.5
        assert abs(r.mean() - 1.0) < 0.005 and abs(r.std() - 0.1) < 0.005
    else:
        assert abs(np.median(np.log(r))) < 0.005 and abs(np.log(r).std() - 0.15) < 0.005
    area = 2 * PI * r if dim == 2 else 4 * PI * r ** 2
    assert c["avg_area_in"] == pytest.approx(area.mean(), rel=1e-12)
    m = math.sqrt(phi * vbox / PI) if dim == 2 else (0.75 * phi * vbox / PI) ** (1 / 3)
    assert c["vol_mul"] == pytest.approx(m, rel=1e-14)
    assert c["skin"] == 0.3 and c["min_rad"] == pytest.approx(0.1) and c["f0"] == 1.0


def test_integrate_runs_and_counts(ora):
    f = ora.Foam.synthetic(dim=2, count=1000, seed=426)
    for _ in range(30):
        f.integrate()
    sc = f.scalars()
    assert sc["steps"] == 30 and sc["attempts"] >= 30
    assert sc["bubble_steps"] == 30 * 1000
    st, _ = f.get_state()
    assert np.all(np.isfinite(st["x"])) and np.all(np.isfinite(st["v"]))


def test_openmp_matches_serial_closely(ora):
    a = ora.Foam.synthetic(dim=3, count=4000, seed=5, radius_dist=1, r_std=0.15, phi=0.84)
    b = ora.Foam.synthetic(dim=3, count=4000, seed=5, radius_dist=1, r_std=0.15, phi=0.84)
    a.run(5)
    ora.set_threads(4)
    try:
        b.run(5)
    finally:
        ora.set_threads(1)
    sa, _ = a.get_state()
    sb, _ = b.get_state()
    assert np.array_equal(sa["idx"], sb["idx"])
    assert np.allclose(sa["x"], sb["x"], rtol=1e-12, atol=1e-12)


# --------------------------------------------------------------------------
# observables of the printed rows and snapshots
# --------------------------------------------------------------------------
@pytest.mark.parametrize("wall", [0, 1])
def test_observe_two_bubble_energy_closed_form(ora, wall):
    # a pair straddling x = 0: 0.4 apart through the periodic boundary, 19.6
    # apart inside the box when x is a wall; a third bubble far from both
    L = 20.0
    pos = [(0.2, 5.0), (L - 0.2, 5.0), (10.0, 12.0)]
    r = [0.5, 0.6, 0.7]
    f = foam(ora, pos, r, 2, L=L, wall=(wall, 0, 0))
    sums, e = f.observe()
    o = 0.0 if wall else (0.5 + 0.6 - 0.4) ** 2
    assert e[0] == pytest.approx(o, rel=1e-13, abs=0)
    assert e[1] == pytest.approx(o, rel=1e-13, abs=0)
    assert e[2] == 0.0
    assert sums[0] == pytest.approx(2 * o, rel=1e-13, abs=0)
    assert sums[1] == pytest.approx(PI * sum(x * x for x in r), rel=1e-15)
    assert sums[2] == pytest.approx(sum(r), rel=1e-15)


@pytest.mark.parametrize("dim,n,phi,seed", [(2, 700, 0.9, 11), (3, 900, 0.9, 12)])
def test_observe_matches_brute_force_minimum_image(ora, dim, n, phi, seed):
    f = ora.Foam.synthetic(dim=dim, count=n, phi=phi, seed=seed, radius_dist=dim - 2, r_std=0.15)
    f.run(25)  # nonzero velocities and paths
    st, sc = f.get_state()
    L = np.asarray(f.consts()["interval"][:dim])
    x = st["x"][:dim].T
    r = st["x"][3]
    d = x[:, None, :] - x[None, :, :]
    d -= L * np.round(d / L)
    o = r[:, None] + r[None, :] - np.sqrt((d * d).sum(-1))
    np.fill_diagonal(o, 0.0)
    e_bf = np.where(o > 0, o * o, 0.0).sum(1)
    sums, e = f.observe()
    assert e_bf.sum() > 0
    assert np.allclose(e, e_bf, rtol=1e-12, atol=1e-15)
    vol = PI * r ** 2 * (4.0 / 3.0 * r if dim == 3 else 1.0)
    assert sums[0] == pytest.approx(e_bf.sum(), rel=1e-12)
    assert sums[1] == pytest.approx(vol.sum(), rel=1e-13)
    assert sums[2] == pytest.approx(r.sum(), rel=1e-13)
    assert sums[3] == pytest.approx(st["path"].sum(), rel=1e-13)
    assert np.abs(st["path"]).sum() > 0
    for q in range(4):
        want = st["v"][q].sum() if (q != 2 or dim == 3) else 0.0
        assert sums[4 + q] == pytest.approx(want, rel=1e-12, abs=1e-12 * np.abs(st["v"][q]).sum())


# --------------------------------------------------------------------------
# imposed flow
# --------------------------------------------------------------------------
def test_flow_closed_forms(ora):
    """Velocities are set, positions are fixed apart: listed pairs that do not
    overlap exert no force, so the predicted velocity is the flow term alone.
    Inside the box the bubble takes the imposed velocity; outside, the sum of
    the velocities of listed neighbours with a smaller index and of
    overlapping ones with a larger index, over its list length."""
    vel = (0.3, -0.2)
    pos = [(4.0, 4.0),                            # 0: inside the flow box
           (6.3, 15.0 - 11.0),                    # 1: 0.3 past a face, r 0.5: inside
           (10.0, 15.0), (11.1, 15.0), (12.2, 15.0),  # 2-4: a chain 1.1 apart
           (15.0, 5.0), (15.9, 5.0)]              # 5-6: overlapping
    r = [0.5] * 7
    v = np.zeros((4, 7))
    v[0, :] = [5.0, 6.0, 1.0, 10.0, 100.0, 7.0, 70.0]
    v[1, :] = [-5.0, -6.0, 2.0, 20.0, 200.0, 8.0, 80.0]
    c = consts(2, flow_lbb=(2.0, 2.0, 0.0), flow_tfr=(6.0, 6.0, 0.0), flow_vel=vel + (0.0,))
    st = make_state(pos, r, 2, v, v)
    # the lists a search gives (listed: within r_i + r_j + skin = 1.3), set
    # by hand so that the indices stay as written
    lists = [[], [], [3], [2, 4], [3], [6], [5]]
    st["nbr_off"] = np.cumsum([0] + [len(x) for x in lists]).astype(np.int32)
    st["nbr"] = np.array(sum(lists, []), np.int32)

    def run(flow):
        f = ora.Foam.from_state(c, st, scalars(7))
        f.attempt(gas=False, path=False, flow=flow)
        return f, f.attempt_out()["vp"]

    f, vp = run(True)
    assert f.get_state()[0]["deg"].tolist() == [0, 0, 1, 2, 1, 1, 1]
    _, vp0 = run(False)
    assert np.all(vp0[:2, :5] == 0.0)            # no overlaps among 0-4
    for i in (0, 1):
        assert tuple(vp[:2, i]) == vel
    assert tuple(vp[:2, 2]) == (0.0, 0.0)        # its neighbour 3 is larger, apart
    assert tuple(vp[:2, 3]) == (0.5, 1.0)        # v2 / 2 (4 is larger, apart)
    assert tuple(vp[:2, 4]) == (10.0, 20.0)      # v3 / 1
    # overlapping: each takes the other's velocity, on top of the pair force
    assert vp[0, 5] - vp0[0, 5] == pytest.approx(70.0, rel=1e-14)
    assert vp[1, 5] - vp0[1, 5] == pytest.approx(80.0, rel=1e-14)
    assert vp[0, 6] - vp0[0, 6] == pytest.approx(7.0, rel=1e-14)
    assert vp[1, 6] - vp0[1, 6] == pytest.approx(8.0, rel=1e-14)


@pytest.mark.parametrize("dim", [2, 3])
def test_flow_box_edges_count_within_radius(ora, dim):
    """One free bubble per position: inside strictly between the faces or
    within its radius of a face, on every axis."""
    lo, hi, rr = 5.0, 9.0, 0.5
    vel = (0.25, 0.5, -0.75)
    cases = [(7.0, True), (9.4, True), (9.5, True), (9.6, False), (4.6, True), (4.4, False)]
    for x, inside in cases:
        p = [(x, 7.0, 7.0)[:dim]]
        kw = dict(flow_lbb=(lo,) * 3, flow_tfr=(hi,) * 3, flow_vel=vel, dt=1e-4)
        f = foam(ora, p, [rr], dim, **kw)
        f.attempt(gas=False, path=False, flow=True)
        vp = f.attempt_out()["vp"]
        want = vel[:dim] if inside else (0.0,) * dim
        assert tuple(vp[:dim, 0]) == want, (x, inside)


# --------------------------------------------------------------------------
# the AB2 velocity bootstrap of init
# --------------------------------------------------------------------------
def test_bootstrap_two_bubble_closed_form(ora):
    """Two bubbles of radius 0.5 0.8 apart (f0 = 1): the pair velocities are
    -/+0.2 along x; after an Euler step of dt they are 0.8 + 0.4 dt apart and
    the velocities there, -/+(0.2 - 0.4 dt), become the old velocities.  A
    third bubble far away keeps zero; the current velocities stay zero."""
    dt = 1e-4
    st = make_state([(10.0, 10.0), (10.8, 10.0), (3.0, 3.0)], [0.5, 0.5, 0.5], 2)
    st["nbr_off"] = np.array([0, 1, 2, 2], np.int32)
    st["nbr"] = np.array([1, 0], np.int32)
    f = ora.Foam.from_state(consts(2), st, scalars(3, dt=dt))
    f.bootstrap()
    got, _ = f.get_state()
    want = 0.2 - 0.4 * dt
    assert got["vo"][0, 0] == pytest.approx(-want, rel=1e-12)
    assert got["vo"][0, 1] == pytest.approx(want, rel=1e-12)
    assert got["vo"][0, 2] == 0.0 and not np.any(got["vo"][1])
    assert not np.any(got["v"])
    assert np.array_equal(got["x"], st["x"])


# --------------------------------------------------------------------------
# the box rescale of init
# --------------------------------------------------------------------------
@pytest.mark.parametrize("dim,rel,phi", [(2, (2.0, 1.0, 1.0), 0.7), (3, (1.0, 1.5, 1.0), 0.8)])
def test_rescale_keeps_relative_positions_and_fills_phi(ora, dim, rel, phi):
    f = ora.Foam.synthetic(dim=dim, count=900, phi=0.5, seed=31, radius_dist=dim - 2, r_std=0.15)
    st0, _ = f.get_state()
    c0 = f.consts()
    vol = f.observe()[0][1]
    f.rescale(phi, rel, vol)
    st1, _ = f.get_state()
    c1 = f.consts()
    L0 = np.array(c0["interval"][:dim])
    L1 = np.array(c1["interval"][:dim])
    # the box holds phi, in the proportions rel
    assert vol / np.prod(L1) == pytest.approx(phi, rel=1e-13)
    assert np.allclose(L1 / np.array(rel[:dim]), L1[0] / rel[0], rtol=1e-14)
    # every bubble keeps its relative place (matched by original index)
    o0 = np.argsort(st0["idx"])
    o1 = np.argsort(st1["idx"])
    rel0 = (st0["x"][:dim, o0] - np.array(c0["lbb"][:dim])[:, None]) / L0[:, None]
    rel1 = (st1["x"][:dim, o1] - np.array(c1["lbb"][:dim])[:, None]) / L1[:, None]
    assert np.allclose(rel0, rel1, rtol=0, atol=1e-14)
    # volume multiplier: M = sqrt(phi V_box / pi) or cbrt(3/4 phi V_box / pi)
    m = phi * np.prod(L1) / PI
    want = np.sqrt(m) if dim == 2 else np.cbrt(0.75 * m)
    assert c1["vol_mul"] == pytest.approx(want, rel=1e-14)
    # the lists of the new box are the brute-force pairs
    assert np.array_equal(f.pairs(), f.brute_pairs())


# --------------------------------------------------------------------------
# walls (reference addWallVelocity, src/kernels.cu:546-593)
#
# Durian's wall term is the pair term of the bubble's mirror image: a bubble
# at distance d < r from a wall overlaps its reflection (same radius, at 2d)
# and f0 (1/(2d) - 1/(2r)) 2d = f0 (1 - d/r), directed away from the wall.
# A touched wall drags the velocity components along it by (1 - drag), so a
# component k ends as (v_k + push_k) * (1 - drag)^(walls touched other than
# k).  "Touching" includes d == r exactly (push 0, drag applied).
# --------------------------------------------------------------------------
WALL_F0, WALL_DRAG = 1.7, 0.3


def _wall_consts(dim, wall, L=10.0):
    return consts(dim, L=L, wall=wall, f0=WALL_F0, drag=WALL_DRAG)


@pytest.mark.parametrize("dim", [2, 3])
@pytest.mark.parametrize("side", ["low", "high"])
def test_wall_single_wall_closed_form(ora, dim, side):
    L, r, d = 10.0, 0.8, 0.35
    v0 = np.array([0.11, -0.07, 0.05])[:dim]
    keep = 1.0 - WALL_DRAG
    for a in range(dim):
        wall = [0, 0, 0]
        wall[a] = 1
        pos = np.full(dim, 4.0)
        pos[a] = d if side == "low" else L - d
        got = ora.wall_velocity(_wall_consts(dim, wall, L), pos, r, v0)[:dim]
        push = WALL_F0 * (1.0 - d / r) * (1.0 if side == "low" else -1.0)
        want = v0 * keep
        want[a] = v0[a] + push
        assert np.allclose(got, want, rtol=1e-14, atol=0), (a, side, got, want)


@pytest.mark.parametrize("dim", [2, 3])
@pytest.mark.parametrize("r,d", [(1.0, 0.3), (0.7, 0.65), (1.3, 0.05)])
def test_wall_push_equals_mirror_image_pair_velocity(ora, dim, r, d):
    # wall box: one bubble at distance d from the x = 0 wall
    L = 10.0
    got = ora.wall_velocity(_wall_consts(dim, [1, 0, 0], L), [d, 4.0, 4.0][:dim], r,
                            np.zeros(dim))[0]
    # periodic box: the same bubble and its mirror image at distance 2 d,
    # through the oracle's pair step (pinned by the two-bubble closed form)
    pos = [[5.0 + d, 5.0, 5.0], [5.0 - d, 5.0, 5.0]]
    f = foam(ora, pos, [r, r], dim, L=20.0, f0=WALL_F0, dt=1e-9, tol=1e9)
    f.attempt(gas=False, path=False)
    st, _ = f.get_state()
    k = int(np.where(st["idx"] == 0)[0][0])
    mirror = f.attempt_out()["vp"][0, k]
    assert got == pytest.approx(mirror, rel=1e-14)
    assert got == pytest.approx(WALL_F0 * (1.0 - d / r), rel=1e-14)


@pytest.mark.parametrize("dim,walls", [(2, (1, 1, 0)), (3, (1, 1, 0)), (3, (1, 0, 1)),
                                       (3, (0, 1, 1)), (3, (1, 1, 1))])
def test_wall_corner_drag_law(ora, dim, walls):
    L, r = 10.0, 0.9
    dist = np.array([0.2, 0.5, 0.7])  # distance to the touched wall per axis
    pos = np.full(3, 5.0)
    low = np.array([True, False, True])
    for a in range(dim):
        if walls[a]:
            pos[a] = dist[a] if low[a] else L - dist[a]
    v0 = np.array([0.21, -0.13, 0.08])
    got = ora.wall_velocity(_wall_consts(dim, list(walls), L), pos, r, v0)
    keep = 1.0 - WALL_DRAG
    touched = [a for a in range(dim) if walls[a]]
    for k in range(dim):
        push = WALL_F0 * (1.0 - dist[k] / r) * (1.0 if low[k] else -1.0) if walls[k] else 0.0
        others = sum(1 for a in touched if a != k)
        want = (v0[k] + push) * keep ** others
        assert got[k] == pytest.approx(want, rel=1e-14), (k, got, want)


@pytest.mark.parametrize("dim", [2, 3])
def test_wall_not_touching_and_touching_at_contact(ora, dim):
    L, r = 10.0, 0.5
    v0 = np.array([0.3, -0.2, 0.1])[:dim]
    c = _wall_consts(dim, [1, 1, 1], L)
    # farther than r from every wall: unchanged, bit for bit
    got = ora.wall_velocity(c, np.full(dim, 5.0), r, v0)[:dim]
    assert np.array_equal(got, v0)
    far = ora.wall_velocity(c, [r + 1e-9] + [5.0] * (dim - 1), r, v0)[:dim]
    assert np.array_equal(far, v0)
    # exactly r from the x = 0 wall (1.5 - 1.0 is exact): no push, drag applies
    at = ora.wall_velocity(_wall_consts(dim, [1, 1, 1], L), [0.5] + [5.0] * (dim - 1), r, v0)[:dim]
    want = v0 * (1.0 - WALL_DRAG)
    want[0] = v0[0]
    assert np.allclose(at, want, rtol=1e-15, atol=0)


def test_wall_velocity_through_an_attempt(ora):
    # one bubble near the lower x wall and the upper y wall, zero history:
    # the predicted velocity is the wall term alone, and the corrector moves
    # the bubble by dt/2 times it (trapezoid with zero current velocity)
    L, r, dt = 10.0, 0.8, 1e-3
    pos = [0.5, L - 0.6, 5.0]
    f = foam(ora, [pos], [r], 3, L=L, wall=(1, 1, 0), f0=WALL_F0, drag=WALL_DRAG, dt=dt, tol=1e9)
    f.attempt(gas=False, path=False)
    o = f.attempt_out()
    keep = 1.0 - WALL_DRAG
    ux = WALL_F0 * (1.0 - 0.5 / r)
    uy = -WALL_F0 * (1.0 - 0.6 / r)
    want = np.array([ux * keep, uy * keep, 0.0])
    assert np.allclose(o["vp"][:3, 0], want, rtol=1e-14, atol=0)
    assert np.allclose(o["xc"][:3, 0], np.array(pos) + 0.5 * dt * want, rtol=0, atol=1e-15)


# --------------------------------------------------------------------------
# expansion since the last search and the search trigger
# (reference src/kernels.cu:369-387, :412, :487-495; src/simulator.cu:323)
#
# A free bubble (no neighbours, no walls) has zero predicted velocity, so the
# trapezoid gives x_c = x + dt/2 v, r_c = r + dt/2 (dr/dt + gas term), and its
# expansion is |x_c - x_saved| + (r_c - r_saved).  Equal radii make the
# liquid-mediated gas term zero (1/rho = 1/r).  A search follows an accepted
# step whose largest expansion is at least half the skin.
# --------------------------------------------------------------------------
def _free_foam(ora, pos, r, v, saved, dt, skin=0.3, gas_drdt=0.0):
    n = len(r)
    vv = np.zeros((4, n))
    vv[:3] = np.asarray(v, float).T
    vv[3] = gas_drdt
    f = foam(ora, pos, r, 3, v=vv, dt=dt, tol=1e9, skin=skin)
    st, sc = f.get_state()
    order = st["idx"]
    st["saved"] = np.asarray(saved, float).T[:, order].copy()
    f.set_state(st, sc)
    return f, order


def test_expansion_free_bubbles_closed_form(ora):
    dt, drdt = 0.2, 0.05
    pos = np.array([[3.0, 3.0, 3.0], [10.0, 10.0, 10.0], [16.0, 4.0, 12.0]])
    r = np.array([1.0, 1.0, 1.0])
    v = np.array([[0.3, -0.4, 0.1], [0.05, 0.02, -0.9], [-0.2, 0.2, 0.2]])
    saved = np.column_stack([pos - np.array([[0.01, 0.02, -0.03], [0.0, 0.0, 0.0],
                                             [0.04, -0.01, 0.02]]),
                             r - np.array([0.002, 0.03, -0.01])])
    f, order = _free_foam(ora, pos, r, v, saved, dt, gas_drdt=drdt)
    f.attempt(gas=True, path=False)
    xc = pos + 0.5 * dt * v
    rc = r + 0.5 * dt * drdt
    exp = np.linalg.norm(xc - saved[:, :3], axis=1) + (rc - saved[:, 3])
    assert f.scalars()["max_exp"] == pytest.approx(exp.max(), rel=1e-12)
    # the bubble with the largest expansion decides (bubble 1: it moved most)
    assert int(np.argmax(exp)) == 1
    # without gas exchange the radius term drops out
    g, _ = _free_foam(ora, pos, r, v, saved, dt, gas_drdt=drdt)
    g.attempt(gas=False, path=False)
    assert g.scalars()["max_exp"] == pytest.approx(
        np.linalg.norm(xc - saved[:, :3], axis=1).max(), rel=1e-12)


@pytest.mark.parametrize("saved_x,want", [(0.25, 1), (0.25 + 2.0 ** -54, 0)])
def test_search_trigger_at_half_skin(ora, saved_x, want):
    # skin 0.5: a bubble at rest at x = 0.5, its saved position 0.25 (exactly
    # half the skin away) or one ulp above it (0.25 - 2^-54 away: both values
    # and their difference are exact doubles); no gas exchange, so no radius
    # term, and sqrt(t * t) == t
    skin = 0.5
    pos = np.array([[0.5, 5.0, 5.0]])
    saved = np.array([[saved_x, 5.0, 5.0, 1.0]])
    f, _ = _free_foam(ora, pos, [1.0], np.zeros((1, 3)), saved, dt=1e-3, skin=skin)
    f.attempt(gas=False, path=False)
    sc = f.scalars()
    assert sc["max_exp"] == 0.5 - saved_x
    assert (sc["max_exp"] >= 0.25) == bool(want)
    assert sc["need_search"] == want
    g, _ = _free_foam(ora, pos, [1.0], np.zeros((1, 3)), saved, dt=1e-3, skin=skin)
    before = g.scalars()["searches"]
    g.integrate(gas=False, path=False)
    assert g.scalars()["searches"] - before == want
    # a search resets the saved positions to the current ones
    if want:
        st, _ = g.get_state()
        assert np.array_equal(st["saved"][:3, 0], st["x"][:3, 0])
