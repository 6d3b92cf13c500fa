"""cubble_find_pairs: the reference's pair output (every pair i < j within
r_i + r_j + skin under the minimum image, grouped by i, ascending in j) for
bubbles in caller-owned device memory, bit-exact against the oracle's O(N^2)
brute force on the wrapped positions.  Device memory comes from torch."""
import numpy as np
import pytest

import cubble_amd as ca

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def _consts(dim, L, wall, skin):
    iv = [L[0], L[1], L[2] if dim == 3 else 0.0]
    return {"dim": dim, "wall": list(wall), "lbb": [0.0, 0.0, 0.0], "tfr": iv, "interval": iv,
            "wall_drag": 0.1, "skin": skin, "min_rad": 0.1, "f0": 1.0, "k_param": 0.1,
            "kappa": 0.01, "avg_area_in": 1.0, "vol_mul": 1.0, "tol": 1e-4, "cell_cap": 1 << 20}


def brute(ora, dim, L, wall, skin, x4):
    """The oracle's brute-force pairs of positions x4 (4 x n, inside the box)."""
    n = x4.shape[1]
    st = {"x": x4.copy(), "v": np.zeros((4, n)), "vo": np.zeros((4, n)), "saved": x4.copy(),
          "path": np.zeros(n), "err": np.zeros(n), "wc": np.zeros((3, n), np.int32),
          "idx": np.arange(n, dtype=np.int32), "nbr_off": np.zeros(n + 1, np.int32),
          "nbr": np.zeros(0, np.int32)}
    sc = {"count": n, "cell_dim": [1, 1, 1], "dt": 1e-4, "time_int": 0, "time_frac": 0.0,
          "max_r_cell": float(x4[3].max()) if n else 1.0, "steps": 0, "attempts": 0,
          "searches": 0, "removals": 0, "bubble_steps": 0.0, "s_a": 0.0, "s_ar": 0.0,
          "s_oa": 0.0, "s_oar": 0.0, "s_v": 0.0, "max_err": 0.0, "max_r": 0.0, "max_exp": 0.0,
          "n_del": 0, "accepted": 0, "need_search": 0, "num_pairs": 0}
    f = ora.Foam.from_state(_consts(dim, L, wall, skin), st, sc)
    return f.brute_pairs()


def gpu_pairs(dim, L, wall, skin, x4_in, cap, find=ca.find_pairs):
    n = x4_in.shape[1]
    dev = [torch.tensor(np.ascontiguousarray(x4_in[k]), dtype=torch.float64, device="cuda")
           for k in range(4)]
    pi = torch.full((max(cap, 1),), -1, dtype=torch.int32, device="cuda")
    pj = torch.full((max(cap, 1),), -1, dtype=torch.int32, device="cuda")
    num = torch.full((1,), -1, dtype=torch.int64, device="cuda")
    box = ca.Box.of(dim, [0.0] * 3, L, wall)
    zp = dev[2].data_ptr() if dim == 3 else 0
    find(dev[0].data_ptr(), dev[1].data_ptr(), zp, dev[3].data_ptr(), n, box, skin, pi.data_ptr(),
         pj.data_ptr(), cap, num.data_ptr(), torch.cuda.current_stream().cuda_stream)
    torch.cuda.synchronize()
    total = int(num.cpu()[0])
    got = np.stack([pi.cpu().numpy()[:min(total, cap)], pj.cpu().numpy()[:min(total, cap)]], 1)
    return total, got


def foam_positions(rng, dim, n, L, rmin=0.4, rmax=1.4):
    x4 = np.zeros((4, n))
    for k in range(dim):
        x4[k] = rng.uniform(0.0, L[k], n)
    x4[3] = rng.uniform(rmin, rmax, n)
    return x4


CASES = [(2, 3000, (60.0, 60.0, 0.0), (0, 0, 0)), (3, 4000, (25.0, 25.0, 25.0), (0, 0, 0)),
         (3, 3000, (22.0, 30.0, 18.0), (1, 0, 1)), (2, 40, (5.5, 5.5, 0.0), (0, 0, 0)),
         (3, 300, (5.0, 5.0, 5.0), (0, 1, 0)), (2, 1, (10.0, 10.0, 0.0), (0, 0, 0))]


@pytest.mark.parametrize("dim,n,L,wall", CASES)
def test_find_pairs_equals_brute_force(ora, dim, n, L, wall):
    print(f"find_pairs dim={dim} n={n} L={L} wall={wall}", flush=True)
    rng = np.random.default_rng(n + dim)
    x4 = foam_positions(rng, dim, n, L)
    skin = 0.3
    want = brute(ora, dim, L, wall, skin, x4)
    total, got = gpu_pairs(dim, L, wall, skin, x4, 64 * n)
    assert total == len(want)
    assert np.array_equal(got, want)


@pytest.mark.parametrize("dim", [2, 3])
def test_find_pairs_wraps_positions_outside_the_box(ora, dim):
    rng = np.random.default_rng(7)
    L = (30.0, 30.0, 30.0)
    n = 3000
    x4 = foam_positions(rng, dim, n, L)
    moved = x4.copy()
    # a tenth of the bubbles one period out of the box, either side; the
    # search wraps them back with one period, (x +- L) -+ L in FP64
    for k in range(dim):
        up, down = rng.random(n) < 0.05, rng.random(n) < 0.05
        moved[k][up] += L[k]
        moved[k][down & ~up] -= L[k]
    wrapped = moved.copy()
    for k in range(dim):
        wrapped[k] = np.where(moved[k] > L[k], moved[k] - L[k],
                              np.where(moved[k] < 0.0, moved[k] + L[k], moved[k]))
    want = brute(ora, dim, L, (0, 0, 0), 0.3, wrapped)
    total, got = gpu_pairs(dim, L, (0, 0, 0), 0.3, moved, 64 * n)
    assert total == len(want) and np.array_equal(got, want)


def test_find_pairs_reports_the_count_beyond_capacity(ora):
    rng = np.random.default_rng(3)
    L = (25.0, 25.0, 25.0)
    x4 = foam_positions(rng, 3, 3000, L)
    want = brute(ora, 3, L, (0, 0, 0), 0.3, x4)
    cap = len(want) // 3
    total, got = gpu_pairs(3, L, (0, 0, 0), 0.3, x4, cap)
    assert total == len(want)
    # the groups that fit entirely are written as they are; the group at the
    # boundary holds some of its pairs; nothing lands beyond the capacity
    b = want[cap - 1, 0]
    whole = int(np.searchsorted(want[:, 0], b))
    assert np.array_equal(got[:whole], want[:whole])
    part = got[whole:]
    assert np.all(part[:, 0] == b)
    assert set(map(tuple, part)) <= set(map(tuple, want[want[:, 0] == b]))


def test_find_pairs_empty_and_bad_arguments():
    total, got = gpu_pairs(3, (10.0, 10.0, 10.0), (0, 0, 0), 0.3, np.zeros((4, 0)), 16)
    assert total == 0 and len(got) == 0
    with pytest.raises(ca.CubbleError) as e:
        gpu_pairs(3, (10.0, 0.0, 10.0), (0, 0, 0), 0.3, np.ones((4, 5)), 16)
    assert e.value.code == ca.E_INVALID and "box" in str(e.value)
