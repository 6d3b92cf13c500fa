"""cubble_find_pairs_noalloc (scratch in a fixed device array, no allocation)
against cubble_find_pairs through the C ABI: bit-equal pairs and counts on
two calls in a row, the second with fewer bubbles (what the first call left
in the shared array must not leak into the second), and the size limit
refused before any work."""
import ctypes

import numpy as np
import pytest

import cubble_amd as ca

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def _run(x4, dim, L, skin, cap, noalloc, stream):
    n = x4.shape[1]
    dev = [torch.tensor(np.ascontiguousarray(x4[k]), dtype=torch.float64, device="cuda")
           for k in range(4)]
    pi = torch.full((cap,), -1, dtype=torch.int32, device="cuda")
    pj = torch.full((cap,), -1, dtype=torch.int32, device="cuda")
    num = torch.full((1,), -1, dtype=torch.int64, device="cuda")
    box = ca.Box.of(dim, [0.0] * 3, L)
    ca.find_pairs(dev[0].data_ptr(), dev[1].data_ptr(), dev[2].data_ptr() if dim == 3 else 0,
                  dev[3].data_ptr(), n, box, skin, pi.data_ptr(), pj.data_ptr(), cap,
                  num.data_ptr(), stream, noalloc=noalloc)
    torch.cuda.synchronize()
    total = int(num.cpu()[0])
    return total, pi.cpu().numpy()[:min(total, cap)], pj.cpu().numpy()[:min(total, cap)]


def _foam(rng, dim, n, L):
    x4 = np.zeros((4, n))
    for k in range(dim):
        x4[k] = rng.uniform(0.0, L[k], n)
    x4[3] = rng.uniform(0.4, 1.4, n)
    return x4


@pytest.mark.parametrize("dim,n1,n2,L0", [(3, 200000, 30000, (120.0, 120.0, 120.0)),
                                         (2, 150000, 2000, (600.0, 600.0, 0.0))])
def test_noalloc_equals_allocating_twice_second_smaller(dim, n1, n2, L0):
    print(f"no-alloc find_pairs dim={dim} n1={n1} n2={n2}", flush=True)
    rng = np.random.default_rng(dim)
    stream = torch.cuda.current_stream().cuda_stream
    for n in (n1, n2):
        # the box shrinks with the count: both calls at the same density
        L = tuple(v * (n / n1) ** (1.0 / dim) for v in L0)
        x4 = _foam(rng, dim, n, L)
        cap = 32 * n
        ta, ia, ja = _run(x4, dim, L, 0.3, cap, False, stream)
        tb, ib, jb = _run(x4, dim, L, 0.3, cap, True, stream)
        assert ta == tb and n // 2 < ta <= cap
        assert np.array_equal(ia, ib) and np.array_equal(ja, jb)
        # grouped by i, ascending in j
        assert np.all(np.diff(ia) >= 0) and np.all(ia < ja)


def test_noalloc_refuses_counts_beyond_its_scratch():
    box = ca.Box.of(3, [0.0] * 3, [10.0] * 3)
    lib = ca.lib()
    rc = lib.cubble_find_pairs_noalloc(None, None, None, None, ca.FIND_PAIRS_NOALLOC_MAX + 1,
                                       ctypes.byref(box), 0.3, None, None, 0, None, None)
    assert rc == ca.E_CAPACITY
    assert "exceed" in lib.cubble_last_error(None).decode()
