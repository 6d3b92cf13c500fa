"""set_state checks caller neighbour lists on the host before any upload
(the kernels index with the entries): malformed lists are refused with
CUBBLE_E_INVALID and the device state stays as it was."""
import numpy as np
import pytest

import cubble_amd as ca
from parity import config

pytestmark = pytest.mark.gpu


def _mutations(st, n):
    """(name, state) pairs, each breaking one rule of the lists."""
    off, nbr = st["nbr_off"], st["nbr"]
    i = int(np.argmax(np.diff(off)))  # a bubble with the longest list (>= 2)
    a, b = int(off[i]), int(off[i + 1])
    assert b - a >= 2
    out = []

    def with_nbr(fn):
        s = dict(st)
        s["nbr"] = nbr.copy()
        fn(s["nbr"])
        return s

    out.append(("out of range", with_nbr(lambda x: x.__setitem__(a, n))))
    out.append(("negative", with_nbr(lambda x: x.__setitem__(a, -1))))
    out.append(("itself", with_nbr(lambda x: x.__setitem__(a, i))))
    out.append(("not ascending", with_nbr(lambda x: x.__setitem__(slice(a, a + 2),
                                                                   x[a:a + 2][::-1].copy()))))
    # replace one entry by another bubble that does not list i back
    others = np.setdiff1d(np.arange(n), np.concatenate([nbr[a:b], [i]]))
    k = next(int(j) for j in others if not np.any(nbr[off[j]:off[j + 1]] == i)
             and int(j) > int(nbr[b - 1]))
    out.append(("not symmetric", with_nbr(lambda x: x.__setitem__(b - 1, k))))
    s = dict(st)
    s["nbr_off"] = off.copy()
    s["nbr_off"][i + 1] = s["nbr_off"][i] - 1
    out.append(("offsets decrease", s))
    s = dict(st)
    s["nbr_off"] = off.copy() + 1
    out.append(("first offset", s))
    return out


@pytest.mark.parametrize("dim,n", [(2, 3000), (3, 4000)])
def test_set_state_rejects_malformed_lists(dim, n):
    print(f"set_state rejection dim={dim} n={n}", flush=True)
    sim = ca.Simulator(config(dim, n, seed=13))
    sim.run_steps(3, ca.IntegrationParams(True, False, True))
    st, sc = sim.get_state()
    for name, bad in _mutations(st, sc["count"]):
        with pytest.raises(ca.CubbleError) as e:
            sim.set_state(bad, sc)
        assert e.value.code == ca.E_INVALID, name
        now, nsc = sim.get_state()
        for key in ("x", "v", "idx", "nbr_off", "nbr"):
            assert np.array_equal(now[key], st[key]), (name, key)
        assert nsc["count"] == sc["count"]
    # the unmodified state is accepted and the foam steps on
    sim.set_state(st, sc)
    sim.run_steps(2, ca.IntegrationParams(True, False, True))
    sim.close()
