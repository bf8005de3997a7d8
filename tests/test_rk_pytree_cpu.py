"""Host logic of rk3_step's pytree handling (no GPU): leaves come out in
order and the structure (tuples, lists, namedtuples, dicts) is rebuilt."""
import collections

import numpy as np

from famrdp_jax.time.rk import _flatten


def test_flatten_rebuild_round_trip():
    P = collections.namedtuple("P", "a b")
    x, y, z, w = (np.full(2, v) for v in (1.0, 2.0, 3.0, 4.0))
    tree = {"s": (x, [y, P(z, w)]), "t": x}
    leaves, rebuild = _flatten(tree)
    assert [float(v[0]) for v in leaves] == [1.0, 2.0, 3.0, 4.0, 1.0]
    out = rebuild([v * 10 for v in leaves])
    assert isinstance(out, dict) and isinstance(out["s"], tuple) and isinstance(out["s"][1], list)
    assert isinstance(out["s"][1][1], P)
    assert float(out["s"][1][1].b[0]) == 40.0 and float(out["t"][0]) == 10.0


def test_single_array_is_one_leaf():
    a = np.arange(3.0)
    leaves, rebuild = _flatten(a)
    assert len(leaves) == 1 and leaves[0] is a and rebuild([a]) is a
