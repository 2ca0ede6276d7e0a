"""The allocation-free KE/PE entry point (iw_kepe_noalloc, a static device
array as its scratch) against iw_kepe, called through the C-ABI: two calls in
a row, the second on a smaller instance, each bit-identical to iw_kepe on the
same state; and the size limit of the static array."""
import ctypes

import numpy as np
import pytest

from helpers import case

pytestmark = pytest.mark.gpu


def _noalloc(rt, g):
    out = np.full(4, np.nan)
    rc = rt.lib().iw_kepe_noalloc(g.h, out.ctypes.data_as(ctypes.c_void_p))
    if rc != 0:
        raise rt.IWError(rt.lib().iw_last_error().decode())
    return out


def test_kepe_noalloc_matches_kepe_on_two_calls_in_a_row():
    from iwrepo_amd import runtime as rt
    res = []
    for name in ("c1", "small"):  # the second call on a smaller instance
        P, flds = case(name)
        print("kepe_noalloc", name, flush=True)
        g = rt.Solver(P)
        g.set_initial(*flds)
        g.run(3)
        a = _noalloc(rt, g)
        b = g.kepe()
        assert np.all(np.isfinite(a))
        assert np.array_equal(a, b), (name, a, b)
        res.append(a)
        g.close()
    assert not np.array_equal(res[0], res[1])
