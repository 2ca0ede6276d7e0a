"""The second stream of the solver.  With the u/v/w groups (no-slip bottom,
several subdomains) iw_run puts the rho solve, rho's coupling recurrences and
rho's finish (coupling correction, Legendre filter) on a second stream beside
the pressure step and the u/v/w solves; with profiling on, every launch stays
on the solver stream and the coupling pass and finish take all four fields
in one launch each.  The two orders must give the same state bit for bit:
the work split between the streams touches disjoint arrays, and each field's
arithmetic is the same in either launch shape."""
import numpy as np
import pytest

from helpers import STATE, case

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def rt():
    from iwrepo_amd import runtime
    runtime.lib()
    return runtime


@pytest.mark.parametrize("name,steps", [("smallns", 12), ("c1ns", 30)])
@pytest.mark.parametrize("mixed", [False, True])
def test_two_streams_bitwise_one_stream(rt, name, steps, mixed):
    P, flds = case(name)
    P.imaxcalc = 0
    P.dtadapt = 0
    out = []
    for one_stream in (False, True):
        g = rt.Solver(P)
        g.set_mixed(mixed)
        g.set_profiling(one_stream)
        g.set_initial(*flds)
        g.run(steps, counts=False)
        out.append([g.get_state(q) for q in range(24)])
        g.close()
    for q in range(24):
        assert np.array_equal(out[0][q], out[1][q]), STATE[q]
    assert np.any(out[0][3] != 0.0)
