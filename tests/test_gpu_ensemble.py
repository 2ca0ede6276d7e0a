"""GPU checks of ensembles of uneven instances advanced by one captured HIP
graph (iw_ensemble_*) and of the transform block size the autotuner picks:
* every member of an ensemble run ends bit-identical to the same instance
  advanced alone by iw_run, with the graph at one chain, at several chains
  and without the graph;
* the transform results do not depend on the block size (64, 128, 256
  threads): whole steps bit-identical;
* the autotuner leaves the members' state as it found it;
* members against the CPU oracle (orders with and without the packed sweeps,
  the 5-member sizes of the batched workload's range at small scale)."""
import numpy as np
import pytest

from helpers import SOLUTION, STATE, rel
from iwrepo_amd.ic import initial_condition
from iwrepo_amd.params import make_params

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def rt():
    from iwrepo_amd import runtime
    runtime.lib()
    return runtime


SHAPES = [(32, 32, 2, 8), (64, 32, 3, 10), (32, 64, 2, 12), (32, 16, 4, 9), (16, 32, 2, 16), (64, 16, 2, 13)]


def members(rt):
    out = []
    for k, (nx, ny, ns, order) in enumerate(SHAPES):
        P = make_params(nx=nx, ny=ny, nsubd=ns, order=order, xlen=4 * np.pi, ylen=2 * np.pi, re=10 ** (3 + 0.4 * k),
                        pr=7.0, dt=1e-3, noslip=k % 2 == 0, imaxcalc=0)
        f = initial_condition(P, 0.02 + 0.015 * k, 1e-3, 50 + k)
        out.append((P, f))
    return out


def _solvers(rt, mem, ft=None):
    ss = []
    for P, f in mem:
        s = rt.Solver(P)
        if ft:
            s.set_fft_threads(ft)
        s.set_initial(*f)
        ss.append(s)
    return ss


def _states(ss):
    return [s.state() for s in ss]


def test_ensemble_equals_members_run_alone(rt):
    mem = members(rt)
    alone = _solvers(rt, mem)
    for s in alone:
        s.run(11, counts=False)
    want = _states(alone)
    for s in alone:
        s.close()
    for lanes in (1, 3, len(mem), 0):
        print("ensemble lanes", lanes, flush=True)
        ss = _solvers(rt, mem)
        e = rt.Ensemble(ss)
        e.set_lanes(lanes)
        e.run(5)   # start-up steps eagerly, one graph replay, the rest eagerly
        e.run(6)   # two replays of a fresh capture
        got = _states(ss)
        for i in range(len(mem)):
            for q in range(24):
                assert np.array_equal(got[i][q], want[i][q]), (lanes, i, STATE[q])
            assert ss[i].time() == pytest.approx(11 * mem[i][0].dt, rel=1e-13)
        e.close()
        for s in ss:
            s.close()


def test_transform_block_size_does_not_change_results(rt):
    mem = members(rt)[:4]
    ref = None
    for ft in (256, 128, 64):
        print("fft threads", ft, flush=True)
        ss = _solvers(rt, mem, ft)
        for s in ss:
            s.run(4, counts=False)
        st = _states(ss)
        if ref is None:
            ref = st
        else:
            for i in range(len(mem)):
                assert all(np.array_equal(st[i][q], ref[i][q]) for q in range(24)), (ft, i)
        for s in ss:
            s.close()
    s = rt.Solver(make_params(nx=1024, ny=8, nsubd=2, order=8))
    with pytest.raises(rt.IWError, match="do not fit"):
        s.set_fft_threads(64)  # 8 x 64 values hold one row of 512 complex values, not two
    s.close()


def test_autotuner_keeps_state_and_members_match_oracle(rt, oracle_mod):
    mem = members(rt)
    ss = _solvers(rt, mem)
    e = rt.Ensemble(ss)
    e.run(3)
    before = _states(ss)
    res = e.tune(6)
    lanes, ft = e.config()
    print("tuned", res, lanes, ft, flush=True)
    after = _states(ss)
    for i in range(len(mem)):
        assert all(np.array_equal(after[i][q], before[i][q]) for q in range(24)), i
    assert 0 < res["ms_per_step"] <= res["default_ms_per_step"]
    e.run(9)
    for i in (0, 1, 2):  # order 8 and 12 (packed sweeps), 10 (band solvers)
        P, f = mem[i]
        o = oracle_mod.Oracle(P)
        o.set_initial(*f)
        o.run(12)
        for q in SOLUTION:
            assert rel(ss[i].get_state(q), o.field(q)) < (1e-10 if q < 12 else 1e-9), (i, STATE[q])
    e.close()
    for s in ss:
        s.close()
