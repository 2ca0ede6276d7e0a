"""GPU checks of the Fortran drop-in boundary and of restart files.

The shim driver (oracle/_shim/<cfg>/shimdrv) is the test harness
oracle/ref_harness.f, which replays main.f:653-1066, linked with the product's
Fortran shims (iwrepo_amd/fortran/iwshim.F) and libiwgpu.so only: every
hot-path routine it calls (horfft, dpdz, convec, calcuvwt, interface_avg3d,
filter_3dlegenfourier, filtering, maxcalc, postp, speclegencalc, the multistep
functions) runs on the GPU, and no object of the reference is linked.  Its
state is compared with the C oracle (pinned to the compiled reference on the
same cases by tests/test_oracle.py), and what the reference prints (maxcalc's
lines and dt log, the energy rows, the Legendre spectra files) with the
reference's own outputs recorded by tests/golden/make_shim_fixtures.py.

Restart files: the host reader and writer of the output.f / restart.f record
layout are checked bit for bit against files the reference wrote and read
(tests/test_oracle.py); here the device reads and writes them and continues
as the oracle restarting from the same state."""
import json
import os

import numpy as np
import pytest

from helpers import GOLDEN, SOLUTION, STATE, case, energy_close, meanwave_fields, rel, routine_inputs

pytestmark = pytest.mark.gpu

FIELD_TOL = 1e-10
HIST_TOL = 1e-9


@pytest.fixture(scope="module")
def rt():
    from iwrepo_amd import runtime
    runtime.lib()
    return runtime


@pytest.fixture(scope="module")
def shim_gold():
    with open(os.path.join(GOLDEN, "shim_outputs.json")) as f:
        return json.load(f)


def _need(O, name):
    if not O.shim_available(name):
        pytest.skip("oracle/_shim/%s/shimdrv not built (built by oracle/build_ref.py with the "
                    "reference's COMMON headers)" % name)


def _cmp_state(ss, so):
    errs = {STATE[q]: rel(ss[q], so[q]) for q in SOLUTION}
    print(errs, flush=True)
    for q in SOLUTION:
        assert errs[STATE[q]] < (FIELD_TOL if q < 12 else HIST_TOL), STATE[q]
    return errs


def _counts(diag):
    return np.array([[int(v) for v in l.split()[1:]] for l in diag if l.strip()])


def _cmp_printed(ms, gold):
    """maxcalc's printed outputs of the shim against the reference's."""
    dr, ds = np.array(gold["dtlog"]), np.array(ms["dtlog"])
    assert np.array_equal(dr, ds)  # t, dt, cdt, ivardt, idtstep after every maxcalc
    cr, cs = np.array(gold["cflz"]), np.array(ms["cflz"])
    assert np.array_equal(cr[:, 0], cs[:, 0])
    assert rel(cs[:, 1], cr[:, 1]) < 1e-12
    assert gold["tloc"] == ms["tloc"]
    f50r, f50s = np.array(gold["fort50"]), np.array(ms["fort50"])
    assert f50r.shape == f50s.shape
    assert (np.abs(f50r - f50s) <= 1.1e-5 * np.abs(f50r)).all()  # at most the last printed digit


@pytest.mark.parametrize("name,steps", [("small", 10), ("c1", 40), ("c1ns", 40)])
def test_fortran_shims_under_the_step_loop(rt, oracle_mod, name, steps):
    """The step loop of main.f calling the shims: state, time and the
    interface-average counts of every step against the oracle."""
    _need(oracle_mod, name)
    P, flds = case(name)
    print("shim driver", name, steps, flush=True)
    ss, ts, ds = oracle_mod.run_ref(name, P, flds, nsteps=steps, exe="shimdrv")
    o = oracle_mod.Oracle(P)
    o.set_initial(*flds)
    co = o.run(steps)
    _cmp_state(ss, o.state())
    assert abs(ts - o.time()) < 1e-15
    assert np.array_equal(_counts(ds), co)


@pytest.mark.parametrize("name", ["small", "c1"])
def test_fortran_shim_routines(rt, oracle_mod, name):
    """horfft, dpdz, filter_3dlegenfourier and interface_avg3d called one at
    a time through the shims, against the oracle's routines."""
    _need(oracle_mod, name)
    P, flds = case(name)
    o, spec = routine_inputs(oracle_mod, P, flds)
    tol = {1: 1e-14, 2: 1e-13, 3: 1e-12, 4: 1e-13}
    for mode in (1, 2, 3, 4):
        print("shim routine", name, "mode", mode, flush=True)
        inp = flds if mode == 1 else spec
        ss, _, ds = oracle_mod.run_ref(name, P, inp, nsteps=0, mode=mode, exe="shimdrv")
        counts = []
        for q in range(4):
            if mode == 1:
                want = o.horfft(inp[q], -1)
            elif mode == 2:
                want = o.dpdz(inp[q])
            elif mode == 3:
                want = o.filter3d(inp[q])
            else:
                want, c = o.interface_avg(inp[q])
                counts.append(c)
            assert rel(ss[q], want) < tol[mode], (mode, q)
        if mode == 4:
            assert [int(x) for x in ds[0].split()] == counts


@pytest.mark.parametrize("name", ["small", "c1", "c2"])
def test_penalty_solve_on_the_device_matches_oracle(rt, oracle_mod, name):
    """iw_penalty_solve (helmv_penalty on one column pair: boundary-row terms,
    assembly, banded LU with interchanges and SGBTRS on the device) and
    iw_penalty_params against the oracle's penalty solve, for the four
    (epsilon, boundary types and values) of harness mode 8 on the z grids of
    the small case, C1 and 256x128x(16x25), within the parity bar (1e-10):
    the Neumann-Neumann system of the 400-point grid is ill-conditioned
    enough that the rounding of the two factorisations (FMA on the device,
    none in the oracle) shows at 7e-12."""
    from helpers import penalty_cases
    P, _ = case(name)
    o, g = oracle_mod.Oracle(P), rt.Solver(P)
    rng = np.random.default_rng(5)
    for c, (eps, bot, top, bc) in enumerate(penalty_cases(P)):
        print("penalty solve", name, "case", c, flush=True)
        rr, ri = rng.standard_normal(P.nz), rng.standard_normal(P.nz)
        fr, fi = o.penalty_solve(eps, bot, top, rr, ri, bc)
        gr, gi = g.penalty_solve(eps, bot, top, (-rr) * eps, (-ri) * eps, bc)
        assert rel(gr, fr) < 1e-10 and rel(gi, fi) < 1e-10, c
        al, be, ga, de, t1, t2 = g.penalty_params(eps, bot, top)
        assert be[0] == (0.0 if bot == 0 else 1.0) and de[-1] == (0.0 if top == 0 else 1.0)
        assert all(x > 0 for x in t1) and all(x > 0 for x in t2)
    with pytest.raises(rt.IWError, match="boundary types"):
        g.penalty_solve(0.01, 2, 0, np.zeros(P.nz), np.zeros(P.nz))
    g.close()


def test_fortran_penalty_shims_as_testhelm_calls_them(rt, oracle_mod):
    """penalty_setup and penalty_main through the Fortran shims in the shim
    driver (no reference object linked), called as testhelm.f:60-100 calls
    them (harness mode 8), against the oracle: solutions within 1e-12, the
    right-hand sides left times -epsilon in place."""
    from helpers import parse_penalty_diag, penalty_cases, penalty_columns
    _need(oracle_mod, "small")
    for name in ("small", "c1ns"):
        if not oracle_mod.shim_available(name) and name != "small":
            continue
        P, flds = case(name)
        o = oracle_mod.Oracle(P)
        print("penalty shims", name, flush=True)
        _, _, diag = oracle_mod.run_ref(name, P, flds, nsteps=0, mode=8, exe="shimdrv")
        got = parse_penalty_diag(diag, P.nz)
        for c, (eps, bot, top, bc) in enumerate(penalty_cases(P)):
            rr, ri = penalty_columns(flds, c)
            fr, fi = o.penalty_solve(eps, bot, top, rr, ri, bc)
            assert rel(got[c, :, 0], fr) < 1e-12 and rel(got[c, :, 1], fi) < 1e-12, (name, c)
            assert np.array_equal(got[c, :, 2], -rr * eps) and np.array_equal(got[c, :, 3], -ri * eps)


def test_fortran_shim_filtering(rt, oracle_mod):
    """filtering(isgscalc) through the shim: the fields and the /comsgs/
    tensors against the oracle."""
    _need(oracle_mod, "small")
    P, flds = case("small")
    _, spec = routine_inputs(oracle_mod, P, flds)
    for isgs in (0, 1):
        print("shim filtering", isgs, flush=True)
        ss, _, _ = oracle_mod.run_ref("small", P, spec, nsteps=0, mode=5, isgs=isgs, exe="shimdrv")
        o = oracle_mod.Oracle(P)
        for q in range(4):
            o.set_field(q, spec[q])
        o.filtering(isgs)
        for q in range(4):
            assert rel(ss[q], o.field(q)) < 1e-13, (isgs, q)
        if isgs:
            for q in range(9):
                assert rel(ss[24 + q][:, :, :P.nx], o.sgs(q)[:, :, :P.nx]) < 1e-13, q


def test_fortran_maxcalc_shim_under_the_step_loop(rt, oracle_mod, shim_gold):
    """maxcalc through the shim (device extrema, round trip, adaptive step
    with Tso kept) under the step loop with every hot-path routine shimmed:
    50 steps of c1 with two dt changes.  The state against the oracle, the
    printed lines and dt log against the reference's."""
    _need(oracle_mod, "c1")
    P, flds = case("c1")
    print("shim maxcalc c1 50", flush=True)
    ss, ts, _, ms = oracle_mod.run_ref("c1", P, flds, nsteps=50, exe="shimdrv", maxcalc=True)
    _cmp_printed(ms, shim_gold["maxcalc_c1_50"])
    P.dtadapt = 1
    o = oracle_mod.Oracle(P)
    o.set_initial(*flds)
    o.run(50)
    _cmp_state(ss, o.state())
    assert abs(ts - o.time()) < 1e-15


def test_fortran_shims_with_meanwave_and_shear(rt, oracle_mod, shim_gold):
    """The shimmed step loop (maxcalc included) with the shear current and
    the mean-wave fields on: 20 steps of c1."""
    _need(oracle_mod, "c1")
    P, flds = case("c1")
    P.fcshear = 0.5
    o = oracle_mod.Oracle(P)
    mw = meanwave_fields(P, o.table(7, P.nzloc * P.nsubd))
    print("shim mean-wave c1 20", flush=True)
    ss, ts, _, ms = oracle_mod.run_ref("c1", P, flds, nsteps=20, exe="shimdrv", maxcalc=True, meanwave=mw)
    _cmp_printed(ms, shim_gold["meanwave_c1_20"])
    P.dtadapt = 1
    o = oracle_mod.Oracle(P)
    o.set_meanwave(mw)
    o.set_initial(*flds)
    o.run(20)
    _cmp_state(ss, o.state())


def test_fortran_postp_and_speclegen_shims(rt, oracle_mod, shim_gold):
    """postp and speclegencalc through the shims inside the step loop
    (speclegencalc every 2 steps, the filter and postp every 3, plotf every
    2): the state against the oracle, the /energetics/ values and the
    speclegen*.dat files against the reference's; then postp alone after
    filtering(1) against the oracle."""
    _need(oracle_mod, "small")
    P, flds = case("small")
    P.ifilter, P.nstepfilter, P.npostp, P.n3doutput, P.ncalclegspec = 1, 3, 3, 2, 2
    print("shim diagnostics loop", flush=True)
    ss, ts, _, es, ls = oracle_mod.run_ref("small", P, flds, nsteps=6, energy=True, speclegen=True,
                                           exe="shimdrv")
    o = oracle_mod.Oracle(P)
    o.set_initial(*flds)
    o.run(6)
    _cmp_state(ss, o.state())
    gold = shim_gold["diag_small_6"]
    er = np.array(gold["energy"])
    assert er.shape == es.shape == (2, 18)
    for r in range(2):
        assert energy_close(es[r, 2:], er[r, 2:], 0) < 1e-9, r
    lr = {tuple(k.split("|")): np.array(v) for k, v in gold["speclegen"].items()}
    assert sorted(lr) == sorted(ls) and len(lr) == 12
    for k in lr:
        a, b = ls[k], lr[k]
        # both rounded to E12.6: at most one unit of the sixth digit apart
        assert np.all(np.abs(a - b) <= 1.1e-5 * np.abs(b) + 1e-10 * np.abs(b).max()), k
    P, flds = case("small")
    _, spec = routine_inputs(oracle_mod, P, flds)
    ss, _, _, es = oracle_mod.run_ref("small", P, spec, nsteps=0, mode=6, isgs=1, energy=True, exe="shimdrv")
    o = oracle_mod.Oracle(P)
    for q in range(4):
        o.set_field(q, spec[q])
    o.filtering(1)
    eo = o.postp()
    assert energy_close(es[0, 2:], eo, 0) < 1e-11
    for q in range(4):
        assert rel(ss[q], o.field(q)) < 1e-13, q


# ------------------------------------------------------------ restart files


def _oracle_restart(O, P, st, t, cdt, nsteps, mw=None):
    o = O.Oracle(P)
    if mw is not None:
        o.set_meanwave(mw)
    st0 = st.copy()
    st0[12:16] = 0.
    o.set_restart(st0, t, cdt)
    c = o.run(nsteps)
    return o, c


@pytest.mark.parametrize("cdt", [1.0, 2.0])
def test_device_restart_from_file(rt, oracle_mod, tmp_path, cdt):
    """iw_load_restart of a start.dat in the output.f record layout (written
    by the host writer from the oracle's state after 5 steps), then 5 steps
    with the restart coefficients (main.f:619-634, 703-734, 818-825),
    against the oracle restarting from the same state."""
    P, flds = case("small")
    o = oracle_mod.Oracle(P)
    o.set_initial(*flds)
    o.run(5)
    st, t = o.state(), o.time()
    path = tmp_path / "start.dat"
    rt.write_restart_file(str(path), P, t, st)
    ref, cref = _oracle_restart(oracle_mod, P, st, t, cdt, 5)
    g = rt.Solver(P)
    g.load_restart(str(path), cdt)
    cg = g.run(5)
    sg = g.state()
    tg = g.time()
    g.close()
    _cmp_state(sg, ref.state())
    assert abs(tg - ref.time()) < 1e-15
    assert np.array_equal(cg, cref)


def test_device_restart_file_read_back(rt, oracle_mod, tmp_path):
    """iw_save_restart after 5 device steps: the host reader returns the
    device state bit for bit, and the oracle restarting from that state
    agrees with the device restarting from the file."""
    P, flds = case("small")
    g = rt.Solver(P)
    g.set_initial(*flds)
    g.run(5)
    path = tmp_path / "dev.dat"
    g.save_restart(str(path))
    sdev, tdev = g.state(), g.time()
    g.close()
    st, t, umax = rt.read_restart_file(str(path), P)
    assert t == tdev and umax == P.umax
    for q in list(range(12)) + list(range(16, 24)):
        assert np.array_equal(st[q], sdev[q]), STATE[q]
    ref, _ = _oracle_restart(oracle_mod, P, st, t, 1.0, 5)
    h = rt.Solver(P)
    h.load_restart(str(path), 1.0)
    h.run(5)
    sh, th = h.state(), h.time()
    h.close()
    assert abs(th - ref.time()) < 1e-15
    _cmp_state(sh, ref.state())


def test_restart_with_meanwave_fields(rt, oracle_mod, tmp_path):
    """A restart file that holds mean-wave fields (output.f:55-59): the
    device writes them with the state, reads them back and continues as
    the oracle restarting from the same state and fields."""
    P, flds = case("small")
    P.fcshear = 0.5
    o = oracle_mod.Oracle(P)
    mw = meanwave_fields(P, o.table(7, P.nzloc * P.nsubd))
    g = rt.Solver(P)
    g.set_meanwave(mw)
    g.set_initial(*flds)
    g.run(3, counts=False)
    path = tmp_path / "mw.dat"
    g.save_restart(str(path))
    st, t = g.state(), g.time()
    g.close()
    h = rt.Solver(P)
    h.load_restart(str(path), 1.0)
    assert np.array_equal(h.meanwave(), mw)
    h.run(4, counts=False)
    ref, _ = _oracle_restart(oracle_mod, P, st, t, 1.0, 4, mw)
    _cmp_state(h.state(), ref.state())
    h.close()
