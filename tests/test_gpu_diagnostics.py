"""GPU parity of the diagnostics: postp's TKE budget (postp.f:9-76,
415-732) and speclegencalc (filtering.f:1110-1221) on the device, through
the C-ABI, against the CPU oracle, against closed forms (also at the full
256x128x(16x25) size) and, where it was built, the compiled reference."""
import numpy as np
import pytest

from helpers import (SOLUTION, STATE, case, energy_close, g3_params, kepe_closed_form, legendre_closed_form,
                     postp_closed_form, rel, routine_inputs)

pytestmark = pytest.mark.gpu

FIELD_TOL = 1e-10
HIST_TOL = 1e-9


@pytest.fixture(scope="module")
def rt():
    from iwrepo_amd import runtime
    runtime.lib()
    return runtime


@pytest.mark.parametrize("name", ["small", "c1"])
def test_postp_matches_oracle(rt, oracle_mod, name):
    """iw_postp on rough spectral fields, without and with the SGS tensors
    of filtering(1): the sixteen values and the state after its round
    trips."""
    P, flds = case(name)
    _, spec = routine_inputs(oracle_mod, P, flds)
    for isgs in (0, 1):
        print("postp", name, isgs, flush=True)
        o, g = oracle_mod.Oracle(P), rt.Solver(P)
        for q in range(4):
            o.set_field(q, spec[q])
            g.set_state(q, spec[q])
        if isgs:
            o.filtering(1)
            g.filtering(1)
        eo, eg = o.postp(), g.postp()
        print(np.c_[eo, eg], flush=True)
        assert energy_close(eg, eo, 0) < 1e-11, isgs
        for q in range(4):
            assert rel(g.get_state(q), o.field(q)) < 1e-13, (isgs, q)
        g.close()


@pytest.mark.parametrize("name", ["c1", "c2"])
def test_postp_closed_form_on_device(rt, oracle_mod, name):
    """The budget of fields known in closed form; c2 is the full
    256x128x(16x25) grid."""
    P, _ = case(name)
    o = oracle_mod.Oracle(P)
    z, x = o.table(7, P.nzloc * P.nsubd), o.table(8, P.nx + 2)
    flds, want = postp_closed_form(P, x, z)
    g = rt.Solver(P)
    spec = []
    for q in range(4):
        spec.append(g.horfft(flds[q], -1))
        g.set_state(q, spec[q])
    print("postp closed form", name, flush=True)
    got = g.postp()
    print(np.c_[got, want], flush=True)
    assert energy_close(got, want, 0) < 1e-11
    for q in range(4):
        assert rel(g.get_state(q), spec[q]) < 1e-13
    g.close()


def test_speclegen_matches_oracle_and_closed_form(rt, oracle_mod):
    P, flds = case("c1")
    _, spec = routine_inputs(oracle_mod, P, flds)
    o, g = oracle_mod.Oracle(P), rt.Solver(P)
    for q in range(4):
        o.set_field(q, spec[q])
        g.set_state(q, spec[q])
    for q in range(4):
        print("speclegen field", q, flush=True)
        a, b = o.speclegen(q), g.speclegen(q)
        assert np.abs(b - a).max() < 1e-12 * np.abs(a).max(), q
        assert rel(g.get_state(q), o.field(q)) < 1e-14, q
    x, zp = o.table(8, P.nx + 2), o.table(0, P.nzloc)
    modes = [(2 * ks + 1) % P.nzloc for ks in range(P.nsubd)]
    modes[-1] = P.nzloc - 1
    f, want = legendre_closed_form(P, zp, x, modes)
    g.set_state(0, g.horfft(f, -1))
    assert np.abs(g.speclegen(0) - want).max() < 1e-13
    g.close()


def test_step_loop_diagnostics_match_oracle(rt, oracle_mod):
    """iw_run with speclegencalc every 2 steps, the estimation filter every 3
    with postp (and the SGS tensors) every 3, plotf every 2: state, energy
    and spectra records."""
    P, flds = case("small")
    P.ifilter, P.nstepfilter, P.npostp, P.n3doutput, P.ncalclegspec = 1, 3, 3, 2, 2
    o, g = oracle_mod.Oracle(P), rt.Solver(P)
    o.set_initial(*flds)
    g.set_initial(*flds)
    print("diagnostics loop", flush=True)
    co, cg = o.run(6), g.run(6)
    so, sg = o.state(), g.state()
    for q in SOLUTION:
        assert rel(sg[q], so[q]) < (FIELD_TOL if q < 12 else HIST_TOL), STATE[q]
    assert np.array_equal(co, cg)
    eo, eg = o.energy_records(), g.energy_records()
    assert eo.shape == eg.shape == (2, 18)
    assert np.array_equal(eo[:, 0], eg[:, 0]) and np.abs(eo[:, 1] - eg[:, 1]).max() < 1e-15
    for r in range(2):
        assert energy_close(eg[r, 2:], eo[r, 2:], 0) < 1e-9, r
    lo, lg = o.speclegen_records(), g.speclegen_records()
    assert lo.shape == lg.shape == (3, 2 + 4 * P.nzloc * P.nsubd)
    assert np.array_equal(lo[:, 0], lg[:, 0])
    assert np.abs(lg[:, 2:] - lo[:, 2:]).max() < 1e-10 * np.abs(lo[:, 2:]).max()
    g.close()


def test_restart_postp_record(rt, oracle_mod, tmp_path):
    """iw_load_restart takes the restart field through plotf and postp
    (main.f:599-635): its energy record against the oracle's and, where
    built, the reference's energy.txt line; then 3 steps."""
    P, flds = case("small")
    o = oracle_mod.Oracle(P)
    o.set_initial(*flds)
    o.run(4)
    g = rt.Solver(P)
    g.set_initial(*flds)
    g.run(4)
    path = tmp_path / "dev.dat"
    g.save_restart(str(path))
    g.close()
    st = o.state()
    st[12:16] = 0.
    o2 = oracle_mod.Oracle(P)
    o2.set_restart(st, o.time(), 1.0)
    h = rt.Solver(P)
    h.load_restart(str(path), 1.0)
    eo, eh = o2.energy_records(), h.energy_records()
    assert eo.shape == eh.shape == (1, 18)
    assert energy_close(eh[0, 2:], eo[0, 2:], 0) < 1e-10
    if oracle_mod.ref_available("small") and oracle_mod.ref_has_postp("small"):
        _, _, _, en = oracle_mod.run_ref("small", P, flds, nsteps=0, start=path.read_bytes(), energy=True)
        assert energy_close(eh[0, 2:], en[0, 2:], 0) < 1e-10
    o2.run(3)
    h.run(3)
    sh, so = h.state(), o2.state()
    for q in SOLUTION:
        assert rel(sh[q], so[q]) < (FIELD_TOL if q < 12 else HIST_TOL), STATE[q]
    h.close()


@pytest.mark.parametrize("name", ["c1", "g3", "c2"])
def test_kepe_on_device(rt, oracle_mod, name):
    """iw_kepe against the oracle on the wave case and against the closed
    form (c2: the full 256x128x(16x25) grid; g3: 2-D, unequal subdomains);
    the state is left as it was."""
    if name == "g3":
        P, flds = g3_params()[0], None
    else:
        P, flds = case(name)
    o, g = oracle_mod.Oracle(P), rt.Solver(P)
    z, x = o.table(7, P.nzloc * P.nsubd), o.table(8, P.nx + 2)
    if name == "c1":
        print("kepe wave", name, flush=True)
        o.set_initial(*flds)
        g.set_initial(*flds)
        ko, kg = o.kepe(), g.kepe()
        assert np.all(np.abs(kg - ko) <= 1e-13 * np.abs(ko)), (ko, kg)
    cf, want = kepe_closed_form(P, x, z)
    spec = [g.horfft(f, -1) for f in cf]
    for q in range(4):
        g.set_state(q, spec[q])
    print("kepe closed form", name, flush=True)
    got = g.kepe()
    assert np.all(np.abs(got - want) <= 1e-12 * np.abs(want).max()), (got, want)
    for q in range(4):
        assert np.array_equal(g.get_state(q), spec[q])
    g.close()

