"""The mixed-precision path (iw_set_mixed): FP32 physical-space working set in
convec, with the z contraction on FP32 matrix cores, against the FP64 oracle.
The bound is the one the mixed path is built for: kinetic and potential
energy within 1e-4 relative of the FP64 run (here far tighter); every field
within 1e-5 relative after the steps below.  It must differ from the plain
path (FP32 rounding is visible), and the plain path must stay untouched."""
import numpy as np
import pytest

from helpers import STATE, case, rel

pytestmark = pytest.mark.gpu

ENERGY_TOL = 1e-4   # the mixed path's bound on KE and PE
FIELD_TOL = 1e-5    # per field, relative to the velocity scale for u, v, w


def _field_errs(g, o, qs):
    """Error of each array: u, v, w (and their histories) relative to the
    norm of the whole velocity of that level (v is noise-level in these
    waves, so its own norm is no scale for FP32 rounding), rho relative to
    its own norm."""
    out = {}
    for q in qs:
        a, b = g.get_state(q), o.field(q)
        lev = q - q % 4
        if q % 4 < 3:
            scale = np.sqrt(sum(np.linalg.norm(o.field(lev + c)) ** 2 for c in range(3)))
        else:
            scale = np.linalg.norm(b)
        out[STATE[q]] = float(np.linalg.norm(a - b) / scale)
    return out


@pytest.fixture(scope="module")
def rt():
    from iwrepo_amd import runtime
    runtime.lib()
    return runtime


def _params(name):
    P, flds = case(name)
    P.imaxcalc = 0
    P.dtadapt = 0
    return P, flds


@pytest.mark.parametrize("name,steps", [("small", 20), ("small13", 20), ("small21", 20), ("c1", 200),
                                        ("c1ns", 200), ("r48", 20), ("r96", 20)])
def test_mixed_path_against_fp64_oracle(rt, oracle_mod, name, steps):
    P, flds = _params(name)
    print("mixed", name, steps, flush=True)
    o = oracle_mod.Oracle(P)
    o.set_initial(*flds)
    g = rt.Solver(P)
    g.set_mixed(True)
    g.set_initial(*flds)
    eo, eg = [], []
    for n in range(steps // 10):
        o.run(10)
        g.run(10, counts=False)
        eo.append(o.kepe())
        eg.append(g.kepe())
    eo, eg = np.array(eo), np.array(eg)
    errs = _field_errs(g, o, range(12))
    ke = [rel(eg[:, q], eo[:, q]) for q in range(4)]
    print("fields", errs, "KE/PE series", ke, flush=True)
    for q in range(4):
        assert ke[q] < ENERGY_TOL, q
    for q in range(12):
        assert errs[STATE[q]] < FIELD_TOL, STATE[q]
    assert max(errs[STATE[q]] for q in range(4)) > 1e-13  # FP32 rounding is there
    g.close()


def test_mixed_path_with_matrix_core_contraction(rt, oracle_mod, monkeypatch):
    """IW_MIXED_MFMA=1 at creation: the z contraction on FP32 matrix cores
    (k_dpdz_conv_mfma) and the products in their own pass instead of the
    fused VALU pass; the same bounds against the FP64 oracle, and the two
    within FP32 rounding of each other (the matrix cores accumulate the same
    products in the same order, so they often agree bit for bit)."""
    P, flds = _params("c1")
    o = oracle_mod.Oracle(P)
    o.set_initial(*flds)
    o.run(50)
    out = {}
    for mfma in ("1", "0"):
        monkeypatch.setenv("IW_MIXED_MFMA", mfma)
        g = rt.Solver(P)
        g.set_mixed(True)
        g.set_initial(*flds)
        g.run(50, counts=False)
        errs = _field_errs(g, o, range(12))
        print("mfma", mfma, errs, flush=True)
        for q in range(12):
            assert errs[STATE[q]] < FIELD_TOL, (mfma, STATE[q])
        out[mfma] = g.get_state(0)
        g.close()
    assert rel(out["1"], out["0"]) < 1e-5


def test_mixed_run_split_over_calls_is_bitwise_one_call(rt):
    """The mixed path keeps the right-hand-side histories S, Sn, Snm in FP32
    during a run and in the FP64 state between runs: 12 steps as 5 + 7 calls
    equal 12 steps in one call bit for bit (all 24 state arrays), and the
    histories between runs are FP32 values."""
    P, flds = _params("small")
    out = []
    for chunks in ((12,), (5, 7)):
        g = rt.Solver(P)
        g.set_mixed(True)
        g.set_initial(*flds)
        for n in chunks:
            g.run(n, counts=False)
        out.append([g.get_state(q) for q in range(24)])
        g.close()
    for q in range(24):
        assert np.array_equal(out[0][q], out[1][q]), STATE[q]
    for q in range(16, 24):
        a = out[0][q]
        assert np.array_equal(a.astype(np.float32).astype(np.float64), a), STATE[q]
        assert np.any(a != 0.0), STATE[q]


def test_mixed_path_switches_back_to_plain(rt, oracle_mod):
    """iw_set_mixed(0) restores the plain path exactly (1e-10 vs the oracle),
    and the mixed path refuses what it does not cover."""
    P, flds = _params("small")
    g = rt.Solver(P)
    g.set_mixed(True)
    g.set_mixed(False)
    g.set_initial(*flds)
    g.run(10)
    o = oracle_mod.Oracle(P)
    o.set_initial(*flds)
    o.run(10)
    for q in range(12):
        assert rel(g.get_state(q), o.field(q)) < 1e-10, STATE[q]
    g.close()
    P2, _ = case("small")  # mean-wave / shear terms: FP64 path only
    P2.fcshear = 0.5
    h = rt.Solver(P2)
    with pytest.raises(rt.IWError, match="mean-wave"):
        h.set_mixed(True)
    h.close()


def test_mixed_path_maxcalc_records_and_adaptive_dt(rt, oracle_mod):
    """maxcalc on the mixed path: the extrema and CFL maxima with their
    positions (from the FP64 values of the inverse transform), the umaxt
    chain on FP32 u, and the adaptive step: c1 with dt changes at steps 22
    and 43 as on the FP64 path."""
    P, flds = case("c1")
    P.dtadapt = 1
    o, g = oracle_mod.Oracle(P), rt.Solver(P)
    g.set_mixed(True)
    o.set_initial(*flds)
    g.set_initial(*flds)
    o.run(50)
    g.run(50, counts=False)
    ro, rg = o.maxcalc_records(), g.maxcalc_records(50)
    assert [int(r[0]) for r in rg if r[37] == 1] == [int(r[0]) for r in ro if r[37] == 1] == [22, 43]
    assert np.array_equal(rg[:, 35], ro[:, 35])  # dt after every call
    for c in (2, 3, 6, 7, 10, 11, 14, 15, 22, 26, 30):  # extrema and CFL maxima
        assert np.abs(rg[:, c] - ro[:, c]).max() <= 1e-5 * np.abs(ro[:, c]).max(), c
    errs = _field_errs(g, o, range(4))
    print(errs, flush=True)
    for q in range(4):
        assert errs[STATE[q]] < FIELD_TOL, q
    g.close()
