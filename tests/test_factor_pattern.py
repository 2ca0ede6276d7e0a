"""The band sweeps store and apply only the structurally nonzero entries of
each factor: for local row r of a subdomain of nl points, nl - r multipliers
below the diagonal, and above it r + 1 entries (systems factored without row
interchanges) or r + nl (the pressure systems, with interchanges), every
pressure interchange staying within the L column.  The packing checks this on
the device and falls back to the band solvers if it fails; these tests pin
the claim on the oracle's own system matrices.

Each matrix is rebuilt from the oracle's penalty solve (penalty.f:520-711): a
unit right-hand side gives a column of (-eps A)^-1, a row scaling of A^-1
that keeps A's zero pattern.  The rebuilt A is factored with LAPACK's DGBTRF,
the routine the reference calls (bandludcmp.f), with kl = ku = nl.  A planted
coupling outside the penalty pattern must break the check."""
import numpy as np
import pytest
from scipy.linalg import lapack

from helpers import case


def _matrix(o, nz, eps, bot, top):
    cols = []
    for k in range(nz):
        e = np.zeros(nz)
        e[k] = 1.0
        fr, _ = o.penalty_solve(eps, bot, top, e, np.zeros(nz))
        cols.append(fr)
    a = np.linalg.inv(np.array(cols).T)
    # inversion noise where A is structurally zero
    a[np.abs(a) < 1e-11 * np.abs(a).max(axis=1, keepdims=True)] = 0.0
    return a


def _pattern_holds(a, nl, pivoting):
    n = a.shape[0]
    kl = ku = nl
    kv = kl + ku
    ab = np.zeros((2 * kl + ku + 1, n))
    for c in range(n):
        for r in range(max(0, c - ku), min(n, c + kl + 1)):
            ab[kv + r - c, c] = a[r, c]
    lu, piv, info = lapack.dgbtrf(ab, kl, ku)
    assert info == 0
    if not pivoting and np.any(piv != np.arange(n)):
        return None  # such a system goes to the band solvers anyway
    for c in range(n):
        r = c % nl
        ef = nl - r
        eb = r + (nl if pivoting else 1)
        for t in range(ef + 1, kl + 1):
            if c + t < n and lu[kv + t, c] != 0.0:
                return False
        for t in range(eb + 1, kv + 1):
            if c - t >= 0 and lu[kv - t, c] != 0.0:
                return False
        if pivoting and not 0 <= piv[c] - c <= ef:
            return False
    return True


def _systems(P):
    """(eps, bottom bc, top bc, interchanges) of pressure and viscous systems
    over a spread of k^2 (calcuvwt.f:105-141, 380-520; 0 = Dirichlet,
    1 = Neumann)."""
    facvel = 11.0 / 6.0 / (P.xnu * P.dt)
    fact = 11.0 / 6.0 / (P.xkappa * P.dt)
    ubot = 0 if P.bc_uv_bottom == 1 else 1
    out = []
    for k2 in (1.0, 17.0, 400.0, 5000.0):
        out.append((1.0 / k2, 1, 1, True))
        out.append((1.0 / (k2 + facvel), ubot, 1, False))
        out.append((1.0 / (k2 + facvel), 0, 0, False))
        out.append((1.0 / (k2 + fact), 1, 1, False))
    return out


@pytest.mark.parametrize("name", ["small", "smallns", "c1"])
def test_factor_pattern_on_oracle_matrices(oracle_mod, name):
    P, _ = case(name)
    o = oracle_mod.Oracle(P)
    nz, nl = P.nz, P.nzloc
    checked = 0
    for eps, bot, top, pivoting in _systems(P):
        a = _matrix(o, nz, eps, bot, top)
        ok = _pattern_holds(a, nl, pivoting)
        if ok is None:
            continue
        assert ok, (name, eps, bot, top)
        checked += 1
    assert checked >= 8


def test_factor_pattern_rejects_a_coupling_outside_it(oracle_mod):
    """The second row of subdomain 2 coupled to the second point of subdomain
    1 (inside the band, outside the interface rows) breaks the L pattern."""
    P, _ = case("small")
    o = oracle_mod.Oracle(P)
    nz, nl = P.nz, P.nzloc
    a = _matrix(o, nz, 1.0 / (1.0 + 11.0 / 6.0 / (P.xnu * P.dt)), 1, 1)
    assert _pattern_holds(a, nl, False)
    a[2 * nl + 1, nl + 1] = 1e-6 * np.abs(a).max()  # small: no interchange
    assert _pattern_holds(a, nl, False) is False
