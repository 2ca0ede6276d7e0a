"""CPU pins of the learned-closure oracle (oracle/mlp_oracle.py): the exact
gradient by central differences, BF16 rounding known answers, and the
coarse-graining against what the mathematics fixes (polynomials within the
coarse order and Fourier content within the coarse band are reproduced, the
next Legendre polynomial is projected out)."""
import numpy as np
from numpy.polynomial import legendre as L

import mlp_oracle as M
from iwrepo_amd.params import make_params


def _random_params(seed):
    rng = np.random.default_rng(seed)
    p = np.zeros(M.NPARAM)
    Ws, _ = M.unpack(p)
    p[M.OW1:M.OB1] = (rng.standard_normal((M.NH, M.NINP)) * 0.4 * (np.arange(M.NINP) < M.NIN)).ravel()
    p[M.OW2:M.OB2] = rng.standard_normal(M.NH * M.NH) * 0.15
    p[M.OW3:M.OB3] = rng.standard_normal(M.NH * M.NH) * 0.15
    p[M.OW4:M.OB4] = (rng.standard_normal((M.NOUTP, M.NH)) * 0.2 * (np.arange(M.NOUTP) < M.NOUT)[:, None]).ravel()
    for a, b in ((M.OB1, M.OW2), (M.OB2, M.OW3), (M.OB3, M.OW4)):
        p[a:b] = rng.standard_normal(b - a) * 0.1
    p[M.OB4:M.OB4 + M.NOUT] = rng.standard_normal(M.NOUT) * 0.1
    norm = np.concatenate([rng.standard_normal(M.NIN) * 0.1, 0.5 + rng.random(M.NIN),
                           rng.standard_normal(M.NOUT) * 0.1, 0.5 + rng.random(M.NOUT)])
    return p, norm, rng


def test_gradient_matches_central_differences():
    p, norm, rng = _random_params(3)
    X = rng.standard_normal((40, M.NIN))
    T = rng.standard_normal((40, M.NOUT))
    loss, g = M.loss_grad(p, norm, X, T, emulate=False)
    idx = np.concatenate([rng.choice(np.arange(a, b), 6, replace=False) for a, b in
                          ((M.OW1, M.OW1 + M.NIN), (M.OB1, M.OW2), (M.OW2, M.OB2), (M.OB2, M.OW3),
                           (M.OW3, M.OB3), (M.OB3, M.OW4), (M.OW4, M.OW4 + M.NH * M.NOUT),
                           (M.OB4, M.OB4 + M.NOUT))])
    h = 1e-6
    for i in idx:
        e = np.zeros(M.NPARAM)
        e[i] = h
        fd = (M.loss_grad(p + e, norm, X, T, emulate=False)[0] - M.loss_grad(p - e, norm, X, T, emulate=False)[0]) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 * max(abs(fd), 1e-3), (i, fd, g[i])
    # padded entries carry no gradient
    W1 = g[M.OW1:M.OB1].reshape(M.NH, M.NINP)
    W4 = g[M.OW4:M.OB4].reshape(M.NOUTP, M.NH)
    assert np.all(W1[:, M.NIN:] == 0) and np.all(W4[M.NOUT:] == 0) and np.all(g[M.OB4 + M.NOUT:] == 0)


def test_bf16_rounding_known_answers():
    x = np.array([1.0, 1 + 2 ** -8, 1 + 2 ** -8 + 2 ** -20, 1 + 3 * 2 ** -8, -2.5, 3.0e-5], dtype=np.float32)
    want = np.array([1.0, 1.0, 1 + 2 ** -7, 1 + 2 ** -6, -2.5, np.float32(3.0e-5)], dtype=np.float64)
    got = M.bf16(x).astype(np.float64)
    assert np.array_equal(got[:5], want[:5])
    assert abs(got[5] - 3.0e-5) <= 2 ** -8 * 3.0e-5  # 8 significant bits


def test_zprojection_reproduces_low_degree_and_removes_the_next():
    of, oc, r = 16, 8, 2
    Mz = M.zprojection(of, oc, r)
    xf, xc = M.gll_nodes(of), M.gll_nodes(oc)
    xi = np.concatenate([-1 + (2 * q + 1 + xf) / r for q in range(r)])
    rng = np.random.default_rng(0)
    c = rng.standard_normal(oc + 1)
    assert np.abs(Mz @ L.legval(xi, c) - L.legval(xc, c)).max() < 1e-12
    e = np.zeros(oc + 2)
    e[-1] = 1.0  # P_{oc+1} is orthogonal to every polynomial of degree <= oc
    assert np.abs(Mz @ L.legval(xi, e)).max() < 1e-12


def test_coarsen_keeps_the_coarse_band_and_drops_the_rest():
    Pf = make_params(nx=32, ny=16, nsubd=4, order=16)
    Pc = make_params(nx=16, ny=8, nsubd=2, order=8)
    rng = np.random.default_rng(1)
    Ff = np.zeros((Pf.nzloc * Pf.nsubd, Pf.ny, Pf.nx + 2))
    Fc_want = np.zeros((Pc.nzloc * Pc.nsubd, Pc.ny, Pc.nx + 2))
    xf = M.gll_nodes(Pf.nzloc - 1)
    xc = M.gll_nodes(Pc.nzloc - 1)
    # modes inside the coarse band with z content of degree <= 8 on each coarse subdomain
    for (ii, j) in [(0, 0), (3, 2), (7, 3), (5, Pc.ny - 3)]:
        jf = j if j < Pc.ny // 2 else Pf.ny - (Pc.ny - j)
        for m in range(Pc.nsubd):
            c = rng.standard_normal(9)
            for q in range(2):
                xi = -1 + (2 * q + 1 + xf) / 2
                k0 = (2 * m + q) * Pf.nzloc
                Ff[k0:k0 + Pf.nzloc, jf, 2 * ii] = L.legval(xi, c)
                Ff[k0:k0 + Pf.nzloc, jf, 2 * ii + 1] = 0.5 * L.legval(xi, c)
            Fc_want[m * Pc.nzloc:(m + 1) * Pc.nzloc, j, 2 * ii] = L.legval(xc, c)
            Fc_want[m * Pc.nzloc:(m + 1) * Pc.nzloc, j, 2 * ii + 1] = 0.5 * L.legval(xc, c)
    # modes outside the band: x beyond the coarse Nyquist, y at and beyond it
    Ff[:, 1, 2 * 9] = 1.0
    Ff[:, Pc.ny // 2, 2 * 2] = 1.0
    Ff[:, Pf.ny // 2, 2] = 1.0
    got = M.coarsen(Ff, Pf, Pc)
    assert np.abs(got - Fc_want).max() < 1e-12
