"""CPU restatement of the learned closure: the per-point network and the
coarse-graining that makes its training samples.

TEST INFRASTRUCTURE ONLY: imported by tests/ as the checker of the device
kernels (iwrepo_amd/csrc/iw_closure.hip); nothing in the product imports it.
The closure has no counterpart in the reference; the tensors it learns are
the ones of sgscalc.f:125-219 (t_ij = f(u_i u_j) - f(u_i) f(u_j), th_i =
f(u_i rho) - f(u_i) f(rho)) with f the coarse-graining below.

* Network: 12 -> 64 -> 64 -> 64 -> 9, tanh on the hidden layers, inputs and
  outputs normalised by a mean and a deviation.  With ``emulate`` the BF16
  rounding of the device's matrix-core forward is applied at the same points
  (normalised inputs, weights, hidden activations; FP64 accumulation here,
  FP32 on the device), and the backward uses the BF16 weights and activations
  (the straight-through rule the device uses).  Without it the network is
  plain FP64 and the gradient is the exact gradient of the loss, which the
  CPU tests pin by central differences.
* Coarse-graining: spectral cut-off in x, y (coarse |kx|, |ky| below their
  Nyquist wavenumbers) and the L2 projection in z onto the Legendre
  polynomials of the coarse order over coarse subdomains made of r fine ones;
  computed here independently of the device's matrix: the fine nodal
  polynomial is fitted exactly (numpy.polynomial.legendre), integrated with a
  64-point Gauss-Legendre rule per fine subdomain and evaluated at the coarse
  GLL nodes.
"""
import numpy as np
from numpy.polynomial import legendre as L

NIN, NINP, NH, NOUT, NOUTP = 12, 16, 64, 9, 16
OW1 = 0
OB1 = OW1 + NH * NINP
OW2 = OB1 + NH
OB2 = OW2 + NH * NH
OW3 = OB2 + NH
OB3 = OW3 + NH * NH
OW4 = OB3 + NH
OB4 = OW4 + NOUTP * NH
NPARAM = OB4 + NOUTP
NNORM = 2 * NIN + 2 * NOUT


def bf16(a):
    """Round to BF16 (nearest even), returned as float32 values."""
    a = np.ascontiguousarray(a, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    u = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return (u & 0xFFFFFFFF).astype(np.uint32).view(np.float32)


def unpack(p):
    p = np.asarray(p, dtype=np.float64)
    W1 = p[OW1:OB1].reshape(NH, NINP)
    W2 = p[OW2:OB2].reshape(NH, NH)
    W3 = p[OW3:OB3].reshape(NH, NH)
    W4 = p[OW4:OB4].reshape(NOUTP, NH)
    return [W1, W2, W3, W4], [p[OB1:OW2], p[OB2:OW3], p[OB3:OW4], p[OB4:NPARAM]]


def _norm_split(norm):
    norm = np.asarray(norm, dtype=np.float32)
    return norm[:NIN], norm[NIN:2 * NIN], norm[2 * NIN:2 * NIN + NOUT], norm[2 * NIN + NOUT:]


def _round(a, emulate):
    return bf16(a).astype(np.float64) if emulate else np.asarray(a, dtype=np.float64)


def forward(p, norm, X, emulate=True):
    """Normalised outputs (n x 9) and the cache for the backward."""
    Ws, bs = unpack(p)
    mx, sx, _, _ = _norm_split(norm)
    xn = (np.asarray(X, dtype=np.float32) - mx) / sx if emulate else (np.asarray(X) - mx.astype(np.float64)) / sx
    a = np.zeros((len(X), NINP))
    a[:, :NIN] = _round(xn, emulate)
    acts, hs = [a], []
    Wr = [_round(W, emulate) for W in Ws]
    for layer in range(3):
        z = acts[-1] @ Wr[layer].T + bs[layer]
        h = np.tanh(z)
        hs.append(h)
        acts.append(_round(h, emulate))
    y = acts[-1] @ Wr[3].T + bs[3]
    return y[:, :NOUT], (acts, hs, Wr)


def loss_grad(p, norm, X, T, emulate=True):
    """Mean squared error of the normalised outputs and its gradient
    (NPARAM values, padded entries zero)."""
    y, (acts, hs, Wr) = forward(p, norm, X, emulate)
    _, _, my, sy = _norm_split(norm)
    tn = (np.asarray(T, dtype=np.float32) - my) / sy if emulate else (np.asarray(T) - my.astype(np.float64)) / sy
    n = len(X)
    e = y - tn
    loss = float(np.sum(e * e) / (n * NOUT))
    d = np.zeros((n, NOUTP))
    d[:, :NOUT] = 2.0 * e / (n * NOUT)
    g = np.zeros(NPARAM)
    offs = [(OW1, OB1, OW2), (OW2, OB2, OW3), (OW3, OB3, OW4), (OW4, OB4, NPARAM)]
    for layer in (3, 2, 1, 0):
        ow, ob, oe = offs[layer]
        g[ow:ob] = (d.T @ acts[layer]).ravel()
        g[ob:oe] = d.sum(axis=0)
        if layer > 0:
            d = (d @ Wr[layer]) * (1.0 - hs[layer - 1] ** 2)
    return loss, g


def predict(p, norm, X, emulate=True):
    y, _ = forward(p, norm, X, emulate)
    _, _, my, sy = _norm_split(norm)
    return my.astype(np.float64) + sy.astype(np.float64) * y


# ------------------------------------------------------------ coarse-graining


def gll_nodes(order):
    c = np.zeros(order + 1)
    c[-1] = 1.0
    return np.concatenate([[-1.0], np.sort(L.legroots(L.legder(c))), [1.0]])


def zprojection(order_f, order_c, r):
    """Matrix (nlc x r nlf): coarse GLL nodal values of the L2 projection onto
    degree order_c of the piecewise polynomial given at the fine GLL nodes
    of r equal fine subdomains of one coarse subdomain."""
    xf, xc = gll_nodes(order_f), gll_nodes(order_c)
    g, w = L.leggauss(64)
    nlf, nlc = order_f + 1, order_c + 1
    M = np.zeros((nlc, r * nlf))
    for q in range(r):
        for k in range(nlf):
            e = np.zeros(nlf)
            e[k] = 1.0
            cf = L.legfit(xf, e, order_f)  # the fine nodal basis function (exact fit)
            fq = L.legval(g, cf)           # on the Gauss points of fine subdomain q
            xi = -1.0 + (2 * q + 1 + g) / r  # those points in the coarse coordinate
            for n in range(order_c + 1):
                cn = np.zeros(n + 1)
                cn[n] = 1.0
                a = (2 * n + 1) / 2.0 * np.sum(w * fq * L.legval(xi, cn)) / r
                M[:, q * nlf + k] += a * L.legval(xc, cn)
    return M


def coarsen(Ff, Pf, Pc):
    """Coarse spectral array (nz_c, ny_c, nxpp_c) from a fine one."""
    r = Pf.nsubd // Pc.nsubd
    M = zprojection(Pf.nzloc - 1, Pc.nzloc - 1, r)
    nlf, nlc = Pf.nzloc, Pc.nzloc
    nyc, nxhc = Pc.ny, Pc.nx // 2
    out = np.zeros((Pc.nzloc * Pc.nsubd, nyc, Pc.nx + 2))
    rows = [j if j < nyc // 2 else Pf.ny - (nyc - j) for j in range(nyc)]
    keep = [j for j in range(nyc) if j != nyc // 2]
    for m in range(Pc.nsubd):
        blk = Ff[m * r * nlf:(m + 1) * r * nlf]  # (r nlf, ny_f, nxpp_f)
        proj = np.tensordot(M, blk, axes=(1, 0))  # (nlc, ny_f, nxpp_f)
        for j in keep:
            out[m * nlc:(m + 1) * nlc, j, :2 * nxhc] = proj[:, rows[j], :2 * nxhc]
    return out
