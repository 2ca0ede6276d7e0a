"""Seeded initial conditions: a linear mode-1 internal wave plus Gaussian noise.

The wave follows the mode-1 polarisation of the reference's virtual paddle
(``forcing.f:325-414``: u = -(W'(z)/k) sin(kx - wt), w = W(z) cos(kx - wt),
rho' = -(1/w) N^2 (rho0/g) W(z) sin(kx - wt), W = sin(m z)), shifted in phase so
that u = A cos(m z) cos(k x) with k = 2 pi/Lx and m = 2 pi/(2 zlen)
(``main.f:337-346``).  Noise is N(0,1)*amplitude from a counter-based
SplitMix64 stream (Box-Muller), drawn on the unique vertical points so the two
stored copies of an interface point agree.  All fields are physical-space
arrays shaped like a(nxpp,ny,nz) (numpy shape (nz, ny, nxpp)) with zero pads.
"""
import math

import numpy as np

_GOLDEN = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)


def splitmix64(seed, n):
    """n outputs of SplitMix64 started at ``seed`` (counter-based, vectorised)."""
    with np.errstate(over="ignore"):
        z = np.uint64(seed) + (np.arange(1, n + 1, dtype=np.uint64) * _GOLDEN)
        z = (z ^ (z >> np.uint64(30))) * _M1
        z = (z ^ (z >> np.uint64(27))) * _M2
        z = z ^ (z >> np.uint64(31))
    return z


def gaussian(seed, n):
    """n standard normal samples (Box-Muller on pairs of 53-bit uniforms)."""
    m = (n + 1) // 2
    bits = splitmix64(seed, 2 * m)
    u = ((bits >> np.uint64(11)).astype(np.float64) + 0.5) * (1.0 / 9007199254740992.0)
    u1, u2 = u[0::2], u[1::2]
    r = np.sqrt(-2.0 * np.log(u1))
    th = 2.0 * math.pi * u2
    out = np.empty(2 * m)
    out[0::2] = r * np.cos(th)
    out[1::2] = r * np.sin(th)
    return out[:n]


def gll_nodes(order):
    """Gauss-Lobatto-Legendre nodes on [-1,1] (roots of (1-x^2) P'_N)."""
    from numpy.polynomial import legendre as L
    c = np.zeros(order + 1)
    c[-1] = 1.0
    inner = np.sort(L.legroots(L.legder(c))) if order > 1 else np.array([])
    return np.concatenate([[-1.0], inner, [1.0]])


def z_grid(P):
    """Stored z coordinates z(k), k over nsubd*nzloc (setup.f:151-160)."""
    x = gll_nodes(P.nzloc - 1)
    z = []
    for ks in range(P.nsubd):
        z.extend(P.z0[ks] + 0.5 * P.zh[ks] * (x + 1.0))
    return np.array(z)


def initial_condition(P, wave=0.05, noise=1e-4, seed=1):
    """Physical (u, v, w, rho') for parameter block ``P``."""
    nx, ny, nzl, ns = P.nx, P.ny, P.nzloc, P.nsubd
    nxpp, nz = nx + 2, nzl * ns
    x = np.arange(nx) * (P.xlen / nx)
    z = z_grid(P)
    k = 2 * math.pi / P.xlen
    m = 2 * math.pi / (2 * P.zlen)
    n2 = P.brunt * P.brunt
    omega = P.brunt * k / math.sqrt(k * k + m * m)
    cz, sz = np.cos(m * z)[:, None, None], np.sin(m * z)[:, None, None]
    cx, sx = np.cos(k * x)[None, None, :], np.sin(k * x)[None, None, :]
    fields = []
    u = wave * cz * cx
    w = (wave * k / m) * sz * sx
    r = (wave * k / (m * omega)) * (n2 * P.rho0 / P.grav) * sz * cx
    base = [u, np.zeros((1, 1, 1)), w, r]
    # unique-point index of each stored z point
    kuniq = np.array([ks * (nzl - 1) + kl for ks in range(ns) for kl in range(nzl)])
    nuniq = ns * (nzl - 1) + 1
    for q in range(4):
        g = gaussian(seed * 16 + q, nuniq * ny * nx).reshape(nuniq, ny, nx) * noise
        f = np.zeros((nz, ny, nxpp))
        f[:, :, :nx] = np.broadcast_to(base[q], (nz, ny, nx)) + g[kuniq]
        fields.append(f)
    fields[2][0] = 0.0       # w = 0 at the walls (Dirichlet, calcuvwt.f:528-532)
    fields[2][nz - 1] = 0.0
    return tuple(np.ascontiguousarray(f) for f in fields)
