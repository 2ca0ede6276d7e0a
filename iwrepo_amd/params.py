"""Run parameters of the stratified-Boussinesq step and the named configurations.

The parameter block mirrors the reference's namelists and ``dim.h``
(``main.f:185-194``, ``input.dat:1-49``, ``dim.h:12-28``) and has the C layout of
``iw_params`` in ``include/iwgpu.h``.
"""
import ctypes
import math

MAXSUB = 64


class IWParams(ctypes.Structure):
    _fields_ = [
        ("nx", ctypes.c_int), ("ny", ctypes.c_int), ("nzloc", ctypes.c_int), ("nsubd", ctypes.c_int),
        ("bc_uv_bottom", ctypes.c_int),
        ("ispecfilternl", ctypes.c_int), ("ispecfilterexp", ctypes.c_int),
        ("nspecfilterexp", ctypes.c_int), ("navgintrfc", ctypes.c_int),
        ("ifilter", ctypes.c_int), ("nstepfilter", ctypes.c_int), ("npostp", ctypes.c_int),
        ("imaxcalc", ctypes.c_int), ("dtadapt", ctypes.c_int), ("n3doutput", ctypes.c_int),
        ("ncalclegspec", ctypes.c_int),
        ("xlen", ctypes.c_double), ("ylen", ctypes.c_double), ("zlen", ctypes.c_double),
        ("xnu", ctypes.c_double), ("xkappa", ctypes.c_double), ("brunt", ctypes.c_double),
        ("grav", ctypes.c_double), ("rho0", ctypes.c_double), ("omz", ctypes.c_double),
        ("dt", ctypes.c_double),
        ("amp", ctypes.c_double), ("p", ctypes.c_double), ("pf", ctypes.c_double),
        ("fackc", ctypes.c_double),
        ("facdir", ctypes.c_double), ("facneumn", ctypes.c_double), ("facrobin", ctypes.c_double),
        ("zcen", ctypes.c_double), ("xbcrit", ctypes.c_double),
        ("xacrit", ctypes.c_double), ("umax", ctypes.c_double), ("fwave", ctypes.c_double),
        ("fackcbv", ctypes.c_double), ("fcshear", ctypes.c_double), ("uo", ctypes.c_double),
        ("z0", ctypes.c_double * MAXSUB), ("zh", ctypes.c_double * MAXSUB),
    ]

    @property
    def nz(self):
        return self.nzloc * self.nsubd

    @property
    def nxpp(self):
        return self.nx + 2

    def shape(self):
        """numpy shape (C order) of one field a(nxpp,ny,nz)."""
        return (self.nz, self.ny, self.nxpp)

    def grid_points(self):
        """Unique grid points Nx*Ny*(nsubd*p+1), the grid-point-steps/s unit."""
        return self.nx * self.ny * (self.nsubd * (self.nzloc - 1) + 1)


def make_params(nx, ny, nsubd, order, xlen=2 * math.pi, ylen=math.pi / 2, zlen=1.0,
                re=1e4, pr=1.0, n2=1.0, dt=1e-3, noslip=False, amp=36.0, p=4.0, pf=4.0,
                fackc=0.0, facdir=1.0, facneumn=1.0, facrobin=1e4, navgintrfc=1,
                ispecfilternl=1, ispecfilterexp=1, nspecfilterexp=1, zcen=0.13, xbcrit=0.84,
                xacrit=0.0, umax=0.05, fwave=0.0, heights=None, grav=1.0, rho0=1.0,
                ifilter=0, nstepfilter=5, npostp=100000000, fackcbv=0.0, imaxcalc=1, dtadapt=0,
                fcshear=0.0, uo=0.07, n3doutput=0, ncalclegspec=0):
    """Build a parameter block.  ``order`` is the GLL polynomial order p
    (nzloc = p+1); z in [0,zlen] is split into equal subdomains unless
    ``heights`` gives them (subd.dat, setup.f:297-330)."""
    P = IWParams()
    P.nx, P.ny, P.nzloc, P.nsubd = nx, ny, order + 1, nsubd
    if nsubd > MAXSUB:
        raise ValueError("at most %d subdomains" % MAXSUB)
    P.bc_uv_bottom = 1 if noslip else 0
    P.ispecfilternl, P.ispecfilterexp = ispecfilternl, ispecfilterexp
    P.nspecfilterexp, P.navgintrfc = nspecfilterexp, navgintrfc
    P.xlen, P.ylen, P.zlen = xlen, ylen, zlen
    P.xnu = 1.0 / re
    P.xkappa = P.xnu / pr
    P.brunt = math.sqrt(n2)
    P.grav, P.rho0, P.omz = grav, rho0, 0.0
    P.dt = dt
    P.amp, P.p, P.pf, P.fackc = amp, p, pf, fackc
    P.facdir, P.facneumn, P.facrobin = facdir, facneumn, facrobin
    P.zcen, P.xbcrit = zcen, xbcrit
    P.xacrit, P.umax, P.fwave = xacrit, umax, fwave
    P.ifilter, P.nstepfilter, P.npostp, P.fackcbv = ifilter, nstepfilter, npostp, fackcbv
    P.imaxcalc, P.dtadapt = imaxcalc, dtadapt
    P.fcshear, P.uo = fcshear, uo
    P.n3doutput, P.ncalclegspec = n3doutput, ncalclegspec
    if heights is None:
        heights = [zlen / nsubd] * nsubd
    z = 0.0
    for k, h in enumerate(heights):
        P.z0[k] = z
        P.zh[k] = h
        z += h
    return P


# Named workloads.  Sizes follow the BASELINE configurations; the wave and noise
# amplitudes feed iwrepo_amd.ic.initial_condition.
CONFIGS = {
    # mode-1 internal wave, free slip, 64x16x(4x17), 200 steps
    "c1": dict(params=dict(nx=64, ny=16, nsubd=4, order=16, xlen=2 * math.pi, ylen=math.pi / 2,
                           re=1e4, pr=1.0, dt=1e-3), wave=0.05, noise=1e-4, steps=200, seed=1),
    # single large instance: 256x128x(16x25), no-slip bottom, 1000 steps
    "c2": dict(params=dict(nx=256, ny=128, nsubd=16, order=24, xlen=4 * math.pi,
                           ylen=2 * math.pi, re=5e4, pr=7.0, dt=5e-4, noslip=True),
               wave=0.1, noise=1e-3, steps=1000, seed=2),
    # largest single-GPU case: 512x256x(32x25), no-slip bottom, 500 steps
    "c5": dict(params=dict(nx=512, ny=256, nsubd=32, order=24, xlen=4 * math.pi,
                           ylen=2 * math.pi, re=2e5, pr=7.0, dt=2.5e-4, noslip=True),
               wave=0.2, noise=1e-3, steps=500, seed=5),
    # learned-closure fine run: 256x256x(16x25), Re 1e5, wave amplitude 0.15,
    # 2000 steps (box, bottom condition, Pr and dt as in the single large instance)
    "c4f": dict(params=dict(nx=256, ny=256, nsubd=16, order=24, xlen=4 * math.pi,
                            ylen=2 * math.pi, re=1e5, pr=7.0, dt=5e-4, noslip=True),
                wave=0.15, noise=1e-3, steps=2000, seed=4),
    # its coarse counterpart: 64x64x(8x13)
    "c4c": dict(params=dict(nx=64, ny=64, nsubd=8, order=12, xlen=4 * math.pi,
                            ylen=2 * math.pi, re=1e5, pr=7.0, dt=5e-4, noslip=True),
                wave=0.15, noise=1e-3, steps=2000, seed=4),
    # small cases for fast parity tests
    "small": dict(params=dict(nx=16, ny=8, nsubd=3, order=8, re=1e3, dt=2e-3),
                  wave=0.05, noise=1e-3, steps=10, seed=7),
    "smallns": dict(params=dict(nx=16, ny=8, nsubd=3, order=8, re=1e3, dt=2e-3, noslip=True),
                    wave=0.05, noise=1e-3, steps=10, seed=8),
    # GLL order 10: its u,v,w,rho solves take the one-wavefront-per-system path
    "small11": dict(params=dict(nx=16, ny=8, nsubd=3, order=10, re=1e3, dt=2e-3),
                    wave=0.05, noise=1e-3, steps=10, seed=9),
    # GLL orders 12 and 20 (13 and 21 points per subdomain) on the packed-factor sweeps
    "small13": dict(params=dict(nx=16, ny=8, nsubd=3, order=12, re=1e3, dt=2e-3, noslip=True),
                    wave=0.05, noise=1e-3, steps=10, seed=10),
    "small21": dict(params=dict(nx=16, ny=8, nsubd=2, order=20, re=1e3, dt=2e-3),
                    wave=0.05, noise=1e-3, steps=10, seed=11),
    # every other order 8..24 on the specialised kernels (orders 9, 11, 14,
    # 15, 18, 19, 22, 23 here; 10, 12, 16, 20 above)
    **{"small%d" % (o + 1): dict(params=dict(nx=16, ny=8, nsubd=2, order=o, re=1e3, dt=2e-3, noslip=o % 2 == 1),
                                 wave=0.05, noise=1e-3, steps=10, seed=20 + o)
       for o in (9, 11, 14, 15, 18, 19, 22, 23)},
    # lengths with factors 3 and 5 (the radix-3/5 passes): x lengths
    # M = nx/2 = 24, 48, 30, 45 and y lengths 24, 40, 18
    "r48": dict(params=dict(nx=48, ny=24, nsubd=2, order=8, re=1e3, dt=2e-3),
                wave=0.05, noise=1e-3, steps=10, seed=40),
    "r96": dict(params=dict(nx=96, ny=40, nsubd=3, order=12, re=1e3, dt=2e-3, noslip=True),
                wave=0.05, noise=1e-3, steps=10, seed=41),
    "r60": dict(params=dict(nx=60, ny=18, nsubd=2, order=10, re=1e3, dt=2e-3),
                wave=0.05, noise=1e-3, steps=10, seed=42),
    "r90": dict(params=dict(nx=90, ny=1, nsubd=3, order=8, re=1e3, dt=2e-3, noslip=True),
                wave=0.05, noise=1e-3, steps=10, seed=43),
    "c1ns": dict(params=dict(nx=64, ny=16, nsubd=4, order=16, xlen=2 * math.pi, ylen=math.pi / 2,
                             re=1e4, pr=1.0, dt=1e-3, noslip=True), wave=0.05, noise=1e-4,
                 steps=200, seed=3),
}


def config_params(name):
    return make_params(**CONFIGS[name]["params"])


def c3_members(n=64, seed=0x1D5EC0F1 + 3):
    """The batched-ensemble workload: n independent instances drawn from one
    seed (NumPy PCG64): Nx, Ny uniform in {32, 64, 128}, 2-8 subdomains of GLL
    order 8-24, wave amplitude uniform in [0.01, 0.1], Re log-uniform in
    [1e3, 1e5]; the box, Pr, dt and the no-slip bottom of the single large
    instance, dt 1e-3; noise 1e-3.  Returns [(params, wave, noise, ic_seed)]."""
    import numpy as np
    rng = np.random.default_rng(seed)
    out = []
    for k in range(n):
        nx, ny = (int(v) for v in rng.choice([32, 64, 128], size=2))
        nsubd = int(rng.integers(2, 9))
        order = int(rng.integers(8, 25))
        wave = float(rng.uniform(0.01, 0.1))
        re = float(10 ** rng.uniform(3.0, 5.0))
        P = make_params(nx=nx, ny=ny, nsubd=nsubd, order=order, xlen=4 * math.pi, ylen=2 * math.pi, re=re,
                        pr=7.0, dt=1e-3, noslip=True, imaxcalc=0)
        out.append((P, wave, 1e-3, 1000 + k))
    return out

