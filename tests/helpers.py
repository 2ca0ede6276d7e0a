"""Shared helpers of the test suite (parameter blocks, norms)."""
import json
import os

import numpy as np

from iwrepo_amd.params import CONFIGS, config_params, make_params
from iwrepo_amd.ic import initial_condition

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")

STATE = "u v w t un vn wn tn unm vnm wnm tnm su sv sw st sun svn swn stn sunm svnm swnm stnm".split()
# arrays that carry the solution (su..st hold scratch copies between steps)
SOLUTION = list(range(12)) + list(range(16, 24))


def rel(a, b):
    n = np.linalg.norm(b)
    return float(np.linalg.norm(a - b) / n) if n > 0 else float(np.linalg.norm(a))


def case(name):
    P = config_params(name)
    c = CONFIGS[name]
    return P, initial_condition(P, c["wave"], c["noise"], c["seed"])


def g3_params():
    d = json.load(open(os.path.join(GOLDEN, "g3_run.json")))
    q = dict(d["params"])
    P = make_params(nx=q["nx"], ny=q["ny"], nsubd=q["nsubd"], order=q["nzloc"] - 1, xlen=q["xlen"],
                    ylen=q["ylen"], zlen=q["zlen"], re=1.0 / q["xnu"], pr=q["xnu"] / q["xkappa"],
                    n2=q["brunt"] ** 2, dt=q["dt"], amp=q["amp"], p=q["p"], pf=q["pf"],
                    fackc=q["fackc"], facdir=q["facdir"], facneumn=q["facneumn"],
                    facrobin=q["facrobin"], navgintrfc=q["navgintrfc"], zcen=q["zcen"],
                    xbcrit=q["xbcrit"], xacrit=q["xacrit"], umax=q["umax"], fwave=q["fwave"],
                    heights=[h for _, h in d["subdomains"]], grav=q["grav"], rho0=q["rho0"])
    for k, (z0, _) in enumerate(d["subdomains"]):
        P.z0[k] = z0
    return P, d


def g1_params(dtadapt=2):
    """G1: the logged 400-step 2-D forced run (tests/golden/g1_run.json)."""
    d = json.load(open(os.path.join(GOLDEN, "g1_run.json")))
    q = d["params"]
    sub = d["subdomains"]
    P = make_params(nx=d["nx"], ny=1, nsubd=len(sub), order=d["nzloc"] - 1, xlen=q["xlen"], ylen=q["ylen"],
                    zlen=q["zlen"], re=1.0 / q["xnu"], pr=q["xnu"] / q["xkappa"], n2=q["brunt"] ** 2,
                    dt=q["dtin"], amp=q["amp"], p=q["p"], pf=q["pf"], fackc=q["fackc"], facdir=q["facdir"],
                    facneumn=q["facneumn"], facrobin=q["facrobin"], navgintrfc=int(q["navgintrfc"]),
                    zcen=q["zcen"], xbcrit=q["xbcrit"], xacrit=q["xacrit"], umax=q["umax"], fwave=8.0,
                    heights=[h for _, h in sub], grav=q["grav"], rho0=q["rho0"], fackcbv=q["fackcbv"],
                    dtadapt=dtadapt)
    P.brunt = q["brunt"]
    for k, (z0, _) in enumerate(sub):
        P.z0[k] = z0
    return P, d


def g2_params():
    d = json.load(open(os.path.join(GOLDEN, "g2_testhelm.json")))
    sub = d["subdomains"]
    P = make_params(nx=4, ny=1, nsubd=len(sub), order=d["nzloc"] - 1, zlen=d["zlen"],
                    heights=[h for _, h in sub], facdir=d["facdir"], facneumn=d["facneumn"],
                    facrobin=d["facrobin"])
    for k, (z0, _) in enumerate(sub):
        P.z0[k] = z0
    return P, d


def routine_inputs(O, P, flds):
    """Physical fields for the horfft check and spectral fields (rough, with
    interface jumps) for the dpdz, filter and interface-averaging checks."""
    o = O.Oracle(P)
    rng = np.random.default_rng(11)
    spec = []
    for q in range(4):
        f = flds[q] + 0.05 * rng.standard_normal(o.shape)
        f[:, :, P.nx:] = 0.0
        spec.append(o.horfft(f, -1))
    return o, spec


def meanwave_fields(P, z):
    """Nine smooth (nz, nxpp) arrays for /sol2dvar/ (COMMON order: umwv,
    wmwv, dumdx, dumdz, dwmdx, dwmdz, rhomwv, drhomdx, drhomdz), periodic in
    x, O(1e-2)."""
    nxpp = P.nx + 2
    x = np.arange(nxpp) * (P.xlen / P.nx)
    kx = 2 * np.pi / P.xlen
    spec = [(0.02, np.cos, np.cos), (0.01, np.sin, np.sin), (0.02, np.sin, np.cos), (0.03, np.cos, np.sin),
            (0.01, np.cos, np.sin), (0.01, np.sin, np.cos), (0.05, np.sin, np.sin), (0.02, np.cos, np.cos),
            (0.04, np.sin, np.cos)]
    z = np.asarray(z)
    return np.stack([a * fz(np.pi * z)[:, None] * fx(kx * x)[None, :] for a, fx, fz in spec])

