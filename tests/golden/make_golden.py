#!/usr/bin/env python3
"""Write the golden fixtures of tests/golden from the reference's recorded outputs.

Run in a container where /root/reference is mounted.  Each fixture is data only
(inputs and expected outputs); provenance is recorded in its header.

* g2_testhelm.json -- G2: testhelm's penalty Helmholtz solution (testout.dat:2-251)
  on the G1 grid: 10 subdomains (output:78-88), nzloc = 25 (output:59),
  epsilon = 1e-2, Dirichlet/Dirichlet, right-hand side -1 (testhelm.f:59-99),
  facdir = 1, facrobin = 1e4 (output:1-48).
* g1_run.json -- G1: the 400-step 2-D forced run logged in ``output``: its
  namelists (output:2-48), subdomains (output:78-88), the per-step maxcalc
  lines (istep-1, tmin, tmax, ktmin, ktmax; the temperature extremum
  locations; dz and CFLZ at the maximum, 16 digits) and the adaptive-dt
  events (dtchange.dat:1-18: time, new dt).
* g3_run.json -- G3: the last 9-step run of the shipped configuration
  (dim.h 256x1, 12x25; input.dat; subd.dat): u at two points after each step
  (uveltest.dat:1041-1049), interface-average fractions (avgint.dat:2-9) and
  the z grid (z_grid.dat:1-300).
"""
import json
import os
import re

REF = "/root/reference/mode1_bbl"
HERE = os.path.dirname(os.path.abspath(__file__))


def g2():
    rows = []
    with open(os.path.join(REF, "testout.dat")) as f:
        for line in f.readlines()[1:]:
            p = line.split()
            if len(p) >= 2:
                rows.append([float(p[0]), float(p[1])])
    sub = []
    with open(os.path.join(REF, "output")) as f:
        lines = f.readlines()
    for line in lines[77:88]:
        m = re.match(r"\s+(\d+)\s+(\S+)\s+(\S+)\s*$", line)
        if m:
            sub.append([float(m.group(2).replace("E", "e")), float(m.group(3))])
    return {"source": "testout.dat:2-251, output:78-88", "nzloc": 25, "epsilon": 1e-2,
            "facdir": 1.0, "facneumn": 1.0, "facrobin": 1e4, "zlen": 5.0,
            "subdomains": sub, "z_fr": rows}


def g3():
    with open(os.path.join(REF, "uveltest.dat")) as f:
        lines = [l for l in f.readlines() if l.strip()]
    probe = [[float(x.replace("D", "E")) for x in l.split()] for l in lines[-9:]]
    with open(os.path.join(REF, "avgint.dat")) as f:
        avg = [[float(x) for x in l.split()] for l in f.readlines()[1:] if l.strip()]
    with open(os.path.join(REF, "z_grid.dat")) as f:
        zg = [float(l.split()[1]) for l in f if l.strip()]
    with open(os.path.join(REF, "subd.dat")) as f:
        sl = f.read().split("\n")
    ns = int(sl[0])
    sub = [[float(v) for v in sl[1 + k].split()] for k in range(ns)]
    params = dict(nx=256, ny=1, nzloc=25, nsubd=12, dt=0.1, xlen=20.0, ylen=1.0, zlen=5.0,
                  brunt=0.2, xnu=1e-5, xkappa=1e-5, grav=9.81, umax=0.01, rho0=1e3,
                  xacrit=2.5, xbcrit=0.84, zcen=0.13, amp=35.0, p=5.0, pf=10.0, fackc=0.0,
                  facdir=1.0, facneumn=1.0, facrobin=1e4, navgintrfc=1, fwave=8.0)
    return {"source": "dim.h:12-20, input.dat:1-49, subd.dat, uveltest.dat:1041-1049, "
                      "avgint.dat:2-9, z_grid.dat:1-300",
            "params": params, "subdomains": sub, "probe_i": 32, "probe_k": [10, 50],
            "probe": probe, "avgint": avg, "z_grid": zg}


def g1():
    with open(os.path.join(REF, "output")) as f:
        lines = f.read().split("\n")
    params = {}
    for l in lines[1:48]:
        m = re.match(r"^\s*([A-Z0-9]+)\s*=\s*([-+0-9.Ee]+)", l)
        if m:
            params[m.group(1).lower()] = float(m.group(2))
    sub = []
    for l in lines[77:88]:
        m = re.match(r"\s+(\d+)\s+(\S+)\s+(\S+)\s*$", l)
        if m:
            sub.append([float(m.group(2)), float(m.group(3))])
    grid = re.findall(r"\d+", lines[58])  # output:59 grid points in x and z
    minmax, tmin, tmax, cflz = [], [], [], []
    for l in lines:
        m = re.match(r"^\s+(\d+)\s+(\S+E[+-]\d\d)\s+(\S+E[+-]\d\d)\s+(\d+)\s+(\d+)\s*$", l)
        if m:
            minmax.append([float(m.group(i)) for i in range(1, 6)])
        elif "Temp. min at i,j,k" in l:
            tmin.append([int(x) for x in l.split("k")[-1].split()])
        elif "Temp. max at i,j,k" in l:
            tmax.append([int(x) for x in l.split("k")[-1].split()])
        elif "At max CFLZ dz=" in l:
            cflz.append([float(x) for x in l.split("=")[1].split()])
    with open(os.path.join(REF, "dtchange.dat")) as f:
        events = [[float(x) for x in l.split()] for l in f if l.strip()]
    return {"source": "output:2-48, 59, 78-88 and its per-step maxcalc lines (various.f:862-885); "
                      "dtchange.dat:1-18",
            "nx": int(grid[0]), "nz": int(grid[1]), "nzloc": 25, "params": params, "subdomains": sub,
            "minmax": minmax, "tmin_loc": tmin, "tmax_loc": tmax, "cflz": cflz, "dt_events": events}


if __name__ == "__main__":
    with open(os.path.join(HERE, "g1_run.json"), "w") as f:
        json.dump(g1(), f, indent=0)
    with open(os.path.join(HERE, "g2_testhelm.json"), "w") as f:
        json.dump(g2(), f, indent=0)
    with open(os.path.join(HERE, "g3_run.json"), "w") as f:
        json.dump(g3(), f, indent=0)
    print("written")
