#!/usr/bin/env python3
"""The batched-ensemble workload: 64 independent instances of uneven sizes
(iwrepo_amd.params.c3_members: Nx, Ny in {32, 64, 128}, 2-8 subdomains of
GLL order 8-24, Re 1e3-1e5, one seed), `--steps` steps each (500).  Timed:
* the members one after another, each advanced by iw_run alone;
* the ensemble with one captured graph per three steps, every member its own
  chain, 256 transform threads (the defaults);
* the ensemble after the autotuner (transform block size per member, number
  of chains or no graph).
Rates are summed unique grid points x steps / s.  Prints one JSON object.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))

from iwrepo_amd import runtime as rt  # noqa: E402
from iwrepo_amd.ic import initial_condition  # noqa: E402
from iwrepo_amd.params import c3_members  # noqa: E402


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def build(mem, fields):
    ss = []
    for (P, _, _, _), f in zip(mem, fields):
        s = rt.Solver(P)
        s.set_initial(*f)
        ss.append(s)
    return ss


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--members", type=int, default=64)
    ap.add_argument("--steps", type=int, default=500)
    ap.add_argument("--tune-steps", type=int, default=9)
    a = ap.parse_args()
    mem = c3_members(a.members)
    pts = sum(P.grid_points() for P, _, _, _ in mem)
    log("ensemble: %d members, %d grid points; initial conditions" % (len(mem), pts))
    fields = [initial_condition(P, w, nz, sd) for P, w, nz, sd in mem]
    out = {"members": len(mem), "grid_points": pts, "steps": a.steps,
           "shapes": [[P.nx, P.ny, P.nsubd, P.nzloc - 1] for P, _, _, _ in mem]}
    # members one after another
    ss = build(mem, fields)
    for s in ss:  # the start-up steps and three more, as the ensemble runs below
        s.run(6, counts=False)
        s.sync()
    t0 = time.perf_counter()
    for s in ss:
        s.run(a.steps, counts=False)
        s.sync()
    out["sequential_s"] = time.perf_counter() - t0
    final_seq = [ss[i].get_state(0) for i in (0, len(ss) - 1)]
    for s in ss:
        s.close()
    log("ensemble: sequential %.2f s" % out["sequential_s"])
    # default ensemble and tuned ensemble
    for mode in ("graph", "tuned"):
        ss = build(mem, fields)
        e = rt.Ensemble(ss)
        e.run(3)
        if mode == "tuned":
            t1 = time.perf_counter()
            out["tune"] = e.tune(a.tune_steps)
            out["tune_s"] = time.perf_counter() - t1
            lanes, ft = e.config()
            out["tuned_lanes"] = lanes
            out["tuned_fft_threads"] = ft.tolist()
        else:
            e.set_lanes(len(ss))
        e.run(3)  # capture
        t0 = time.perf_counter()
        e.run(a.steps)
        out[mode + "_s"] = time.perf_counter() - t0
        same = all(np.array_equal(ss[i].get_state(0), f) for i, f in zip((0, len(ss) - 1), final_seq))
        out[mode + "_state_equals_sequential"] = bool(same)
        e.close()
        for s in ss:
            s.close()
        log("ensemble: %s %.2f s" % (mode, out[mode + "_s"]))
    for k in ("sequential", "graph", "tuned"):
        out[k + "_gps"] = pts * a.steps / out[k + "_s"]
    print(json.dumps(out))


if __name__ == "__main__":
    main()
