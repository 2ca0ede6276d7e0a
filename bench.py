#!/usr/bin/env python3
"""Throughput of the stratified-Boussinesq time step on one MI355X.

Workload: 256 x 128 Fourier modes on a 4pi x 2pi box, z in [0,1] as 16 GLL
subdomains of order 24 (385 unique points), no-slip bottom, Re=5e4, Pr=7,
N^2=1, dt=5e-4; mode-1 wave of amplitude 0.1 plus seeded noise 1e-3.  Plain
FP64 path.  A step is one full pass of the reference step (maxcalc round
trip, convec, calcuvwt, interface averaging) over the device-resident state.

Prints ONE JSON line.  ``value`` = grid-point-steps/s (unique grid points
Nx*Ny*(nsubd*p+1) per second) measured over exactly --steps steps after
--warmup untimed steps (the warmup covers the three start-up factorisations
of the implicit operators, whose factors are then reused while dt is fixed).
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

MI355X_HBM_GBS = 8000.0  # published HBM3E peak, GB/s


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def cpu_baseline(P, fields, nsteps=1):
    """The CPU oracle (C restatement of the reference step, per-step
    refactorisation as in the reference) on the host cores, bounded sample."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    o = O.Oracle(P)
    o.set_initial(*fields)
    t0 = time.perf_counter()
    o.run(nsteps)
    dt = time.perf_counter() - t0
    threads = O.lib().iwo_nthreads()
    return {"value": P.grid_points() * nsteps / dt, "unit": "grid-point-steps/s", "cores": threads,
            "kind": "port",
            "sample": "%d step(s) of the same 256x128x(16x25) workload from the same initial "
                      "condition, %.1f s wall on %d OpenMP threads" % (nsteps, dt, threads)}


def load_traffic(kernel, workload):
    path = os.path.join(ROOT, "profiles", "traffic_%s_%s.json" % (kernel, workload))
    if not os.path.exists(path):
        return None
    try:
        d = json.load(open(path))
        return float(d["hbm_bytes_per_launch"])
    except Exception:
        return None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--config", default="c2")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--profile-steps", type=int, default=3)
    a = ap.parse_args()
    if a.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py: this build runs on one GPU only; --gpus %d is not supported "
              "(the step is measured on one MI355X)" % a.gpus, file=sys.stderr)
        sys.exit(2)

    from iwrepo_amd.params import CONFIGS, config_params
    from iwrepo_amd.ic import initial_condition
    from iwrepo_amd.runtime import Solver

    P = config_params(a.config)
    c = CONFIGS[a.config]
    log("bench: workload %s, initial condition" % a.config)
    fields = initial_condition(P, c["wave"], c["noise"], c["seed"])
    s = Solver(P)
    s.set_initial(*fields)
    log("bench: warmup %d steps" % a.warmup)
    s.run(a.warmup, counts=False)
    s.sync()
    log("bench: timing %d steps" % a.steps)
    t0 = time.perf_counter()
    s.run(a.steps, counts=False)
    s.sync()
    elapsed = time.perf_counter() - t0
    ms_per_step = 1e3 * elapsed / a.steps
    gps = P.grid_points() * a.steps / elapsed
    u = s.get_state(0)
    import numpy as np
    if not np.all(np.isfinite(u)):
        raise SystemExit("bench: non-finite state")

    # per-kernel device times (HIP events on the solver stream), separate short run
    s.set_profiling(True)
    s.run(a.profile_steps, counts=False)
    stats = s.kernel_stats()
    s.set_profiling(False)
    tot = {k: v["avg_ms"] * v["launches"] for k, v in stats.items()}
    dom = max(tot, key=tot.get)
    dstat = stats[dom]
    achieved = dstat["avg_bytes"] / (dstat["avg_ms"] * 1e-3) / 1e9
    traffic = load_traffic(dom, a.config)
    roofline = {"bound": "hbm", "kernel": dom, "achieved": round(achieved, 1),
                "peak": MI355X_HBM_GBS, "unit": "GB/s", "frac": round(achieved / MI355X_HBM_GBS, 4),
                "traffic": traffic, "algorithmic_bytes_per_launch": dstat["avg_bytes"],
                "avg_launch_ms": round(dstat["avg_ms"], 4),
                "step_share": round(tot[dom] / sum(tot.values()), 3)}
    kernels = {k: {"avg_ms": round(v["avg_ms"], 4), "launches_per_step": v["launches"] / a.profile_steps,
                   "GBps": round(v["avg_bytes"] / max(v["avg_ms"], 1e-9) / 1e6, 1)}
               for k, v in sorted(stats.items(), key=lambda kv: -tot[kv[0]])}
    devbytes = s.device_bytes()
    s.close()

    cpu = None
    if not a.no_cpu_baseline:
        log("bench: cpu baseline")
        cpu = cpu_baseline(P, fields, 5)

    out = {"metric": "grid-point-steps/s on 1 GPU, mixed/learned vs plain FP64 at fixed error",
           "value": gps, "unit": "grid-point-steps/s", "n_gpus": 1, "steps": a.steps,
           "warmup": a.warmup, "ms_per_step": ms_per_step, "higher_is_better": True,
           "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
           "config": {"workload": "256x128 Fourier x 16 GLL subdomains of order 24 (385 points), "
                                  "4pi x 2pi x 1, no-slip bottom, Re=5e4, Pr=7, N^2=1, dt=5e-4, "
                                  "mode-1 wave 0.1 + seeded noise 1e-3; plain FP64 path",
                      "grid_points": P.grid_points(), "path": "plain-fp64",
                      "factor_cache": "implicit-operator factors reused while dt is fixed"},
           "roofline": roofline, "cpu_baseline": cpu, "kernels": kernels,
           "device_bytes": devbytes}
    print(json.dumps(out))


if __name__ == "__main__":
    main()
