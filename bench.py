#!/usr/bin/env python3
"""Throughput of the stratified-Boussinesq time step on one MI355X, mixed
precision against plain FP64 at equal error.

Workload (default ``--config c2``): 256 x 128 Fourier modes on a 4pi x 2pi box,
z in [0,1] as 16 GLL subdomains of order 24 (385 unique points), no-slip
bottom, Re=5e4, Pr=7, N^2=1, dt=5e-4; mode-1 wave of amplitude 0.1 plus
seeded noise 1e-3.  ``--config c5`` runs 512 x 256 x (32 x 25) (Re 2e5,
dt 2.5e-4).  A step is one full pass of the reference step (maxcalc with its
records and round trip, convec, calcuvwt, interface averaging) over the
device-resident state.

Both forms of the step are timed: the plain FP64 path, and the mixed path
(FP32 physical-space working set in convec, FP32 matrix-core z contraction,
FP64 state, filters, banded solves and projection).  The equal-error check
runs both for ``--error-steps`` steps (1000 at the default workload) from the
same initial condition and compares the kinetic and potential energy series
(every 10 steps) of the mixed run with the plain run's; the plain path is the
FP64 reference here (it matches the CPU oracle within 1e-10 per field in the
parity tests).  ``value`` is the mixed path's rate when every KE/PE series
is within 1e-4 relative L2, else the plain path's.

Prints ONE JSON line.  Rates are unique grid points Nx*Ny*(nsubd*p+1) per
second over exactly --steps timed steps after --warmup untimed ones (the
warmup covers the three start-up factorisations of the implicit operators).
"""
import argparse
import json
import os
import platform
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

MI355X_HBM_GBS = 8000.0  # published HBM3E peak, GB/s
STATE_BYTES_PER_POINT = 224.0  # compulsory FP64 state traffic of one SS3/BDF3 step
ENERGY_TOL = 1e-4

WORKLOADS = {
    "c2": "256x128 Fourier x 16 GLL subdomains of order 24 (385 points), 4pi x 2pi x 1, no-slip bottom, "
          "Re=5e4, Pr=7, N^2=1, dt=5e-4, mode-1 wave 0.1 + seeded noise 1e-3",
    "c5": "512x256 Fourier x 32 GLL subdomains of order 24 (769 points), 4pi x 2pi x 1, no-slip bottom, "
          "Re=2e5, Pr=7, N^2=1, dt=2.5e-4, mode-1 wave 0.2 + seeded noise 1e-3",
    "c4f": "256x256 Fourier x 16 GLL subdomains of order 24 (385 points), 4pi x 2pi x 1, no-slip bottom, "
           "Re=1e5, Pr=7, N^2=1, dt=5e-4, mode-1 wave 0.15 + seeded noise 1e-3",
}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or "unknown"


def cpu_baseline(P, fields, nsteps, threads, label):
    """The CPU oracle (C restatement of the reference step, per-step
    refactorisation as in the reference) on the host cores, bounded sample."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    L = O.lib()
    if threads:
        L.iwo_set_threads(threads)
    o = O.Oracle(P)
    o.set_initial(*fields)
    t0 = time.perf_counter()
    o.run(nsteps)
    dt = time.perf_counter() - t0
    n = L.iwo_nthreads()
    return {"value": P.grid_points() * nsteps / dt, "unit": "grid-point-steps/s", "cores": n,
            "kind": "port",
            "sample": "%d step(s) of the same %s workload from the same initial condition, %.1f s wall on %d "
                      "OpenMP thread(s) of %s" % (nsteps, label, dt, n, cpu_model())}


def load_traffic(kernel, workload, mixed=False):
    """HBM bytes per launch of a kernel class from its committed counter
    summary; on the mixed path the mixed form's file when the class has one
    (bdf_filter reads FP32 histories there)."""
    names = (["traffic_%s_mixed_%s.json" % (kernel, workload)] if mixed else []) + \
        ["traffic_%s_%s.json" % (kernel, workload)]
    for name in names:
        path = os.path.join(ROOT, "profiles", name)
        if not os.path.exists(path):
            continue
        try:
            d = json.load(open(path))
            return float(d["hbm_bytes_per_launch"]), "profiles/" + name
        except Exception:
            return None, None
    return None, None


def time_path(Solver, P, fields, mixed, a):
    s = Solver(P)
    s.set_mixed(mixed)
    s.set_initial(*fields)
    s.run(a.warmup, counts=False)
    s.sync()
    t0 = time.perf_counter()
    s.run(a.steps, counts=False)
    s.sync()
    elapsed = time.perf_counter() - t0
    import numpy as np
    if not np.all(np.isfinite(s.get_state(0))):
        raise SystemExit("bench: non-finite state")
    # per-kernel device times (HIP events on the solver stream), separate short run
    s.set_profiling(True)
    s.run(a.profile_steps, counts=False)
    stats = s.kernel_stats()
    s.set_profiling(False)
    devbytes = s.device_bytes()
    s.close()
    return elapsed, stats, devbytes


def energy_series(Solver, P, fields, mixed, nsteps, every=10):
    import numpy as np
    s = Solver(P)
    s.set_mixed(mixed)
    s.set_initial(*fields)
    ke = []
    for n in range(nsteps // every):
        s.run(every, counts=False)
        ke.append(s.kepe())
    st = [s.get_state(q) for q in range(4)]
    s.close()
    return np.array(ke), st


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--config", default="c2")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--profile-steps", type=int, default=3)
    ap.add_argument("--error-steps", type=int, default=None,
                    help="steps of the equal-error check (default: min(1000, the workload's steps))")
    a = ap.parse_args()
    if a.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py: this build runs on one GPU only; --gpus %d is not supported "
              "(the step is measured on one MI355X)" % a.gpus, file=sys.stderr)
        sys.exit(2)

    import numpy as np
    from iwrepo_amd.params import CONFIGS, config_params
    from iwrepo_amd.ic import initial_condition
    from iwrepo_amd.runtime import Solver

    P = config_params(a.config)
    c = CONFIGS[a.config]
    nerr = a.error_steps if a.error_steps is not None else min(1000, c["steps"])
    log("bench: workload %s, initial condition" % a.config)
    fields = initial_condition(P, c["wave"], c["noise"], c["seed"])

    res = {}
    for name, mixed in (("plain", False), ("mixed", True)):
        log("bench: %s path, warmup %d, timing %d steps" % (name, a.warmup, a.steps))
        res[name] = time_path(Solver, P, fields, mixed, a)

    # equal error: KE/PE series of the mixed run against the plain FP64 run
    err = None
    if nerr >= 10:
        log("bench: equal-error check, %d steps of each path" % nerr)
        kp, sp = energy_series(Solver, P, fields, False, nerr)
        km, sm = energy_series(Solver, P, fields, True, nerr)
        series = [float(np.linalg.norm(km[:, q] - kp[:, q]) / np.linalg.norm(kp[:, q])) for q in range(4)]
        final = [float(np.linalg.norm(km[-1, q] - kp[-1, q]) / abs(kp[-1, q])) for q in range(4)]
        fields_rel = [float(np.linalg.norm(x - y) / np.linalg.norm(y)) for x, y in zip(sm, sp)]
        err = {"steps": nerr, "every": 10, "reference": "plain FP64 path",
               "kepe_series_rel_l2": dict(zip(["KE_u", "KE_v", "KE_w", "PE"], series)),
               "kepe_final_rel": dict(zip(["KE_u", "KE_v", "KE_w", "PE"], final)),
               "field_final_rel_l2": dict(zip(["u", "v", "w", "rho"], fields_rel)),
               "bound": ENERGY_TOL, "within": bool(max(series) <= ENERGY_TOL)}

    rates = {k: P.grid_points() * a.steps / v[0] for k, v in res.items()}
    ms = {k: 1e3 * v[0] / a.steps for k, v in res.items()}
    head = "mixed" if (err and err["within"]) else "plain"
    elapsed, stats, devbytes = res[head]

    tot = {k: v["avg_ms"] * v["launches"] for k, v in stats.items()}
    dom = max(tot, key=tot.get)
    dstat = stats[dom]
    achieved = dstat["avg_bytes"] / (dstat["avg_ms"] * 1e-3) / 1e9
    traffic, tsrc = load_traffic(dom, a.config, mixed=head == "mixed")
    step_gbs = rates[head] * STATE_BYTES_PER_POINT / 1e9
    roofline = {"bound": "hbm", "kernel": dom, "achieved": round(achieved, 1),
                "peak": MI355X_HBM_GBS, "unit": "GB/s", "frac": round(achieved / MI355X_HBM_GBS, 4),
                "traffic": traffic, "traffic_source": tsrc,
                "algorithmic_bytes_per_launch": dstat["avg_bytes"],
                "avg_launch_ms": round(dstat["avg_ms"], 4),
                "step_share": round(tot[dom] / sum(tot.values()), 3),
                "step_state_GBps": round(step_gbs, 1),
                "step_frac": round(step_gbs / MI355X_HBM_GBS, 4)}
    kernels = {}
    for name in ("plain", "mixed"):
        st = res[name][1]
        t2 = {k: v["avg_ms"] * v["launches"] for k, v in st.items()}
        kernels[name] = {k: {"avg_ms": round(v["avg_ms"], 4), "launches_per_step": v["launches"] / a.profile_steps,
                             "GBps": round(v["avg_bytes"] / max(v["avg_ms"], 1e-9) / 1e6, 1)}
                         for k, v in sorted(st.items(), key=lambda kv: -t2[kv[0]])}

    cpu = cpu1 = None
    if not a.no_cpu_baseline:
        log("bench: cpu baseline (all host threads)")
        cpu = cpu_baseline(P, fields, 3 if a.config == "c2" else 1, 0, a.config)
        log("bench: cpu baseline (1 thread)")
        cpu1 = cpu_baseline(P, fields, 1, 1, a.config)

    out = {"metric": "grid-point-steps/s on 1 GPU, mixed/learned vs plain FP64 at fixed error",
           "value": rates[head], "unit": "grid-point-steps/s", "n_gpus": 1, "steps": a.steps,
           "warmup": a.warmup, "ms_per_step": ms[head], "higher_is_better": True,
           "scaling": "weak", "vs_baseline": None, "dtype": "f64+f32" if head == "mixed" else "f64",
           "data": "synthetic",
           "config": {"workload": WORKLOADS.get(a.config, a.config), "grid_points": P.grid_points(),
                      "path": head, "factor_cache": "implicit-operator factors reused while dt is fixed"},
           "plain_fp64": {"value": rates["plain"], "ms_per_step": ms["plain"]},
           "mixed": {"value": rates["mixed"], "ms_per_step": ms["mixed"]},
           "mixed_over_plain": rates["mixed"] / rates["plain"],
           "equal_error": err,
           "roofline": roofline, "cpu_baseline": cpu, "cpu_baseline_1core": cpu1, "kernels": kernels,
           "device_bytes": devbytes}
    print(json.dumps(out))


if __name__ == "__main__":
    main()
