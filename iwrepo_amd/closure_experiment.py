#!/usr/bin/env python3
"""The learned-closure workload: a fine FP64 run coarse-grained into
training samples, the closure trained on the device, then coarse runs with
the closure, with the estimation-model fallback (ifilter = 1: filtering.f:8-93
every step) and with no model, each compared with the coarse-grained fine
run.

Default sizes (``--size c4``): fine 256 x 256 x (16 x 25), Re 1e5, wave 0.15,
2000 steps at dt 5e-4 (config c4f); coarse 64 x 64 x (8 x 13) (config c4c).
Samples are taken every ``--every`` fine steps from step ``--spinup`` to
``--train-end``; the a priori test uses the held-out snapshots after it; the
coarse runs start from the coarse-grained fine state at ``--train-end`` and
run to the end of the fine run.  ``--size small`` runs the same on
32 x 16 x (4 x 17) / 16 x 8 x (2 x 9).  Prints one JSON object.
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
from iwrepo_amd.params import CONFIGS, config_params, make_params  # noqa: E402


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def params(size):
    if size == "c4":
        Pf, Pc = config_params("c4f"), config_params("c4c")
        c = CONFIGS["c4f"]
        return Pf, Pc, c["wave"], c["noise"], c["seed"]
    box = dict(xlen=4 * np.pi, ylen=2 * np.pi, re=1e5, pr=7.0, dt=5e-4, noslip=True)
    return (make_params(nx=32, ny=16, nsubd=4, order=16, **box), make_params(nx=16, ny=8, nsubd=2, order=8, **box),
            0.15, 1e-3, 4)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--size", default="c4")
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--spinup", type=int, default=200)
    ap.add_argument("--train-end", type=int, default=1000)
    ap.add_argument("--every", type=int, default=50)
    ap.add_argument("--train-steps", type=int, default=20000)
    ap.add_argument("--batch", type=int, default=16384)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--lr-end", type=float, default=1e-4)
    a = ap.parse_args()
    Pf, Pc, wave, noise, seed = params(a.size)
    for P in (Pf, Pc):
        P.imaxcalc = 0
    t0 = time.time()
    log("closure: fine initial condition")
    f = initial_condition(Pf, wave, noise, seed)
    fine = rt.Solver(Pf)
    fine.set_initial(*f)
    del f
    coarse = rt.Solver(Pc)
    net = rt.Closure(11)
    held = rt.Closure(12)
    step = 0
    fine.run(a.spinup, counts=False)
    step = a.spinup
    ref = {}     # coarse-grained fine states at the comparison times
    nheld = 0
    while step <= a.steps:
        if step <= a.train_end:
            net.sample(fine, coarse)
        else:
            held.sample(fine, coarse)
            nheld += 1
        if step >= a.train_end:
            ref[step] = [coarse.get_state(q) for q in range(4)] + [coarse.kepe().tolist()]
        if step == a.steps:
            break
        fine.run(a.every, counts=False)
        step += a.every
        log("closure: fine step", step)
    fine.close()
    ntrain = int(rt.lib().iw_closure_dataset(net.h, None, None, 0))
    log("closure: training on", ntrain, "samples")
    t1 = time.time()
    # ten chunks with the learning rate decaying geometrically from lr to
    # lr_end (the Adam state carries over between chunks)
    hist = np.concatenate([net.train(a.train_steps // 10, batch=a.batch,
                                     lr=a.lr * (a.lr_end / a.lr) ** (k / 9.0), seed=100 + k) for k in range(10)])
    train_s = time.time() - t1
    # in-distribution check: a random subset of the training samples
    Xs, Ts = net.dataset()
    sel = np.random.default_rng(0).choice(len(Xs), min(len(Xs), 200000), replace=False)
    ys = net.forward(Xs[sel])
    fit = [float(np.linalg.norm(ys[:, c] - Ts[sel, c]) / np.linalg.norm(Ts[sel, c])) for c in range(9)]
    del Xs, Ts
    # a priori on the held-out snapshots (the coarse instance holds the last)
    _, Th = held.dataset()
    n1 = Pc.nx * Pc.ny * Pc.nzloc * Pc.nsubd
    true = Th[-n1:].reshape(Pc.nzloc * Pc.nsubd, Pc.ny, Pc.nx, 9)
    pred = net.predict(coarse)
    saved = [coarse.get_state(q) for q in range(4)]
    coarse.filtering(1)
    fb = np.stack([coarse.sgs(q)[:, :, :Pc.nx] for q in range(9)], axis=-1)
    for q in range(4):
        coarse.set_state(q, saved[q])

    def err(p, t):
        return [float(np.linalg.norm(p[..., c] - t[..., c]) / np.linalg.norm(t[..., c])) for c in range(9)]
    ap_net, ap_fb = err(pred, true), err(fb, true)
    # a posteriori: coarse runs from the coarse-grained state at train_end
    start = ref[a.train_end]
    phys0 = [coarse.horfft(start[q], 1) for q in range(4)]
    post = {}
    for mode in ("none", "closure", "fallback"):
        P = params(a.size)[1]
        P.imaxcalc = 0
        if mode == "fallback":
            P.ifilter, P.nstepfilter = 1, 1
        c = rt.Solver(P)
        if mode == "closure":
            rt.set_closure(c, net)
        c.set_initial(*phys0)
        series, t2 = [], time.time()
        s = a.train_end
        while s < a.steps:
            c.run(a.every, counts=False)
            s += a.every
            st = [c.get_state(q) for q in range(4)]
            w = ref[s]
            e = float(np.sqrt(sum(np.linalg.norm(x - y) ** 2 for x, y in zip(st, w[:4])) /
                              sum(np.linalg.norm(y) ** 2 for y in w[:4])))
            ke = c.kepe()
            series.append({"step": s, "state_rel_l2": e, "kepe": ke.tolist(), "kepe_ref": w[4],
                           "finite": bool(all(np.all(np.isfinite(x)) for x in st))})
        c.sync()
        post[mode] = {"series": series, "ms_per_step": 1e3 * (time.time() - t2) / max(a.steps - a.train_end, 1)}
        c.close()
        log("closure: coarse run", mode, series[-1]["state_rel_l2"] if series else None)
    out = {"size": a.size, "fine": [Pf.nx, Pf.ny, Pf.nsubd, Pf.nzloc], "coarse": [Pc.nx, Pc.ny, Pc.nsubd, Pc.nzloc],
           "fine_steps": a.steps, "samples": ntrain, "held_out_snapshots": nheld,
           "train": {"steps": a.train_steps, "batch": a.batch, "lr": a.lr, "seconds": train_s,
                     "loss_first": float(hist[0]), "loss_last_100": float(hist[-100:].mean()),
                     "training_set_tau_rel_err": fit},
           "a_priori_tau_rel_err": {"network": ap_net, "estimation_model": ap_fb,
                                    "network_mean": float(np.mean(ap_net)), "estimation_model_mean": float(np.mean(ap_fb)),
                                    "none": 1.0},
           "a_posteriori": post, "wall_s": time.time() - t0}
    print(json.dumps(out))


if __name__ == "__main__":
    main()
