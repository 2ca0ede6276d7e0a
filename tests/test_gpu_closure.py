"""GPU checks of the learned sub-grid closure (iwrepo_amd/csrc/iw_closure.hip)
against its CPU oracle (oracle/mlp_oracle.py, pinned in test_mlp_oracle.py)
and the C oracle's transforms:
* the BF16 matrix-core forward and the FP32 backward of the network;
* the coarse-graining of a fine state and the training samples (SGS tensors
  and features) it yields;
* training on the device lowers the loss, and the trained network predicts
  the SGS tensors of a held-out coarse-grained state better than the
  estimation-model tensors of the fallback path (filtering.f:8-93,
  sgscalc.f:125-219) and than no model;
* coarse runs with the closure stay finite; their error against the
  coarse-grained fine run is reported next to those of the plain coarse run
  and of the fallback."""
import json
import os

import numpy as np
import pytest

from helpers import rel
from iwrepo_amd.ic import initial_condition
from iwrepo_amd.params import make_params

pytestmark = pytest.mark.gpu

BOX = dict(xlen=4 * np.pi, ylen=2 * np.pi, re=1e4, pr=7.0, dt=1e-3, noslip=True, imaxcalc=0)


def pair():
    Pf = make_params(nx=32, ny=16, nsubd=4, order=16, **BOX)
    Pc = make_params(nx=16, ny=8, nsubd=2, order=8, **BOX)
    return Pf, Pc


@pytest.fixture(scope="module")
def rt():
    from iwrepo_amd import runtime
    runtime.lib()
    return runtime


@pytest.fixture(scope="module")
def M():
    import mlp_oracle
    return mlp_oracle


def _random_setup(M, seed):
    rng = np.random.default_rng(seed)
    p = np.zeros(M.NPARAM, dtype=np.float32)
    p[M.OW1:M.OB1] = (rng.standard_normal((M.NH, M.NINP)) * 0.4 * (np.arange(M.NINP) < M.NIN)).ravel()
    p[M.OW2:M.OB2] = rng.standard_normal(M.NH * M.NH) * 0.15
    p[M.OW3:M.OB3] = rng.standard_normal(M.NH * M.NH) * 0.15
    p[M.OW4:M.OB4] = (rng.standard_normal((M.NOUTP, M.NH)) * 0.2 * (np.arange(M.NOUTP) < M.NOUT)[:, None]).ravel()
    for a, b in ((M.OB1, M.OW2), (M.OB2, M.OW3), (M.OB3, M.OW4)):
        p[a:b] = rng.standard_normal(b - a) * 0.1
    p[M.OB4:M.OB4 + M.NOUT] = rng.standard_normal(M.NOUT) * 0.1
    norm = np.concatenate([rng.standard_normal(M.NIN) * 0.1, 0.5 + rng.random(M.NIN),
                           rng.standard_normal(M.NOUT) * 0.1, 0.5 + rng.random(M.NOUT)]).astype(np.float32)
    return p, norm, rng


def test_network_forward_and_gradient_match_oracle(rt, M):
    """Forward within 2e-3 (BF16 rounding points emulated by the oracle, FP32
    vs FP64 accumulation), gradient of the loss within 2e-3 relative L2."""
    p, norm, rng = _random_setup(M, 5)
    net = rt.Closure(1)
    net.set_params(np.concatenate([p, norm]))
    assert np.array_equal(net.params(), np.concatenate([p, norm]))
    X = rng.standard_normal((3000, M.NIN)).astype(np.float32)
    T = rng.standard_normal((3000, M.NOUT)).astype(np.float32)
    y = net.forward(X)
    yo = M.predict(p, norm, X)
    print("forward rel", rel(y, yo), flush=True)
    assert rel(y, yo) < 2e-3
    loss, g = net.gradient(X, T)
    lo, go = M.loss_grad(p, norm, X, T)
    print("loss", loss, lo, "grad rel", rel(g, go), flush=True)
    assert abs(loss - lo) < 2e-3 * lo
    assert rel(g, go) < 2e-3
    for a, b in ((M.OW1, M.OB1), (M.OW2, M.OB2), (M.OW3, M.OB3), (M.OW4, M.OB4)):
        assert rel(g[a:b], go[a:b]) < 5e-3, a
    net.close()


def _fine_run(rt, Pf, steps, seed=4):
    f = initial_condition(Pf, 0.15, 1e-3, seed)
    s = rt.Solver(Pf)
    s.set_initial(*f)
    if steps:
        s.run(steps, counts=False)
    return s


def _oracle_targets_and_inputs(O, M, Pf, Pc, fine_spec):
    of, oc = O.Oracle(Pf), O.Oracle(Pc)
    phys = [of.horfft(fine_spec[q], 1) for q in range(4)]
    u, v, w, t = phys
    prods = [u * u, u * v, u * w, v * v, v * w, w * w, u * t, v * t, w * t]
    cstate = [M.coarsen(fine_spec[q], Pf, Pc) for q in range(4)]
    cphys = [oc.horfft(c, 1) for c in cstate]
    fprod = [oc.horfft(M.coarsen(of.horfft(pq, -1), Pf, Pc), 1) for pq in prods]
    cu, cv, cw, ct = cphys
    fp = [cu * cu, cu * cv, cu * cw, cv * cv, cv * cw, cw * cw, cu * ct, cv * ct, cw * ct]
    nx = Pc.nx
    tau = np.stack([(a - b)[:, :, :nx] for a, b in zip(fprod, fp)], axis=-1)
    # features: spectral x/y derivatives, dpdz in z
    al, be = 2 * np.pi / Pc.xlen, 2 * np.pi / Pc.ylen
    kx = np.arange(Pc.nx // 2 + 1) * al
    ky = np.array([j if j < Pc.ny // 2 else j - Pc.ny for j in range(Pc.ny)]) * be

    def deriv(spec, k, axis):
        c = spec[:, :, 0::2] + 1j * spec[:, :, 1::2]
        c = c * (1j * (k[None, None, :] if axis == "x" else k[None, :, None]))
        out = np.zeros_like(spec)
        out[:, :, 0::2], out[:, :, 1::2] = c.real, c.imag
        return oc.horfft(out, 1)
    g = []
    for q in range(3):
        g += [deriv(cstate[q], kx, "x"), deriv(cstate[q], ky, "y"), oc.dpdz(cphys[q])]
    rx, ry, rz = deriv(cstate[3], kx, "x"), deriv(cstate[3], ky, "y"), oc.dpdz(cphys[3])
    z = oc.table(7, Pc.nzloc * Pc.nsubd)
    feats = g + [rz, np.sqrt(rx * rx + ry * ry), np.broadcast_to(z[:, None, None], rz.shape)]
    X = np.stack([f[:, :, :nx] for f in feats], axis=-1)
    return cstate, tau, X


def test_coarse_graining_and_samples_match_oracle(rt, oracle_mod, M):
    Pf, Pc = pair()
    fine = _fine_run(rt, Pf, 3)
    spec = [fine.get_state(q) for q in range(4)]
    coarse = rt.Solver(Pc)
    net = rt.Closure(2)
    print("closure sample", flush=True)
    net.sample(fine, coarse)
    X, T = net.dataset()
    cstate, tau, Xo = _oracle_targets_and_inputs(oracle_mod, M, Pf, Pc, spec)
    for q in range(4):
        assert rel(coarse.get_state(q), cstate[q]) < 1e-12, q
    n = Pc.nx * Pc.ny * Pc.nzloc * Pc.nsubd
    assert X.shape == (n, 12) and T.shape == (n, 9)
    T = T.reshape(tau.shape)
    X = X.reshape(Xo.shape)
    for c in range(9):
        assert rel(T[..., c], tau[..., c]) < 1e-5, c
    for c in range(12):
        assert rel(X[..., c], Xo[..., c]) < 1e-5, c
    fine.close()
    coarse.close()
    net.close()


def _sgs_fallback(coarse):
    """The estimation model's tensors of the coarse state (filtering(1),
    sgscalc.f:125-219) at the sample points; the state is restored."""
    saved = [coarse.get_state(q) for q in range(4)]
    coarse.filtering(1)
    P = coarse.P
    tau = np.stack([coarse.sgs(q)[:, :, :P.nx] for q in range(9)], axis=-1)
    for q in range(4):
        coarse.set_state(q, saved[q])
    return tau


def _tau_err(pred, true):
    return float(np.mean([rel(pred[..., c], true[..., c]) for c in range(9)]))


def _record(tag, data):
    if os.path.isdir("gpurun_out"):
        with open(os.path.join("gpurun_out", tag + ".json"), "w") as f:
            json.dump(data, f, indent=1)


def test_training_and_a_priori_gain_over_fallback(rt):
    Pf, Pc = pair()
    Pc.nstepfilter = 1
    fine = _fine_run(rt, Pf, 5)
    coarse = rt.Solver(Pc)
    net = rt.Closure(3)
    for k in range(10):  # training snapshots every 4 steps
        net.sample(fine, coarse)
        fine.run(4, counts=False)
    fine.run(10, counts=False)
    held = rt.Closure(4)  # held-out snapshot: its targets
    held.sample(fine, coarse)
    _, Th = held.dataset()
    true = Th.reshape(Pc.nzloc * Pc.nsubd, Pc.ny, Pc.nx, 9)
    print("training", net.dataset()[0].shape, flush=True)
    hist = net.train(600, batch=4096, lr=2e-3)
    print("loss", hist[0], hist[-1], flush=True)
    assert np.all(np.isfinite(hist))
    assert hist[-50:].mean() < 0.5 * hist[:10].mean()
    pred = net.predict(coarse)
    fb = _sgs_fallback(coarse)
    e_net, e_fb, e_zero = _tau_err(pred, true), _tau_err(fb, true), 1.0
    print("a priori tau error: network", e_net, "estimation model", e_fb, "none", e_zero, flush=True)
    _record("closure_a_priori_small", {"network": e_net, "estimation_model": e_fb, "none": e_zero,
                                       "loss_first": float(hist[0]), "loss_last": float(hist[-1])})
    assert e_net < e_fb and e_net < e_zero
    # a posteriori: coarse runs from the coarse-grained state against the
    # coarse-grained fine run 20 steps later
    phys0 = [coarse.horfft(coarse.get_state(q), 1) for q in range(4)]
    fine.run(20, counts=False)
    ref = rt.Solver(Pc)
    rt.coarse_grain(fine, ref)
    want = [ref.get_state(q) for q in range(4)]
    errs = {}
    for mode in ("plain", "closure", "fallback"):
        P = pair()[1]
        if mode == "fallback":
            P.ifilter, P.nstepfilter = 1, 1
        c = rt.Solver(P)
        if mode == "closure":
            rt.set_closure(c, net)
        c.set_initial(*phys0)
        c.run(20, counts=False)
        st = [c.get_state(q) for q in range(4)]
        assert all(np.all(np.isfinite(x)) for x in st), mode
        errs[mode] = float(np.sqrt(sum(np.linalg.norm(a - b) ** 2 for a, b in zip(st, want)) /
                                   sum(np.linalg.norm(b) ** 2 for b in want)))
        c.close()
    print("a posteriori state error vs the coarse-grained fine run:", errs, flush=True)
    _record("closure_a_posteriori_small", errs)
    for s in (fine, coarse, ref):
        s.close()
    net.close()
    held.close()
