"""GPU checks of the learned warm start: the BF16 matrix-core GEMM, the MLP
gradients against the CPU oracle (same bfloat16 rounding points, FP64
accumulation), training, and what the trained network gains over the flat
start on the 944-bus network.

Tolerances: GEMM |dC| <= 1e-4 max|C| (FP32 vs FP64 accumulation of the same
bf16 operands); gradients: relative L2 error <= 2e-2 and cosine >= 0.999
(a bf16 rounding decision can flip where FP32 and FP64 pre-activations
straddle a rounding boundary)."""
import os
import sys

import numpy as np
import pytest

from conftest import ROOT

pytestmark = pytest.mark.gpu
sys.path.insert(0, os.path.join(ROOT, "oracle"))


def _say(*a):
    print("[gpu-mlp]", *a, flush=True)


@pytest.mark.parametrize("ta,tb", [(False, False), (True, False), (False, True)])
def test_gemm_bf16(pf, ta, tb):
    import mlp_oracle as M
    w = pf.WarmStart(16, 16, 16, 1)
    rng = np.random.default_rng(3)
    Mm, N, K = 72, 136, 200
    A = rng.standard_normal((K, Mm) if ta else (Mm, K)).astype(np.float32)
    B = rng.standard_normal((N, K) if tb else (K, N)).astype(np.float32)
    C = w.gemm_probe(A, B, ta, tb)
    Aq, Bq = M.bf16(A), M.bf16(B)
    ref = (Aq.T if ta else Aq) @ (Bq.T if tb else Bq)
    assert C.shape == ref.shape
    assert np.abs(C - ref).max() <= 1e-4 * np.abs(ref).max()


def test_mlp_gradient_vs_oracle(pf):
    import mlp_oracle as M
    w = pf.WarmStart(40, 24, 64, 2, seed=5)
    rng = np.random.default_rng(4)
    X = rng.standard_normal((32, 40)).astype(np.float32) * 2 + 1
    T = rng.standard_normal((32, 24)).astype(np.float32)
    xm, xs = X.mean(0), X.std(0)
    ym, ys = T.mean(0), T.std(0)
    w.set_norm(xm, xs, ym, ys)
    p = w.params()
    g, loss = w.grad(X, T)
    lo, go = M.loss_and_grad(p.astype(np.float64), w.layer_shapes(), X, T, xm, xs, ym, ys, quantize=True)
    assert abs(loss - lo) <= 1e-4 * lo
    rel = np.linalg.norm(g - go) / np.linalg.norm(go)
    cos = float(g @ go / (np.linalg.norm(g) * np.linalg.norm(go)))
    _say("grad rel err", rel, "cos", cos)
    assert rel <= 2e-2 and cos >= 0.999
    # prediction = oracle forward
    Yp = w.predict(X)
    Yo = M.predict(p.astype(np.float64), w.layer_shapes(), X, xm, xs, ym, ys, quantize=True)
    assert np.abs(Yp - Yo).max() <= 2e-2 * np.abs(Yo).max()


def test_mlp_training_reduces_loss(pf):
    rng = np.random.default_rng(6)
    X = rng.standard_normal((1024, 32)).astype(np.float32)
    Wt = rng.standard_normal((32, 16)).astype(np.float32) / 6
    T = np.tanh(X @ Wt).astype(np.float32)
    w = pf.WarmStart(32, 16, 64, 2, seed=2)
    w.set_norm(X.mean(0), X.std(0), T.mean(0), T.std(0))
    losses = w.fit(X, T, epochs=20, batch=64, lr=2e-3)
    _say("losses", losses[0], losses[-1])
    assert losses[-1] < 0.25 * losses[0]


def test_warm_start_gains_on_944(pf, synth, cases):
    """Train on NR solutions the batch solver produces, then compare NR
    iterations from the learned start and from the flat start on held-out
    scenarios (same tolerance, flat-start fallback on)."""
    net = pf.Network(cases.path("c944"))
    is_pv = np.zeros(net.nB, bool)
    is_pv[net.pv] = True
    vset = np.where(is_pv, net.V, 0.0)
    vset[net.slack] = net.V[net.slack]
    opts = pf.options(tolerancia=1e-8, no_max_iter=60, lim_inj_reat=1, precision=pf.MIXED)

    def solve(seeds, V0=None, T0=None):
        P, Q, L = synth.scenario_batch(net.Pg, net.Pload, net.Qg, net.Qload, is_pv, seeds)
        b = pf.Batch([net], opts)
        b.set_scenarios(np.zeros(len(seeds), np.int32), P.ravel(), Q.ravel(), L.ravel(),
                        None if V0 is None else V0.ravel(), None if T0 is None else T0.ravel())
        ms = b.solve()
        V, th, it, cv, mm = b.results()
        return P, Q, V.reshape(len(seeds), -1), th.reshape(len(seeds), -1), it, cv, ms

    train = np.arange(4000000, 4002048)
    P, Q, V, th, it, cv, _ = solve(train)
    assert cv.all()
    X = np.concatenate([P, Q, np.repeat(vset[None, :], len(P), 0)], 1).astype(np.float32)
    Y = np.concatenate([V, th], 1).astype(np.float32)
    w = pf.WarmStart(3 * net.nB, 2 * net.nB, 512, 3, seed=11)
    w.set_norm(X.mean(0), X.std(0), Y.mean(0), Y.std(0))
    losses = w.fit(X, Y, epochs=30, batch=128, lr=1e-3, weight_decay=1e-5)
    _say("warm-start training loss", losses[0], losses[-1])
    test = np.arange(5000000, 5000256)
    Pt, Qt, L = synth.scenario_batch(net.Pg, net.Pload, net.Qg, net.Qload, is_pv, test)[:3]
    Xt = np.concatenate([Pt, Qt, np.repeat(vset[None, :], len(Pt), 0)], 1).astype(np.float32)
    Yt = w.predict(Xt)
    _, _, Vf, thf, itf, cvf, msf = solve(test)
    _, _, Vw, thw, itw, cvw, msw = solve(test, Yt[:, :net.nB].astype(np.float64), Yt[:, net.nB:].astype(np.float64))
    _say("iterations flat", itf.mean(), "warm", itw.mean(), "ms flat", msf, "warm", msw)
    assert cvf.all() and cvw.all()
    assert itw.mean() < itf.mean()


def test_warm_start_predict_into_batch(pf, synth):
    """Inference on features staged on the GPU, written straight into a
    batch's starting state (no host round trip), solves every scenario to
    the same bits as the host path (predict, then the start passed to
    pf_batch_set_scenarios).  The network is the 118-bus case, a short fit
    on 512 solutions; a predict call in between invalidates the staged rows."""
    import os
    net = pf.Network(os.path.join(os.path.dirname(__file__), "golden", "c118.cdf"))
    is_pv = np.zeros(net.nB, bool)
    is_pv[net.pv] = True
    vset = np.where(is_pv, net.V, 0.0)
    vset[net.slack] = net.V[net.slack]
    opts = pf.options(tolerancia=1e-8, no_max_iter=60, lim_inj_reat=0, precision=pf.MIXED)
    n_in, n_out = 3 * net.nB, 2 * net.nB
    n_in8, n_out8 = -(-n_in // 8) * 8, -(-n_out // 8) * 8

    def feats(P, Q):
        X = np.zeros((len(P), n_in8), np.float32)
        X[:, :n_in] = np.concatenate([P, Q, np.repeat(vset[None, :], len(P), 0)], 1)
        return X

    P, Q, L = synth.scenario_batch(net.Pg, net.Pload, net.Qg, net.Qload, is_pv, np.arange(6100000, 6100512))[:3]
    b = pf.Batch([net], opts)
    b.set_scenarios(np.zeros(len(P), np.int32), P.ravel(), Q.ravel(), L.ravel())
    b.solve()
    V, th, it, cv, mm = b.results()
    assert cv.all()
    X = feats(P, Q)
    Y = np.zeros((len(P), n_out8), np.float32)
    Y[:, :n_out] = np.concatenate([V.reshape(len(P), -1), th.reshape(len(P), -1)], 1)
    w = pf.WarmStart(n_in8, n_out8, 256, 2, seed=5)
    w.set_norm(X.mean(0), X.std(0) + 1e-6, Y.mean(0), Y.std(0) + 1e-9)
    w.fit(X, Y, epochs=4, batch=128, lr=1e-3, weight_decay=1e-5)

    Pt, Qt, Lt = synth.scenario_batch(net.Pg, net.Pload, net.Qg, net.Qload, is_pv, np.arange(6200000, 6200256))[:3]
    Xt = feats(Pt, Qt)
    Yt = w.predict(Xt)
    host = pf.Batch([net], opts)
    host.set_scenarios(np.zeros(len(Pt), np.int32), Pt.ravel(), Qt.ravel(), Lt.ravel(),
                       Yt[:, :net.nB].astype(np.float64).ravel(), Yt[:, net.nB:n_out].astype(np.float64).ravel())
    host.solve()
    r_host = host.results()

    dev = pf.Batch([net], opts)
    dev.set_scenarios(np.zeros(len(Pt), np.int32), Pt.ravel(), Qt.ravel(), Lt.ravel())
    dev.solve()
    r_flat = dev.results()
    w.stage(Xt)
    w.predict_into(dev, 0, net.nB)
    dev.solve()
    r_dev = dev.results()
    _say("warm into batch: iterations flat", r_flat[2].mean(), "warm", r_dev[2].mean())
    for x, y in zip(r_host, r_dev):
        assert np.array_equal(x, y)
    assert r_dev[3].all() and r_dev[2].mean() < r_flat[2].mean()
    # a predict call replaces the staged rows: predict_into refuses until the next stage
    w.predict(Xt[:8])
    with pytest.raises(pf.PfError):
        w.predict_into(dev, 0, net.nB)
    # a row narrower than a scenario's buses is refused
    w.stage(Xt)
    with pytest.raises(pf.PfError):
        w.predict_into(dev, 0, n_out8 - net.nB + 1)
