"""GPU parity of the batched persistent solver (one workgroup per scenario)
against the oracle solving each scenario on its own, and of its warm start /
flat-start fallback.  Tolerance: |dV| <= 1e-6 pu, |dtheta| <= 1e-6 rad."""
import os
import sys

import numpy as np
import pytest

from conftest import GOLDEN

pytestmark = pytest.mark.gpu


def _say(*a):
    print("[gpu-batch]", *a, flush=True)
    sys.stdout.flush()


def _scenarios(synth, net, n, seed0):
    """Perturbed injections: per-bus load factors clip(exp(0.15 N), 0.7, 1.3),
    PV dispatch scaled by the ratio of total loads."""
    is_pv = np.zeros(net.nB, bool)
    is_pv[net.pv] = True
    return synth.scenario_batch(net.Pg, net.Pload, net.Qg, net.Qload, is_pv, np.arange(seed0, seed0 + n))


def _oracle_solve(oc, P, Q, L, lim=True):
    out = []
    for s in range(len(P)):
        out.append(oc.solve(tol=1e-8, lim=lim, Pliq=P[s], Qliq=Q[s], Qload=L[s], reuse_symbolic=True))
    return out


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_batch_uneven(pf, oracle, synth, cases, precision):
    p118 = os.path.join(GOLDEN, "c118.cdf")
    p944 = cases.path("c944")
    n118, n944 = pf.Network(p118), pf.Network(p944)
    o118, o944 = oracle.Case(p118), oracle.Case(p944)
    A = _scenarios(synth, n118, 24, 3000000)
    B = _scenarios(synth, n944, 4, 3100000)
    net_of = np.array([0] * 24 + [1] * 4, np.int32)
    order = np.random.default_rng(0).permutation(len(net_of))   # interleave the two sizes
    net_of = net_of[order]
    rows = {0: list(range(24)), 1: list(range(4))}
    P, Q, L = [], [], []
    idx = {0: 0, 1: 0}
    src = {0: A, 1: B}
    picked = []
    for k in net_of:
        i = rows[k][idx[k]]
        idx[k] += 1
        picked.append((k, i))
        P.append(src[k][0][i]); Q.append(src[k][1][i]); L.append(src[k][2][i])
    opts = pf.options(tolerancia=1e-8, no_max_iter=60, lim_inj_reat=1,
                      precision=pf.MIXED if precision == "mixed" else pf.FP64)
    _say("batch", precision, len(net_of))
    b = pf.Batch([n118, n944], opts)
    b.set_scenarios(net_of, np.concatenate(P), np.concatenate(Q), np.concatenate(L))
    ms = b.solve()
    V, th, it, cv, mm = b.results()
    lq, fl = b.limq()
    _say("elapsed ms", ms, "iters", it.tolist())
    off = 0
    for s, (k, i) in enumerate(picked):
        nb = (n118, n944)[k].nB
        oc = (o118, o944)[k]
        r = oc.solve(tol=1e-8, lim=True, Pliq=src[k][0][i], Qliq=src[k][1][i], Qload=src[k][2][i],
                     reuse_symbolic=True)
        assert r["converged"]
        assert cv[s], (s, k, i, it[s], mm[s])
        assert mm[s] <= 1e-8
        assert np.abs(V[off:off + nb] - r["V"]).max() <= 1e-6
        assert np.abs(th[off:off + nb] - r["theta"]).max() <= 1e-6
        assert np.array_equal(lq[s], r["limQ"]), (s, k, i)
        if precision == "fp64":
            assert it[s] == r["iterations"]
            assert fl[s] == r["n_switch_events"]
        off += nb
    # replaying the captured graph gives the same answer
    b.solve()
    V2, th2, it2, cv2, _ = b.results()
    assert np.array_equal(V, V2) and np.array_equal(it, it2)


def test_batch_warm_start_flat_parity_without_limits(pf, oracle, synth, cases):
    """Without Q limits the power-flow answer does not depend on the start:
    every warm-started scenario must reproduce the FLAT-start oracle (the
    reference path, CPU/estrutura/barra.cpp:20-25) within 1e-6 pu / 1e-6 rad,
    in fewer iterations."""
    for name, n in (("c118", 16), ("c944", 6)):
        p = os.path.join(GOLDEN, "c118.cdf") if name == "c118" else cases.path(name)
        net = pf.Network(p)
        oc = oracle.Case(p)
        P, Q, L = _scenarios(synth, net, n, 3250000)
        ref = _oracle_solve(oc, P, Q, L, lim=False)
        opts = pf.options(tolerancia=1e-8, no_max_iter=60, lim_inj_reat=0, precision=pf.MIXED)
        b = pf.Batch([net], opts)
        net_of = np.zeros(n, np.int32)
        b.set_scenarios(net_of, P.ravel(), Q.ravel(), L.ravel())
        b.solve()
        _, _, it_flat, cv_flat, _ = b.results()
        rng = np.random.default_rng(5)
        V0 = np.concatenate([r["V"] for r in ref]) + 5e-3 * rng.standard_normal(n * net.nB)
        T0 = np.concatenate([r["theta"] for r in ref]) + 5e-3 * rng.standard_normal(n * net.nB)
        b.set_scenarios(net_of, P.ravel(), Q.ravel(), L.ravel(), V0, T0)
        b.solve()
        V, th, it, cv, mm = b.results()
        nb = net.nB
        dv = max(float(np.abs(V[nb * s:nb * (s + 1)] - ref[s]["V"]).max()) for s in range(n))
        dt = max(float(np.abs(th[nb * s:nb * (s + 1)] - ref[s]["theta"]).max()) for s in range(n))
        _say("warm (no limits)", name, "iters flat", it_flat.tolist(), "warm", it.tolist(), "dV", dv, "dth", dt)
        assert cv_flat.all() and cv.all() and (mm <= 1e-8).all()
        assert all(r["converged"] for r in ref)
        assert dv <= 1e-6 and dt <= 1e-6
        assert it.mean() < it_flat.mean()


def test_batch_warm_start_and_fallback(pf, oracle, synth):
    """With Q limits on, the answer depends on the switching path: a bus
    switched back to PV keeps the |V| it reached while limited
    (CPU/limInjReat.cpp:62-87), and every scenario here switches back at
    least once.  So a warm-started solve is checked against the oracle run
    from the same start (the same rules applied along the same path), and
    the flat-start fallback against the flat-start oracle."""
    p118 = os.path.join(GOLDEN, "c118.cdf")
    net = pf.Network(p118)
    oc = oracle.Case(p118)
    P, Q, L = _scenarios(synth, net, 8, 3200000)
    ref = _oracle_solve(oc, P, Q, L)
    opts = pf.options(tolerancia=1e-8, no_max_iter=60, lim_inj_reat=1, precision=pf.MIXED)
    b = pf.Batch([net], opts)
    net_of = np.zeros(8, np.int32)
    b.set_scenarios(net_of, P.ravel(), Q.ravel(), L.ravel())
    b.solve()
    _, _, it_flat, cv_flat, _ = b.results()
    assert cv_flat.all()
    # warm start near the oracle's answer: fewer iterations than the flat start.
    # The start sets |V| of PQ buses and all non-slack angles; PV buses begin at
    # their setpoints.  Back-switched buses keep the |V| they reached while
    # limited (CPU/limInjReat.cpp:62-87), so the answer depends on the start:
    # compare with the oracle started from the same state.
    rng = np.random.default_rng(9)
    V0 = np.concatenate([r["V"] for r in ref]) + 1e-3 * rng.standard_normal(8 * 118)
    T0 = np.concatenate([r["theta"] for r in ref]) + 1e-3 * rng.standard_normal(8 * 118)
    b.set_scenarios(net_of, P.ravel(), Q.ravel(), L.ravel(), V0, T0)
    b.solve()
    V, th, it, cv, mm = b.results()
    lq_w, _ = b.limq()
    assert cv.all() and (mm <= 1e-8).all()
    assert it.mean() < it_flat.mean()
    is_pq = np.zeros(118, bool)
    is_pq[net.pq] = True
    for s in range(8):
        v0 = np.where(is_pq, V0[118 * s:118 * (s + 1)], net.V)
        t0 = T0[118 * s:118 * (s + 1)].copy()
        t0[net.slack] = net.theta[net.slack]
        r = oc.solve(tol=1e-8, lim=True, Pliq=P[s], Qliq=Q[s], Qload=L[s], V0=v0, theta0=t0,
                     reuse_symbolic=True)
        assert r["converged"]
        assert np.abs(V[118 * s:118 * (s + 1)] - r["V"]).max() <= 1e-6
        assert np.abs(th[118 * s:118 * (s + 1)] - r["theta"]).max() <= 1e-6
        assert np.array_equal(lq_w[s], r["limQ"])
    # a useless start (|V| = 0.05, random angles) diverges and falls back to
    # flat start: then the answer and the decisions are the flat-start oracle's
    bad_V = np.full_like(V0, 0.05)
    bad_T = np.random.default_rng(1).uniform(-3, 3, T0.shape)
    b.set_scenarios(net_of, P.ravel(), Q.ravel(), L.ravel(), bad_V, bad_T)
    b.solve()
    V, th, it, cv, mm = b.results()
    lq, fl = b.limq()
    assert cv.all()
    for s in range(8):
        assert np.abs(V[118 * s:118 * (s + 1)] - ref[s]["V"]).max() <= 1e-6
        assert np.abs(th[118 * s:118 * (s + 1)] - ref[s]["theta"]).max() <= 1e-6
        assert np.array_equal(lq[s], ref[s]["limQ"])


def test_batch_autotune(pf, synth, cases, tmp_path):
    """The autotuner changes only launch shapes: every candidate does the same
    arithmetic in the same order (per-row and per-node lane groups do not
    depend on the workgroup size or on where the working set lives), so the
    tuned answers equal the default-shape answers bit for bit; a second batch
    with the same cache applies the cached shapes without measuring."""
    p118 = os.path.join(GOLDEN, "c118.cdf")
    n118, n944 = pf.Network(p118), pf.Network(cases.path("c944"))
    A = _scenarios(synth, n118, 64, 3000000)
    B = _scenarios(synth, n944, 8, 3200000)
    net_of = np.array([1] * 8 + [0] * 64, np.int32)
    P = np.concatenate([B[0].ravel(), A[0].ravel()])
    Q = np.concatenate([B[1].ravel(), A[1].ravel()])
    L = np.concatenate([B[2].ravel(), A[2].ravel()])
    opts = pf.options(tolerancia=1e-8, no_max_iter=60, lim_inj_reat=1, precision=pf.MIXED)
    cache = str(tmp_path / "tune.txt")

    def run(tune):
        b = pf.Batch([n118, n944], opts)
        b.set_scenarios(net_of, P, Q, L)
        n = b.autotune(cache, reps=1) if tune else 0
        b.solve()
        return b, n, b.results()

    _say("autotune default shapes")
    b0, _, r0 = run(False)
    _say("default", b0.shape(0), b0.shape(1))
    _say("autotune measuring")
    b1, n1, r1 = run(True)
    _say("tuned", b1.shape(0), b1.shape(1), "candidates", n1)
    assert n1 >= 2
    for x, y in zip(r0, r1):
        assert np.array_equal(x, y)
    _say("autotune from cache")
    b2, n2, r2 = run(True)
    assert n2 == 0
    assert b2.shape(0) == b1.shape(0) and b2.shape(1) == b1.shape(1)
    for x, y in zip(r0, r2):
        assert np.array_equal(x, y)
    assert r0[3].all()


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_batch_every_shape(pf, oracle, synth, cases, precision):
    """Every launch shape (working set in LDS / step vectors and level data in
    LDS / all global; workgroup size) and both LU modes (the dense separator
    block on or off) solve each scenario like the oracle.  All results are
    printed before the first assertion."""
    p118, p944 = os.path.join(GOLDEN, "c118.cdf"), cases.path("c944")
    n118, n944 = pf.Network(p118), pf.Network(p944)
    o118, o944 = oracle.Case(p118), oracle.Case(p944)
    A = _scenarios(synth, n118, 6, 3300000)
    B = _scenarios(synth, n944, 2, 3400000)
    ref = [o118.solve(tol=1e-8, lim=True, Pliq=A[0][s], Qliq=A[1][s], Qload=A[2][s], reuse_symbolic=True)
           for s in range(6)]
    ref += [o944.solve(tol=1e-8, lim=True, Pliq=B[0][s], Qliq=B[1][s], Qload=B[2][s], reuse_symbolic=True)
            for s in range(2)]
    net_of = np.array([0] * 6 + [1] * 2, np.int32)
    P = np.concatenate([A[0].ravel(), B[0].ravel()])
    Q = np.concatenate([A[1].ravel(), B[1].ravel()])
    L = np.concatenate([A[2].ravel(), B[2].ravel()])
    rows = []
    answers = {}
    for lu_blocked in (0, 1):
        opts = pf.options(tolerancia=1e-8, no_max_iter=60, lim_inj_reat=1, lu_blocked=lu_blocked,
                          precision=pf.MIXED if precision == "mixed" else pf.FP64)
        b = pf.Batch([n118, n944], opts)
        b.set_scenarios(net_of, P, Q, L)
        for mode in (2, 1, 0):
            for block in (256, 1024):
                try:
                    b.set_shape(0, mode, block)
                    b.set_shape(1, mode if mode < 2 else 0, block)
                except pf.PfError:
                    continue
                b.solve()
                V, th, it, cv, mm = b.results()
                lq, _ = b.limq()
                dv = dt = 0.0
                off = 0
                same_lq = True
                for s, r in enumerate(ref):
                    nb = (n118, n944)[net_of[s]].nB
                    if cv[s] and r["converged"]:
                        dv = max(dv, float(np.abs(V[off:off + nb] - r["V"]).max()))
                        dt = max(dt, float(np.abs(th[off:off + nb] - r["theta"]).max()))
                    same_lq &= bool(np.array_equal(lq[s], r["limQ"]))
                    off += nb
                answers.setdefault(lu_blocked, []).append((V.copy(), th.copy()))
                row = (lu_blocked, b.shape(0)["mode"], b.shape(1)["mode"], block, int(cv.sum()), dv, dt,
                       it.tolist(), [r["iterations"] for r in ref], same_lq)
                _say("lu_blocked %d modes %d/%d block %d converged %d dV %.3g dth %.3g iters %s oracle %s "
                     "same limQ %s" % row)
                rows.append(row)
    assert rows
    for row in rows:
        assert row[4] == len(ref) and row[5] <= 1e-6 and row[6] <= 1e-6 and row[9], row
    # every launch shape of one schedule gives the same answer bit for bit
    for lb, sols in answers.items():
        for V, th in sols[1:]:
            assert np.array_equal(V, sols[0][0]) and np.array_equal(th, sols[0][1]), lb


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_batch_1888(pf, oracle, synth, cases, precision):
    """The 1,888-bus (16-copy) scenarios of the uneven benchmark batch, in the
    launch shape the autotuner picks for them (step vectors and level data in
    LDS, the factor in a global slice, 1024 threads), against the oracle:
    V / theta within 1e-6, the same final PV->PQ set, and in FP64 the same
    iteration and switch counts."""
    p = cases.path("c1888")
    net = pf.Network(p)
    oc = oracle.Case(p)
    n = 10
    P, Q, L = _scenarios(synth, net, n, 3000000 + 2048)
    ref = _oracle_solve(oc, P, Q, L)
    opts = pf.options(tolerancia=1e-8, no_max_iter=60, lim_inj_reat=1,
                      precision=pf.MIXED if precision == "mixed" else pf.FP64)
    b = pf.Batch([net], opts)
    b.set_scenarios(np.zeros(n, np.int32), P.ravel(), Q.ravel(), L.ravel())
    b.autotune(None, reps=1)
    shape = b.shape(0)
    _say("1888", precision, "shape", shape)
    b.solve()
    V, th, it, cv, mm = b.results()
    lq, fl = b.limq()
    nb = net.nB
    rows = []
    for s in range(n):
        r = ref[s]
        dv = float(np.abs(V[nb * s:nb * (s + 1)] - r["V"]).max())
        dt = float(np.abs(th[nb * s:nb * (s + 1)] - r["theta"]).max())
        rows.append((s, bool(cv[s]), r["converged"], int(it[s]), r["iterations"], int(fl[s]),
                     r["n_switch_events"], dv, dt, bool(np.array_equal(lq[s], r["limQ"]))))
        _say("scenario %d conv %s/%s iters %d/%d flips %d/%d dV %.3g dth %.3g same limQ %s" % rows[-1])
    assert shape["mode"] == 1
    assert sum(r[2] for r in rows) >= n - 2
    for row in rows:
        assert row[1] == row[2], row
        assert row[9], row
        if row[2]:
            assert row[7] <= 1e-6 and row[8] <= 1e-6, row
        if precision == "fp64":
            assert row[3] == row[4] and row[5] == row[6], row


def test_batch_global_modes_workgroup_sizes(pf, synth, cases):
    """The placements outside LDS (modes 1 and 0) take at least 256
    threads: their sweeps give each node of a narrow level (up to 16 nodes)
    its own 16 lanes.  Smaller workgroups are refused; the sizes that run
    give the same bits."""
    n944 = pf.Network(cases.path("c944"))
    P, Q, L = _scenarios(synth, n944, 3, 3700000)
    b = pf.Batch([n944], pf.options(tolerancia=1e-8, no_max_iter=60, lim_inj_reat=1, precision=pf.MIXED))
    b.set_scenarios(np.zeros(3, np.int32), P.ravel(), Q.ravel(), L.ravel())
    for mode in (1, 0):
        for block in (64, 128, 192):
            with pytest.raises(pf.PfError):
                b.set_shape(0, mode, block)
    outs = []
    for mode in (1, 0):
        for block in (256, 320, 512, 1024):
            b.set_shape(0, mode, block)
            _say("c944 mode", mode, "block", block)
            b.solve()
            outs.append(b.results())
    assert outs[0][3].all()
    for o in outs[1:]:
        for x, y in zip(outs[0], o):
            assert np.array_equal(x, y)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_batch_odd_size_network(pf, oracle, synth, precision):
    """A network with an odd bus count (IEEE 57-bus): the per-scenario working
    set's complex arrays sit after 5 nB doubles, so they must be realigned to
    16 bytes.  Every placement (LDS, step vectors in LDS, all global) solves
    each scenario like the oracle.  The two placements in global memory are
    refused for a working set that fits in LDS, as this one does, so the
    LDS placement carries the check; all placements share the layout code
    (carve)."""
    p57 = os.path.join(GOLDEN, "ieee", "ieee57.cdf")
    n57 = pf.Network(p57)
    assert n57.nB % 2 == 1
    o57 = oracle.Case(p57)
    P, Q, L = _scenarios(synth, n57, 6, 3600000)
    ref = _oracle_solve(o57, P, Q, L, lim=True)
    opts = pf.options(tolerancia=1e-8, no_max_iter=60, lim_inj_reat=1,
                      precision=pf.MIXED if precision == "mixed" else pf.FP64)
    b = pf.Batch([n57], opts)
    b.set_scenarios(np.zeros(6, np.int32), P.ravel(), Q.ravel(), L.ravel())
    ran = []
    for mode in (2, 1, 0):
        try:
            b.set_shape(0, mode, 256)
        except pf.PfError:
            continue
        ran.append(mode)
        _say("ieee57", precision, "mode", mode)
        b.solve()
        V, th, it, cv, mm = b.results()
        lq, _ = b.limq()
        V, th = V.reshape(6, -1), th.reshape(6, -1)
        for s, r in enumerate(ref):
            assert bool(cv[s]) == bool(r["converged"]), (mode, s)
            assert np.array_equal(lq[s], r["limQ"]), (mode, s)
            if r["converged"]:
                assert np.abs(V[s] - r["V"]).max() <= 1e-6, (mode, s)
                assert np.abs(th[s] - r["theta"]).max() <= 1e-6, (mode, s)
                if precision == "fp64":
                    assert it[s] == r["iterations"], (mode, s)
    assert 2 in ran
