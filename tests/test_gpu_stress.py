"""GPU parity near voltage collapse.  Loads and dispatch are scaled by a
factor lam toward the nose of the P-V curve, where the Jacobian is nearly
singular: the case that tests hardest a factorisation that never pivots
(2x2 block pivots in elimination order, the dense chain's Gauss-Jordan) and
the FP32 factor refined in FP64.  The oracle (CPU/newtonRaphson.cpp:173-314
restated) decides which factors converge; the GPU paths must converge on the
same ones, to the same answer (|dV| <= 1e-6 pu, |dtheta| <= 1e-6 rad, the same
final PV->PQ set, the same iteration count: in mixed precision too, the
refined step leaves the Newton path unchanged up to the nose), and fail on
the others (the batch kernel stops a blown-up scenario early, the oracle runs
it to the limit).
The factors were chosen on the oracle: the last ones below the nose in steps
of 0.01-0.05 (min |V| 0.55-0.67 pu there), and one past it.  At 15,104 buses
the dense chain is a 202 x 202 FP32 Gauss-Jordan without pivoting; there the
flat-start Newton path is erratic near the nose (without Q limits 1.28-1.34
diverge while 1.36-1.38 converge), and the test takes converging factors on
both sides of that band and a diverging one inside it."""
import os
import sys

import numpy as np
import pytest

from conftest import GOLDEN

pytestmark = pytest.mark.gpu

# network: {Q limits on: (converging factors, a failing factor), off: ...}
LOADS = {
    "c118": {True: ([1.5, 1.6, 1.7], 1.9), False: ([2.0, 2.1, 2.15, 2.18], 2.4)},
    "c1888": {True: ([1.2, 1.24, 1.28], 1.6), False: ([1.5, 1.6, 1.62, 1.64], 1.8)},
    "c15104": {True: ([1.1, 1.15], 1.19), False: ([1.2, 1.24, 1.38], 1.3)},
}
_REF = {}   # (network, limits): the oracle's solves, shared by the parametrisations


def _say(*a):
    print("[gpu-stress]", *a, flush=True)
    sys.stdout.flush()


def _path(cases, name):
    return os.path.join(GOLDEN, "c118.cdf") if name == "c118" else cases.path(name)


def _injections(net, lam):
    """Pliq, Qliq, Qload with every load and every generator's P scaled by lam."""
    return lam * (net.Pg - net.Pload), net.Qg - lam * net.Qload, lam * net.Qload


def _reference(oracle, p, net, name, lim):
    conv, fail = LOADS[name][lim]
    lams = conv + [fail]
    if (name, lim) not in _REF:
        oc = oracle.Case(p)
        _REF[name, lim] = [oc.solve(tol=1e-8, max_iter=60, lim=lim, reuse_symbolic=True, **dict(zip(
            ("Pliq", "Qliq", "Qload"), _injections(net, lam)))) for lam in lams]
    ref = _REF[name, lim]
    # the chosen factors are what they claim to be
    assert [r["converged"] for r in ref] == [True] * len(conv) + [False]
    return lams, ref


def _check(name, lam, r, conv, V, th, lq, it):
    if not r["converged"]:
        assert not conv, (name, lam)
        return
    assert conv, (name, lam, it)
    assert np.abs(V - r["V"]).max() <= 1e-6, (name, lam)
    assert np.abs(th - r["theta"]).max() <= 1e-6, (name, lam)
    assert np.array_equal(lq, r["limQ"]), (name, lam)
    assert it == r["iterations"], (name, lam, it, r["iterations"])


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
@pytest.mark.parametrize("lim", [True, False])
@pytest.mark.parametrize("name", ["c118", "c1888", "c15104"])
def test_batch_near_collapse(pf, oracle, cases, name, lim, precision):
    p = _path(cases, name)
    net = pf.Network(p)
    lams, ref = _reference(oracle, p, net, name, lim)
    inj = [_injections(net, lam) for lam in lams]
    opts = pf.options(tolerancia=1e-8, no_max_iter=60, lim_inj_reat=int(lim),
                      precision=pf.MIXED if precision == "mixed" else pf.FP64)
    _say("batch", name, "limits", lim, precision, "factors", lams)
    b = pf.Batch([net], opts)
    b.set_scenarios(np.zeros(len(lams), np.int32), *[np.concatenate([x[k] for x in inj]) for k in range(3)])
    b.solve()
    V, th, it, cv, mm = b.results()
    lq, _ = b.limq()
    _say("iterations", it.tolist(), "oracle", [r["iterations"] for r in ref])
    nb = net.nB
    for s, (lam, r) in enumerate(zip(lams, ref)):
        if cv[s]:
            assert mm[s] <= 1e-8
        _check(name, lam, r, bool(cv[s]), V[s * nb:(s + 1) * nb], th[s * nb:(s + 1) * nb], lq[s], int(it[s]))


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
@pytest.mark.parametrize("lim", [True, False])
@pytest.mark.parametrize("name", ["c118", "c1888", "c15104"])
def test_solver_near_collapse(pf, oracle, cases, name, lim, precision):
    p = _path(cases, name)
    net = pf.Network(p)
    lams, ref = _reference(oracle, p, net, name, lim)
    s = pf.Solver(net, pf.options(tolerancia=1e-8, no_max_iter=60, lim_inj_reat=int(lim),
                                  precision=pf.MIXED if precision == "mixed" else pf.FP64))
    for lam, r in zip(lams, ref):
        _say("solver", name, "limits", lim, precision, "factor", lam)
        s.set_injections(*_injections(net, lam))
        s.reset()
        g = s.solve()
        V, th, lq = s.state()
        _say("iterations", g["iterations"], "oracle", r["iterations"], "status", g["status"])
        _check(name, lam, r, g["converged"], V, th, lq, g["iterations"])
