"""Pins the oracle (the CPU restatement of the reference's sparse NR path)
against what fixes it independently: the IEEE data files, numpy golden
vectors (dense Ybus, complex power, finite-difference Jacobian, a plain NR),
closed-form invariants and hand-worked known answers for the reactive-limit
rules (CPU/limInjReat.cpp:38-131)."""
import os

import numpy as np
import pytest

import sys

from conftest import GOLDEN, IEEE, ROOT, csr_to_dense, power, ybus_dense

sys.path.insert(0, os.path.join(ROOT, "oracle"))

TOL_NR = 1e-8


def golden(name):
    return np.load(os.path.join(GOLDEN, f"{name}_golden.npz"))


def case_path(name):
    return os.path.join(IEEE, name + ".cdf") if name.startswith("ieee") else os.path.join(GOLDEN, name + ".cdf")


# ---------------------------------------------------------------- reader

# counts computed from the CDF text (bus types, distinct branch pairs)
READER_COUNTS = {
    "ieee14": (14, 4, 9, 20, 0),
    "ieee30": (30, 5, 24, 41, 0),
    "ieee57": (57, 6, 50, 78, 2),
    "ieee118": (118, 53, 64, 179, 7),
}


@pytest.mark.parametrize("name", sorted(READER_COUNTS))
def test_reader_counts(oracle, name):
    c = oracle.Case(case_path(name))
    assert (c.nB, c.nPV, c.nPQ, c.nL, c.n_duplicates) == READER_COUNTS[name]
    assert c.nB == c.nPV + c.nPQ + 1


def test_reader_known_fields(oracle):
    # ieee14.cdf, bus 2: PV 1.045 pu, Qmax 50 MVAr, Qmin -40 MVAr, load 21.7 + j12.7,
    # generation 40.0 + j42.4; bus 1 slack at 1.060; baseMVA 100
    c = oracle.Case(case_path("ieee14"))
    assert c.slack == 0 and c.V[0] == 1.060
    i = list(c.pv).index(1)
    assert c.V[1] == 1.045 and c.Vset[i] == 1.045
    assert c.Qmax[i] == 0.5 and c.Qmin[i] == -0.4
    assert c.Pload[1] == pytest.approx(0.217, abs=1e-15)
    assert c.Pliq[1] == pytest.approx(0.400 - 0.217, abs=1e-15)
    assert c.Qliq[1] == pytest.approx(0.424 - 0.127, abs=1e-15)
    assert c.bsh[8] == 0.19                      # bus 9 shunt, taken as pu
    # flat start for PQ buses, theta_inicial for everything but the slack
    assert c.V[3] == 1.0 and np.all(c.theta[1:] == 0.0)


def test_reader_parallel_merge(oracle):
    # ieee118 declares 7 parallel pairs; the reader keeps the first, with
    # z = z1 z2 / (z1 + z2) and the reference's bsh = 1 / (b1 + b2)
    lines = open(case_path("ieee118"), newline="").read().split("\n")
    start = next(i for i, l in enumerate(lines) if l.startswith("BRANCH")) + 1
    rows = {}
    for l in lines[start:]:
        if "-999" in l:
            break
        rows.setdefault((int(l[0:4]), int(l[4:9])), []).append(l)
    dup = [(k, v) for k, v in rows.items() if len(v) == 2]
    assert len(dup) == 7
    c = oracle.Case(case_path("ieee118"))
    ids = {}
    for i in range(c.nB):
        ids.setdefault(i + 1, i)
    (a, b), (l1, l2) = dup[0]
    k = [j for j in range(c.nL) if c.frm[j] == ids[a] and c.to[j] == ids[b]][0]
    z1 = complex(float(l1[19:29]), float(l1[29:40]))
    z2 = complex(float(l2[19:29]), float(l2[29:40]))
    z = z1 * z2 / (z1 + z2)
    assert complex(c.r[k], c.x[k]) == pytest.approx(z, rel=1e-14)
    b1, b2 = float(l1[40:50]), float(l2[40:50])
    assert c.bch[k] == pytest.approx(1.0 / (b1 + b2), rel=1e-14)


def _write(tmp_path, name, text):
    p = os.path.join(str(tmp_path), name)
    with open(p, "w") as f:
        f.write(text)
    return p


def _mutate(path, fn, tmp_path, name):
    lines = open(path).read().split("\n")
    return _write(tmp_path, name, "\n".join(fn(lines)))


@pytest.mark.parametrize("name", ["ieee14", "ieee118"])
def test_reader_cdfx_recolumned(oracle, name):
    """The CDFX column map (CPU/leitura/cdf.cpp:329-493) pinned against the
    CDF one: tests/golden/make_cdfx.py re-columns the IEEE file (one more
    column for the bus number, two more for the branch's to-bus), and the
    oracle reads the same network from both layouts."""
    sys.path.insert(0, GOLDEN)
    import make_cdfx
    src = os.path.join(IEEE, name + ".cdf")
    with open(src, newline="") as f:
        text = f.read()
    with open(os.path.join(IEEE, name + ".cdfx"), newline="") as f:
        assert f.read() == make_cdfx.recolumn(text)
    a = oracle.Case(src)
    b = oracle.Case(os.path.join(IEEE, name + ".cdfx"))
    for attr in ("nB", "nL", "nPV", "nPQ", "slack", "n_duplicates"):
        assert getattr(a, attr) == getattr(b, attr), attr
    for attr in ("V", "theta", "Pliq", "Qliq", "Pload", "Qload", "Pg", "Qg", "Vbase", "gsh", "bsh", "pv", "pq",
                 "Qmin", "Qmax", "Vset", "frm", "to", "r", "x", "bch", "tap_re", "tap_im", "phi"):
        assert np.array_equal(getattr(a, attr), getattr(b, attr)), attr


def test_reader_errors(oracle, tmp_path):
    src = case_path("c14")
    two_slack = _mutate(src, lambda L: L[:3] + [L[3][:24] + " 3" + L[3][26:]] + L[4:], tmp_path, "a.cdf")
    bad_type = _mutate(src, lambda L: L[:3] + [L[3][:24] + " 7" + L[3][26:]] + L[4:], tmp_path, "b.cdf")
    i_br = None

    def bad_id(L):
        j = next(i for i, l in enumerate(L) if l.startswith("BRANCH")) + 1
        return L[:j] + ["9999" + L[j][4:]] + L[j + 1:]

    unknown = _mutate(src, bad_id, tmp_path, "c.cdf")
    for p in (two_slack, bad_type, unknown):
        with pytest.raises(ValueError):
            oracle.Case(p)
    with pytest.raises(ValueError):
        oracle.Case(os.path.join(str(tmp_path), "missing.cdf"))
    del i_br


# ----------------------------------------------------------------- Ybus

@pytest.mark.parametrize("name", ["ieee14", "c14", "c118"])
def test_ybus_golden(oracle, name):
    c = oracle.Case(case_path(name))
    rp, ci, y = c.ybus()
    g = golden(name)
    Yg = np.zeros((c.nB, c.nB), complex)
    Yg[g["Y_i"], g["Y_j"]] = g["Y_v"]
    Y = csr_to_dense(rp, ci, y, c.nB)
    assert np.abs(Y - Yg).max() <= 1e-12 * np.abs(Yg).max()
    # pattern: diagonal + both directions of every branch
    assert len(ci) == c.nB + 2 * len({(min(a, b), max(a, b)) for a, b in zip(c.frm, c.to)})


@pytest.mark.parametrize("name", ["ieee57", "ieee118", "c118"])
def test_ybus_invariants(oracle, name):
    c = oracle.Case(case_path(name))
    rp, ci, y = c.ybus()
    Y = csr_to_dense(rp, ci, y, c.nB)
    assert np.abs(Y - ybus_dense(c)).max() <= 1e-12 * np.abs(Y).max()
    # no phase shifters: Y is symmetric
    assert np.abs(Y - Y.T).max() <= 1e-12 * np.abs(Y).max()
    # with every tap = 1, row sums = shunt + j sum(b/2)
    c2 = oracle.Case(case_path(name))
    c2.tap_re[:] = 1.0
    c2.tap_im[:] = 0.0
    Yd = ybus_dense(c2)
    bsum = np.zeros(c.nB)
    np.add.at(bsum, c.frm, c.bch / 2)
    np.add.at(bsum, c.to, c.bch / 2)
    assert np.allclose(Yd.sum(1), c.gsh + 1j * (c.bsh + bsum), atol=1e-9)


# --------------------------------------------------------------- P, Q, J

@pytest.mark.parametrize("name", ["ieee14", "c14", "c118"])
def test_pq_golden(oracle, name):
    c = oracle.Case(case_path(name))
    rp, ci, y = c.ybus()
    g = golden(name)
    P, Q = oracle.pq(rp, ci, y, g["V"], g["th"])
    scale = np.abs(g["P"]).max() + np.abs(g["Q"]).max()
    assert np.abs(P - g["P"]).max() <= 1e-12 * scale
    assert np.abs(Q - g["Q"]).max() <= 1e-12 * scale


@pytest.mark.parametrize("name", ["ieee14", "c14", "c118"])
@pytest.mark.parametrize("switched", [False, True])
def test_jacobian_fd(oracle, name, switched):
    c = oracle.Case(case_path(name))
    rp, ci, y = c.ybus()
    g = golden(name)
    V, th = g["V"], g["th"]
    P, Q = oracle.pq(rp, ci, y, V, th)
    pqlim = g["pqlim1"] if switched else c.pq
    Jref = g["J1"] if switched else g["J0"]
    rpJ, cJ, vJ = oracle.jacobian(c.nB, c.slack, rp, ci, y, V, th, P, Q, pqlim)
    J = csr_to_dense(rpJ, cJ, vJ, len(rpJ) - 1)
    assert J.shape == Jref.shape
    assert np.abs(J - Jref).max() <= 1e-6 * np.abs(Jref).max()
    # every structural zero of J is a zero of the finite-difference Jacobian
    mask = J == 0
    assert np.abs(Jref[mask]).max() <= 1e-6 * np.abs(Jref).max()


@pytest.mark.parametrize("name", ["c14", "c118", "ieee118"])
def test_sparse_solve(oracle, name):
    c = oracle.Case(case_path(name))
    rp, ci, y = c.ybus()
    rng = np.random.default_rng(5)
    V = 1 + 0.02 * rng.standard_normal(c.nB)
    th = 0.1 * rng.standard_normal(c.nB)
    P, Q = oracle.pq(rp, ci, y, V, th)
    rpJ, cJ, vJ = oracle.jacobian(c.nB, c.slack, rp, ci, y, V, th, P, Q, c.pq)
    n = len(rpJ) - 1
    b = rng.standard_normal(n)
    x = oracle.sparse_solve(rpJ, cJ, vJ, b)
    J = csr_to_dense(rpJ, cJ, vJ, n)
    xr = np.linalg.solve(J, b)
    assert np.abs(x - xr).max() <= 1e-10 * np.abs(xr).max()
    perm = oracle.min_degree(rpJ, cJ)
    assert sorted(perm) == list(range(n))


def test_mismatch_layout(oracle):
    c = oracle.Case(case_path("c14"))
    rp, ci, y = c.ybus()
    g = golden("c14")
    P, Q = oracle.pq(rp, ci, y, g["V"], g["th"])
    ql = c.Qliq.copy()
    ql[c.pv[0]] = 0.3
    pqlim = np.sort(np.append(c.pq, c.pv[0]))
    mis = oracle.mismatch(c.nB, c.slack, c.Pliq, ql, P, Q, pqlim)
    nons = [k for k in range(c.nB) if k != c.slack]
    assert np.array_equal(mis[:c.nB - 1], c.Pliq[nons] - P[nons])
    assert np.array_equal(mis[c.nB - 1:], ql[pqlim] - Q[pqlim])


# ------------------------------------------------------------------ NR

@pytest.mark.parametrize("name", ["ieee14", "c14", "c118"])
def test_nr_nolim_golden(oracle, name):
    c = oracle.Case(case_path(name))
    r = c.solve(tol=1e-10, lim=False)
    g = golden(name)
    assert r["converged"]
    assert np.abs(r["V"] - g["V_sol"]).max() < 1e-8
    assert np.abs(r["theta"] - g["th_sol"]).max() < 1e-8


def test_ieee14_archive_solution(oracle):
    # the CDF bus records hold the archive's solved |V| (3 decimals) and angle
    # (2 decimals); no generator of ieee14 sits at a Q limit there
    c = oracle.Case(case_path("ieee14"))
    r = c.solve(tol=1e-8, lim=False)
    lines = open(case_path("ieee14"), newline="").read().split("\n")[2:16]
    Va = np.array([float(l[27:33]) for l in lines])
    Aa = np.array([float(l[33:40]) for l in lines])
    assert np.abs(r["V"] - Va).max() < 2e-3
    assert np.abs(np.degrees(r["theta"]) - Aa).max() < 0.03


def _check_solution(c, r, tol):
    """Invariants that pin a solution with reactive limits: the mismatch of the
    final PV/PQ set is within tol, every free PV bus is inside its limits,
    every limited bus sits at its limit.  (A bus switched back to PV keeps the
    |V| it reached while limited: the reference's back-switch,
    CPU/limInjReat.cpp:62-87, does not restore the setpoint.)"""
    Y = ybus_dense(c)
    P, Q = power(Y, r["V"], r["theta"])
    nons = [k for k in range(c.nB) if k != c.slack]
    assert np.abs(c.Pliq[nons] - P[nons]).max() <= 10 * tol
    for k in c.pq:
        assert abs(c.Qliq[k] - Q[k]) <= 10 * tol
    for i, b in enumerate(c.pv):
        qg = Q[b] + c.Qload[b]
        if r["limQ"][i]:
            assert min(abs(qg - c.Qmin[i]), abs(qg - c.Qmax[i])) <= 10 * tol
        else:
            assert c.Qmin[i] - 1e-4 - 10 * tol <= qg <= c.Qmax[i] + 1e-4 + 10 * tol


@pytest.mark.parametrize("name", ["c14", "c118"])
def test_nr_limits_invariants(oracle, name):
    c = oracle.Case(case_path(name))
    r = c.solve(tol=1e-8, lim=True)
    assert r["converged"] and r["limQ"].any() and r["n_switch_events"] > 0
    _check_solution(c, r, 1e-8)


def test_nr_reuse_symbolic_same_answer(oracle):
    c = oracle.Case(case_path("c118"))
    a = c.solve(tol=1e-8, lim=True)
    b = c.solve(tol=1e-8, lim=True, reuse_symbolic=True)
    assert a["iterations"] == b["iterations"]
    assert np.array_equal(a["limQ"], b["limQ"])
    assert np.abs(a["V"] - b["V"]).max() < 1e-10


def test_nr_not_converged(oracle):
    c = oracle.Case(case_path("c118"))
    r = c.solve(tol=1e-8, lim=True, max_iter=1)
    assert not r["converged"] and r["iterations"] == 1


def test_nr_batch_matches_single(oracle):
    c = oracle.Case(case_path("c14"))
    rng = np.random.default_rng(3)
    f = rng.uniform(0.8, 1.2, (4, c.nB))
    Pl, Ql, Qld = c.Pliq * f, c.Qliq * f, c.Qload * f
    V, th, it, cv = c.solve_batch(Pl, Ql, Qld, nthreads=2)
    for s in range(4):
        r = c.solve(Pliq=Pl[s], Qliq=Ql[s], Qload=Qld[s])
        assert cv[s] == r["converged"] and it[s] == r["iterations"]
        assert np.array_equal(V[s], r["V"])


# ------------------------------------- NR with limits: a traced known answer

def _nr_prescribed(Y, V, th, Psp, Qsp, is_q, schedule, max_iter=20, tol=1e-8):
    """Dense polar NR in numpy (finite-difference Jacobian) whose PV/PQ changes
    are PRESCRIBED: schedule[pass] = [(bus, has_q_equation, Q spec)], applied
    before that pass's mismatch.  No limit rule is evaluated here; a bus that
    loses its Q equation keeps the |V| it has.  Returns the final state and
    per-pass records (max |g| at the pass, max |dx| of its step, Q and V)."""
    V, th, is_q, Qsp = V.copy(), th.copy(), is_q.copy(), Qsp.copy()
    n, rec = len(V), []
    for it in range(max_iter):
        for b, q, qs in schedule.get(it, []):
            is_q[b], Qsp[b] = q, qs
        rp = list(range(1, n))
        rq = [k for k in rp if is_q[k]]

        def calc(t_, v_):
            E = v_ * np.exp(1j * t_)
            S = E * np.conj(Y @ E)
            return np.concatenate([S.real[rp], S.imag[rq]]), S

        F, S = calc(th, V)
        g = np.concatenate([Psp[rp], Qsp[rq]]) - F
        rec.append({"maxg": float(np.abs(g).max()), "Q": S.imag.copy(), "V": V.copy()})
        if np.abs(g).max() <= tol:
            return V, th, it, rec
        x0 = np.concatenate([th[rp], V[rq]])
        J = np.zeros((len(x0), len(x0)))
        for j in range(len(x0)):
            e = np.zeros(len(x0))
            e[j] = 1e-7

            def at(x):
                t_, v_ = th.copy(), V.copy()
                t_[rp], v_[rq] = x[:len(rp)], x[len(rp):]
                return calc(t_, v_)[0]
            J[:, j] = (at(x0 + e) - at(x0 - e)) / 2e-7
        dx = np.linalg.solve(J, g)
        th[rp] += dx[:len(rp)]
        V[rq] += dx[len(rp):]
        rec[-1]["dx"] = float(np.abs(dx).max())
    return V, th, max_iter, rec


def test_nr_limits_traced_switch_and_back_switch(oracle):
    """tests/golden/c3_limits.cdf: slack bus 1, PV bus 2 (|V| set 0.97,
    64 MW, Q limits [-1, 100] MVAr), PQ bus 3 (160 MW + 60 MVAr load).
    Traced by hand from the plain NR passes (numpy, below):
      pass 0: no check (the reference checks from iteration 1 on);
      pass 1: the previous step is max|dx| = 0.086 < 0.1, so chkLimQ runs
              (CPU/newtonRaphson.cpp:202-212) although max|g| = 0.125 >= 0.1
              at this pass; Q_2 = -0.0181 < Qmin - folga = -0.0101 -> bus 2
              becomes PQ at Qmin (CPU/limInjReat.cpp:44-51);
      pass 2: bus 2 now absorbs 0.01 pu only, so its |V| fell to 0.96447
              < Vset - folga; it is limited at Qmin, so it goes back to PV
              (CPU/limInjReat.cpp:73-82) KEEPING |V| = 0.96447;
      then:   Q_2 stays within its limits; converged after 4 updates.
    A trigger on the current mismatch instead of the previous step would
    skip pass 1 and never switch; an undo/back-switch with the wrong
    voltage test would leave bus 2 limited.  Both fail here."""
    c = oracle.Case(os.path.join(GOLDEN, "c3_limits.cdf"))
    assert (c.nB, c.nPV, c.nPQ) == (3, 1, 1)
    Y = ybus_dense(c)
    is_q = np.zeros(3, bool)
    is_q[c.pq] = True
    folga = 1e-4
    # the plain NR passes fix the trace (no limit logic involved)
    _, _, _, plain = _nr_prescribed(Y, c.V, c.theta, c.Pliq, c.Qliq, is_q, {})
    assert plain[0]["dx"] < 0.1 <= plain[1]["maxg"]
    assert plain[1]["Q"][1] + c.Qload[1] < c.Qmin[0] - folga
    Vp, thp, itp, rec = _nr_prescribed(Y, c.V, c.theta, c.Pliq, c.Qliq, is_q,
                                       {1: [(1, True, c.Qmin[0] - c.Qload[1])], 2: [(1, False, 0.0)]})
    assert rec[2]["V"][1] < c.Vset[0] - folga
    Vkeep = rec[2]["V"][1]
    # the oracle takes the switches at exactly those passes
    events = []
    for k in range(1, 6):
        r = c.solve(tol=1e-8, lim=True, max_iter=k)
        events.append((r["n_switch_events"], bool(r["limQ"][0])))
    assert events == [(0, False), (1, True), (2, False), (2, False), (2, False)]
    r = c.solve(tol=1e-8, lim=True)
    assert r["converged"] and r["iterations"] == itp == 4
    assert not r["limQ"][0] and r["n_switch_events"] == 2
    assert abs(r["V"][1] - Vkeep) <= 1e-9 and abs(r["V"][1] - c.Vset[0]) > 5e-3
    assert np.abs(r["V"] - Vp).max() <= 1e-9 and np.abs(r["theta"] - thp).max() <= 1e-9


# ------------------------------------------- reactive-limit rules by hand

def test_limit_rules_known_answers(oracle):
    """CPU/limInjReat.cpp:38-131 on hand-set states of the 14-bus case."""
    c = oracle.Case(case_path("c14"))
    pv = c.pv
    n = c.nPV
    V = c.V.copy()
    Q = np.zeros(c.nB)
    limQ = np.zeros(n, bool)
    QL = c.Qliq.copy()
    # bus 0: Qg below Qmin - folga -> PQ at Qmin; bus 1: above Qmax + folga -> PQ at Qmax
    # bus 2: inside (just within folga of Qmax) -> stays PV; bus 3: inside -> PV
    Q[pv[0]] = c.Qmin[0] - 2e-4 - c.Qload[pv[0]]
    Q[pv[1]] = c.Qmax[1] + 2e-4 - c.Qload[pv[1]]
    Q[pv[2]] = c.Qmax[2] + 0.5e-4 - c.Qload[pv[2]]
    Q[pv[3]] = 0.5 * (c.Qmin[3] + c.Qmax[3]) - c.Qload[pv[3]]
    lq, ql, f = oracle_limits = c.limits(0, Q, V, limQ, QL)
    assert f == 2 and lq.tolist() == [True, True, False, False]
    assert ql[pv[0]] == c.Qmin[0] - c.Qload[pv[0]]
    assert ql[pv[1]] == c.Qmax[1] - c.Qload[pv[1]]
    assert ql[pv[2]] == c.Qliq[pv[2]]
    del oracle_limits
    # back-switch: limited at Qmax with V > Vset + folga -> PV; at Qmin with
    # V < Vset - folga -> PV; the opposite sides stay limited
    V2 = V.copy()
    V2[pv[0]] = c.Vset[0] - 2e-4      # limited at Qmin, V below -> back to PV
    V2[pv[1]] = c.Vset[1] - 2e-4      # limited at Qmax, V below -> stays
    lq2, ql2, f2 = c.limits(0, Q, V2, lq, ql)
    assert f2 == 1 and lq2.tolist() == [False, True, False, False]
    assert ql2[pv[0]] == c.Qliq[pv[0]]
    V3 = V.copy()
    V3[pv[1]] = c.Vset[1] + 2e-4      # limited at Qmax, V above -> back to PV
    lq3, ql3, f3 = c.limits(1, Q, V3, lq, ql)   # undo_chkLimQ: back-switch rules only
    assert f3 == 1 and lq3.tolist() == [True, False, False, False]
    # undo never switches a free bus, whatever its Q
    lq4, _, f4 = c.limits(1, Q, V, np.zeros(n, bool), QL)
    assert f4 == 0 and not lq4.any()


# ------------------------------------------------------ warm-start MLP oracle

def _flows_closed_form(c, V, th, phase=None):
    """Flows of every branch from its own parameters (numpy complex), in the
    reference's convention: with Y_km the branch's Ybus term, psi_km the phase
    added by phif (+phi from the from side, -phi from the to side), and the
    full line charging b:  S = conj(Y_km) (V_k^2 - V_k V_m e^{j(th_k - th_m + psi_km)}),
    P = Re S, Q = Im S - V_k^2 b."""
    out = [np.zeros(c.nL) for _ in range(4)]
    for l in range(c.nL):
        f, t = c.frm[l], c.to[l]
        y = 1.0 / complex(c.r[l], c.x[l])
        tap = complex(c.tap_re[l], c.tap_im[l])
        for (k, m, Ykm, psi, P, Q) in ((f, t, -y / np.conj(tap), c.phi[l], out[0], out[1]),
                                      (t, f, -y / tap, -c.phi[l], out[2], out[3])):
            S = np.conj(Ykm) * (V[k] ** 2 - V[k] * V[m] * np.exp(1j * (th[k] - th[m] + psi)))
            P[l] = S.real
            Q[l] = S.imag - V[k] ** 2 * c.bch[l]
    return out


@pytest.mark.parametrize("name", ["ieee14", "ieee30", "c118"])
def test_flows_closed_form(oracle, name):
    """Reported branch flows against the closed form from the branch data (the
    reference's sign and its full line charging in Q), at the NR solution;
    branches sharing a bus pair are skipped (their Ybus term is the sum)."""
    path = os.path.join(IEEE, name + ".cdf") if name.startswith("ieee") else os.path.join(GOLDEN, name + ".cdf")
    c = oracle.Case(path)
    r = c.solve(tol=1e-10, lim=False)
    assert r["converged"]
    got = c.flows(r["V"], r["theta"])
    ref = _flows_closed_form(c, r["V"], r["theta"])
    pairs = [tuple(sorted(p)) for p in zip(c.frm, c.to)]
    alone = np.array([pairs.count(p) == 1 for p in pairs])
    assert alone.sum() >= c.nL - 8
    for g, e in zip(got, ref):
        assert np.abs(g[alone] - e[alone]).max() <= 1e-12
    # the reported flow is minus the usual k -> m flow: Pdp + Ppd = -(series losses)
    E = r["V"] * np.exp(1j * r["theta"])
    for l in np.where(alone & (c.tap_re == 1) & (c.tap_im == 0))[0]:
        f, t = c.frm[l], c.to[l]
        loss = (abs(E[f] - E[t]) ** 2 / complex(c.r[l], c.x[l]).conjugate()).real
        assert abs(got[0][l] + got[2][l] + loss) <= 1e-12


def test_flows_phase_shifter(oracle, tmp_path):
    """A 5-degree phase shift on branch 4-7 of ieee14: phif adds +phi from the
    from side and -phi from the to side, on top of the phase already in Ybus."""
    lines = open(os.path.join(IEEE, "ieee14.cdf"), "rb").read().decode().split("\n")
    k = next(i for i, l in enumerate(lines) if l.startswith("   4    7"))
    lines[k] = lines[k][:83] + f"{5.0:<7.1f}" + lines[k][90:]
    p = tmp_path / "ps14.cdf"
    p.write_bytes("\n".join(lines).encode())
    c = oracle.Case(str(p))
    l = int(np.where((c.frm == 3) & (c.to == 6))[0][0])
    assert abs(c.phi[l] - np.deg2rad(5.0)) < 1e-15
    rng = np.random.default_rng(5)
    V = 1 + 0.05 * rng.standard_normal(c.nB)
    th = 0.1 * rng.standard_normal(c.nB)
    got = c.flows(V, th)
    ref = _flows_closed_form(c, V, th)
    for g, e in zip(got, ref):
        assert np.abs(g - e).max() <= 1e-12


def test_mlp_oracle_gradient_fd():
    """The MLP oracle's backward pass against central differences (FP64)."""
    import mlp_oracle as M
    rng = np.random.default_rng(0)
    shapes = [(16, 24), (24, 24), (24, 8)]
    n = sum(k * m + m for k, m in shapes)
    p = rng.standard_normal(n) * 0.3
    X = rng.standard_normal((8, 16))
    T = rng.standard_normal((8, 8))
    xm, xs = X.mean(0), X.std(0) + 0.5
    ym, ys = T.mean(0), T.std(0) + 0.5
    loss, g = M.loss_and_grad(p, shapes, X, T, xm, xs, ym, ys, quantize=False)
    h = 1e-6
    for i in rng.choice(n, 40, replace=False):
        e = np.zeros(n)
        e[i] = h
        lp, _ = M.loss_and_grad(p + e, shapes, X, T, xm, xs, ym, ys, quantize=False)
        lm, _ = M.loss_and_grad(p - e, shapes, X, T, xm, xs, ym, ys, quantize=False)
        assert abs((lp - lm) / (2 * h) - g[i]) <= 1e-6 * (1 + abs(g[i])), i


def test_bf16_rounding_known_answers():
    import mlp_oracle as M
    # 1 + 2^-8 is exactly halfway between bf16 neighbours 1 and 1 + 2^-7: ties to even -> 1
    assert M.bf16(np.float32(1 + 2 ** -8)) == 1.0
    assert M.bf16(np.float32(1 + 3 * 2 ** -8)) == 1 + 2 ** -6
    assert M.bf16(np.float32(1 + 2 ** -7)) == 1 + 2 ** -7
    assert M.bf16(np.float32(-3.0)) == -3.0


# ------------------------------------------------ whole-program FP32 (SPFP)

@pytest.mark.parametrize("name", ["ieee14", "c118"])
def test_spfp_restatement(oracle, name):
    """oracle/spfp.py (the reference's single-precision build restated):
    with dtype float64 it is an FP64 NR and reproduces the oracle without
    limits (1e-9); with float32 it converges at the reference's default
    tolerance 1e-4 (CPU/global.cpp:10) to within FP32 accuracy of it."""
    import spfp
    c = oracle.Case(case_path(name))
    o = c.solve(tol=1e-8, lim=False)
    V64, th64, it64, ok64 = spfp.solve(c, np.float64, tol=1e-8)
    assert ok64 and it64 == o["iterations"]
    assert np.abs(V64 - o["V"]).max() <= 1e-9 and np.abs(th64 - o["theta"]).max() <= 1e-9
    V32, th32, it32, ok32 = spfp.solve(c, np.float32, tol=1e-4)
    assert ok32
    err = max(np.abs(V32 - o["V"]).max(), np.abs(th32 - o["theta"]).max())
    assert 1e-9 < err < 1e-3
