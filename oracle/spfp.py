"""Whole-program single-precision NR: the reference's SPFP build restated
(test infrastructure, like the rest of oracle/).

The reference's SPFP build compiles every floating-point value as float
(DOUBLE_MODE false: CPU/opcoesDeCompilacao.h:7, CPU/global.h float_type;
CPU/buildParallelSPFP.sh:2, GPU/buildSPFP.sh:2) and asks PARDISO for a
single-precision factorisation (iparm[27] = 1, CPU/esparso/spSolver.cpp:133).
This restatement keeps every array in one dtype: the case data rounded to it
on reading (atof into a float_type field), Ybus in the MATPOWER form
(CPU/esparso/spCalcYbus.cpp:13-61), P and Q (CPU/esparso/potencia/
spCalcPQliq.cpp:3-45), the mismatch of the non-slack P and the PQ Q rows,
the Jacobian by the closed forms of CPU/esparso/spCalcJac.cpp:418-548, a
dense LU solve in the same dtype (LAPACK getrf/getrs), and the update
(CPU/limInjReat.cpp:209-232).  It covers the NR loop without reactive
limits (the accuracy of the arithmetic, not the switching rules, is what
the SPFP build is a reference point for).  Dense, so for cases up to a few
thousand buses.

With dtype float64 the same code is an FP64 NR; tests/test_oracle.py pins
it against the oracle there, so the float32 run differs from it only in
the dtype.
"""
from __future__ import annotations

import numpy as np


def ybus(case, dtype):
    c = np.complex64 if dtype == np.float32 else np.complex128
    n = case.nB
    Y = np.zeros((n, n), c)
    for l in range(case.nL):
        f, t = int(case.frm[l]), int(case.to[l])
        z = c(complex(dtype(case.r[l]), dtype(case.x[l])))
        y = c(1) / z
        tap = c(complex(dtype(case.tap_re[l]), dtype(case.tap_im[l])))
        b = c(complex(0, dtype(case.bch[l]) / dtype(2)))
        Y[f, f] += (y + b) / c(abs(tap) ** 2)
        Y[f, t] += -y / np.conj(tap)
        Y[t, f] += -y / tap
        Y[t, t] += y + b
    Y[np.arange(n), np.arange(n)] += (case.gsh.astype(dtype) + 1j * case.bsh.astype(dtype)).astype(c)
    return Y


def solve(case, dtype=np.float32, tol=1e-4, max_iter=100):
    """NR from the case's flat start; returns (V, theta, iterations, converged)
    as float64 arrays holding the dtype's values."""
    dtype = np.dtype(dtype).type
    Y = ybus(case, dtype)
    G, B = Y.real.astype(dtype), Y.imag.astype(dtype)
    n, sl = case.nB, case.slack
    V = case.V.astype(dtype).copy()
    th = case.theta.astype(dtype).copy()
    Pliq, Qliq = case.Pliq.astype(dtype), case.Qliq.astype(dtype)
    npv = [k for k in range(n) if k != sl]
    pq = sorted(int(k) for k in case.pq)
    it = 0
    for it in range(max_iter + 1):
        d = th[:, None] - th[None, :]
        cs, sn = np.cos(d).astype(dtype), np.sin(d).astype(dtype)
        VV = (V[:, None] * V[None, :]).astype(dtype)
        P = (VV * (G * cs + B * sn)).sum(1).astype(dtype)
        Q = (VV * (G * sn - B * cs)).sum(1).astype(dtype)
        g = np.concatenate([Pliq[npv] - P[npv], Qliq[pq] - Q[pq]]).astype(dtype)
        if np.abs(g).max() <= tol:
            return V.astype(np.float64), th.astype(np.float64), it, True
        if it == max_iter:
            break
        # H, N, M, L (CPU/esparso/spCalcJac.cpp:418-548), off-diagonal and diagonal
        H = (VV * (G * sn - B * cs)).astype(dtype)
        N = (V[:, None] * (G * cs + B * sn)).astype(dtype)
        M = (-VV * (G * cs + B * sn)).astype(dtype)
        L = (V[:, None] * (G * sn - B * cs)).astype(dtype)
        dg = np.arange(n)
        Gd, Bd = np.diag(G), np.diag(B)
        H[dg, dg] = -Q - V * V * Bd
        N[dg, dg] = (P + V * V * Gd) / V
        M[dg, dg] = P - V * V * Gd
        L[dg, dg] = (Q - V * V * Bd) / V
        J = np.block([[H[np.ix_(npv, npv)], N[np.ix_(npv, pq)]],
                      [M[np.ix_(pq, npv)], L[np.ix_(pq, pq)]]]).astype(dtype)
        dx = np.linalg.solve(J, g).astype(dtype)
        th[npv] = (th[npv] + dx[:len(npv)]).astype(dtype)
        V[pq] = (V[pq] + dx[len(npv):]).astype(dtype)
    return V.astype(np.float64), th.astype(np.float64), it, False
