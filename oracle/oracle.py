"""ctypes binding of the oracle (liboracle.so).

TEST INFRASTRUCTURE ONLY: the CPU checker for the HIP product path and the
CPU baseline of bench.py.  The product never imports this module.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

_lib = None


class OrcOpts(C.Structure):
    _fields_ = [("tol", C.c_double), ("max_iter", C.c_int), ("lim_inj_reat", C.c_int),
                ("openmp", C.c_int), ("reuse_symbolic", C.c_int)]


class OrcResult(C.Structure):
    _fields_ = [("iterations", C.c_int), ("converged", C.c_int), ("singular", C.c_int),
                ("max_mismatch", C.c_double), ("n_switch_events", C.c_int)]


def build(force: bool = False) -> str:
    if force or not os.path.exists(LIB_PATH):
        subprocess.run(["make", "-s", "-C", HERE, "-B" if force else "all"], check=True)
    return LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        P = C.c_void_p
        dp = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
        ip = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
        L.orc_case_read.argtypes = [C.c_char_p, C.c_int, C.c_double, C.c_double, C.POINTER(P)]
        L.orc_case_free.argtypes = [P]
        L.orc_case_get.argtypes = [P, P, P, P, P, P, P, P]
        L.orc_ybus_nnz.argtypes = [P]
        L.orc_ybus.argtypes = [P, ip, ip, dp]
        L.orc_pq.argtypes = [C.c_int, ip, ip, dp, dp, dp, dp, dp, C.c_int]
        L.orc_mismatch.argtypes = [C.c_int, C.c_int, dp, dp, dp, dp, C.c_int, ip, dp]
        L.orc_jacobian.argtypes = [C.c_int, C.c_int, ip, ip, dp, dp, dp, dp, dp, C.c_int, ip,
                                   ip, ip, dp, C.c_int]
        L.orc_sparse_solve.argtypes = [C.c_int, ip, ip, dp, dp, dp]
        L.orc_min_degree.argtypes = [C.c_int, ip, ip, ip]
        L.orc_limits.argtypes = [P, C.c_int, P, P, P, P, P, P]
        L.orc_nr.argtypes = [P, C.POINTER(OrcOpts), P, P, P, P, P, P, C.POINTER(OrcResult)]
        L.orc_nr_batch.argtypes = [P, C.POINTER(OrcOpts), C.c_int, P, P, P, P, P, P, P, P, C.c_int, C.c_int]
        L.orc_flows.argtypes = [P, P, P, P, P, P, P]
        _lib = L
    return _lib


def _ptr(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


class Case:
    """A network read by the oracle's CDF/CDFX reader."""

    def __init__(self, path: str, wide: int = -1, v0: float = 1.0, theta0: float = 0.0):
        h = C.c_void_p()
        rc = lib().orc_case_read(path.encode(), wide, v0, theta0, C.byref(h))
        if rc != 0:
            raise ValueError(f"oracle reader failed on {path}: status {rc}")
        self.h = h
        ints = np.zeros(6, np.int32)
        lib().orc_case_get(h, _ptr(ints), None, None, None, None, None, None)
        self.nB, self.nL, self.nPV, self.nPQ, self.slack, self.n_duplicates = map(int, ints)
        bus = np.zeros((11, self.nB))
        pv = np.zeros(self.nPV, np.int32)
        pq = np.zeros(self.nPQ, np.int32)
        lim = np.zeros((3, self.nPV))
        ft = np.zeros((2, self.nL), np.int32)
        br = np.zeros((6, self.nL))
        lib().orc_case_get(h, None, _ptr(bus), _ptr(pv), _ptr(pq), _ptr(lim), _ptr(ft), _ptr(br))
        (self.V, self.theta, self.Pliq, self.Qliq, self.Pload, self.Qload, self.Pg, self.Qg,
         self.Vbase, self.gsh, self.bsh) = bus
        self.pv, self.pq = pv, pq
        self.Qmin, self.Qmax, self.Vset = lim
        self.frm, self.to = ft
        self.r, self.x, self.bch, self.tap_re, self.tap_im, self.phi = br

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.orc_case_free(self.h)
            self.h = None

    def ybus(self):
        nnz = lib().orc_ybus_nnz(self.h)
        rp = np.zeros(self.nB + 1, np.int32)
        ci = np.zeros(nnz, np.int32)
        y = np.zeros(2 * nnz)
        lib().orc_ybus(self.h, rp, ci, y)
        return rp, ci, y[0::2] + 1j * y[1::2]

    def solve(self, tol=1e-8, max_iter=100, lim=True, openmp=False, reuse_symbolic=False,
              Pliq=None, Qliq=None, Qload=None, V0=None, theta0=None):
        o = OrcOpts(tol, max_iter, int(lim), int(openmp), int(reuse_symbolic))
        r = OrcResult()
        V = np.array(self.V if V0 is None else V0, dtype=np.float64)
        th = np.array(self.theta if theta0 is None else theta0, dtype=np.float64)
        limQ = np.zeros(max(self.nPV, 1), np.uint8)
        args = [None if a is None else np.ascontiguousarray(a, np.float64) for a in (Pliq, Qliq, Qload)]
        lib().orc_nr(self.h, C.byref(o), *[_ptr(a) for a in args], _ptr(V), _ptr(th),
                     _ptr(limQ), C.byref(r))
        return dict(V=V, theta=th, limQ=limQ[:self.nPV].astype(bool), iterations=r.iterations,
                    converged=bool(r.converged), singular=bool(r.singular),
                    max_mismatch=r.max_mismatch, n_switch_events=r.n_switch_events)

    def limits(self, mode, Q, V, limQ, QliqLim, Qload=None, Qliq=None):
        lq = np.ascontiguousarray(limQ, np.uint8).copy()
        ql = np.ascontiguousarray(QliqLim, np.float64).copy()
        Qload = self.Qload if Qload is None else Qload
        Qliq = self.Qliq if Qliq is None else Qliq
        a = [np.ascontiguousarray(x, np.float64) for x in (Qload, Qliq, Q, V)]
        f = lib().orc_limits(self.h, mode, *[_ptr(x) for x in a], _ptr(lq), _ptr(ql))
        return lq.astype(bool), ql, f

    def flows(self, V, theta):
        """Branch flows as the reference reports them: (Pdp, Qdp, Ppd, Qpd) [nL]."""
        V, theta = (np.ascontiguousarray(a, np.float64) for a in (V, theta))
        out = [np.zeros(self.nL) for _ in range(4)]
        lib().orc_flows(self.h, _ptr(V), _ptr(theta), *[_ptr(a) for a in out])
        return tuple(out)

    def solve_batch(self, Pliq, Qliq, Qload, tol=1e-8, max_iter=100, lim=True,
                    reuse_symbolic=False, nthreads=1, inner=False, want_limq=False):
        """nthreads scenario-parallel threads, or (inner=True) scenarios one
        after another with nthreads OpenMP threads inside each solve."""
        nsc = Pliq.shape[0]
        o = OrcOpts(tol, max_iter, int(lim), 0, int(reuse_symbolic))
        V = np.zeros((nsc, self.nB))
        th = np.zeros((nsc, self.nB))
        it = np.zeros(nsc, np.int32)
        cv = np.zeros(nsc, np.int32)
        lq = np.zeros((nsc, max(self.nPV, 1)), np.uint8)
        P_, Q_, L_ = (np.ascontiguousarray(a, np.float64) for a in (Pliq, Qliq, Qload))
        lib().orc_nr_batch(self.h, C.byref(o), nsc, _ptr(P_), _ptr(Q_), _ptr(L_), _ptr(V),
                           _ptr(th), _ptr(it), _ptr(cv), _ptr(lq), nthreads, int(inner))
        if want_limq:
            return V, th, it, cv.astype(bool), lq[:, :self.nPV].astype(bool)
        return V, th, it, cv.astype(bool)


def pq(rowptr, colind, y, V, theta, openmp=False):
    n = len(rowptr) - 1
    yy = np.empty(2 * len(y))
    yy[0::2] = y.real
    yy[1::2] = y.imag
    P = np.zeros(n)
    Q = np.zeros(n)
    lib().orc_pq(n, np.ascontiguousarray(rowptr, np.int32), np.ascontiguousarray(colind, np.int32),
                 yy, np.ascontiguousarray(V, np.float64), np.ascontiguousarray(theta, np.float64),
                 P, Q, int(openmp))
    return P, Q


def mismatch(nB, slack, Pliq, QliqLim, P, Q, pqlim):
    pqlim = np.ascontiguousarray(pqlim, np.int32)
    g = np.zeros(nB - 1 + len(pqlim))
    lib().orc_mismatch(nB, slack, *(np.ascontiguousarray(a, np.float64) for a in (Pliq, QliqLim, P, Q)),
                       len(pqlim), pqlim, g)
    return g


def jacobian(nB, slack, rowptr, colind, y, V, theta, P, Q, pqlim):
    """Reduced Jacobian in the reference's CSR layout -> (rowptr, col, val)."""
    pqlim = np.ascontiguousarray(pqlim, np.int32)
    yy = np.empty(2 * len(y))
    yy[0::2] = y.real
    yy[1::2] = y.imag
    n = nB - 1 + len(pqlim)
    cap = 4 * len(y) + 16
    rp = np.zeros(n + 1, np.int32)
    cj = np.zeros(cap, np.int32)
    vj = np.zeros(cap)
    nnz = lib().orc_jacobian(nB, slack, np.ascontiguousarray(rowptr, np.int32),
                             np.ascontiguousarray(colind, np.int32), yy,
                             *(np.ascontiguousarray(a, np.float64) for a in (V, theta, P, Q)),
                             len(pqlim), pqlim, rp, cj, vj, cap)
    return rp, cj[:nnz].copy(), vj[:nnz].copy()


def sparse_solve(rowptr, colind, val, b):
    n = len(rowptr) - 1
    x = np.zeros(n)
    rc = lib().orc_sparse_solve(n, np.ascontiguousarray(rowptr, np.int32),
                                np.ascontiguousarray(colind, np.int32),
                                np.ascontiguousarray(val, np.float64),
                                np.ascontiguousarray(b, np.float64), x)
    if rc:
        raise np.linalg.LinAlgError("oracle LU: zero pivot")
    return x


def min_degree(rowptr, colind):
    n = len(rowptr) - 1
    p = np.zeros(n, np.int32)
    lib().orc_min_degree(n, np.ascontiguousarray(rowptr, np.int32),
                         np.ascontiguousarray(colind, np.int32), p)
    return p
