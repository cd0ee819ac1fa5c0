"""The allocation-free batch entry point (pf_batch_solve_noalloc) against the
allocating one (pf_batch_set_scenarios + pf_batch_solve + pf_batch_get /
pf_batch_get_limq) on the same scenarios: every output equal bit for bit,
on two calls in a row through the same handle, the second with fewer
scenarios.  Inputs and outputs live in device memory the caller owns,
allocated here through the HIP runtime libpf itself links; the entry point
allocates nothing."""
import ctypes as C
import os

import numpy as np
import pytest

from conftest import GOLDEN

pytestmark = pytest.mark.gpu


def _scen(synth, net, n, seed0):
    is_pv = np.zeros(net.nB, bool)
    is_pv[net.pv] = True
    return synth.scenario_batch(net.Pg, net.Pload, net.Qg, net.Qload, is_pv, np.arange(seed0, seed0 + n))


def _reference(pf, nets, opts, net_of, P, Q, L):
    b = pf.Batch(nets, opts)
    b.set_scenarios(net_of, P, Q, L)
    b.solve()
    V, th, it, cv, mm = b.results()
    lq, fl = b.limq()
    return V, th, it, cv, mm, np.concatenate(lq), fl


class _Dev:
    """Caller-owned device buffers through the HIP runtime that libpf links."""

    def __init__(self):
        self.hip = C.CDLL("libamdhip64.so.7")
        self.hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.hip.hipFree.argtypes = [C.c_void_p]
        self.hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.ptrs = []

    def put(self, a):
        a = np.ascontiguousarray(a)
        p = C.c_void_p()
        assert self.hip.hipMalloc(C.byref(p), max(a.nbytes, 1)) == 0
        self.ptrs.append(p)
        assert self.hip.hipMemcpy(p, a.ctypes.data_as(C.c_void_p), a.nbytes, 1) == 0   # host to device
        return p

    def get(self, p, like):
        out = np.empty_like(like)
        assert self.hip.hipMemcpy(out.ctypes.data_as(C.c_void_p), p, out.nbytes, 2) == 0   # device to host
        return out

    def free(self):
        for p in self.ptrs:
            self.hip.hipFree(p)
        self.ptrs = []


def test_noalloc_matches_batch(pf, synth, cases):
    lib = pf.lib()
    P_ = C.c_void_p
    lib.pf_batch_solve_noalloc.argtypes = [P_, C.c_int] + [P_] * 14
    p118 = os.path.join(GOLDEN, "c118.cdf")
    n118, n944 = pf.Network(p118), pf.Network(cases.path("c944"))
    nets = [n118, n944]
    opts = pf.options(tolerancia=1e-8, no_max_iter=60, lim_inj_reat=1, precision=pf.MIXED)
    handle = pf.Batch(nets, opts)   # no set_scenarios: the no-alloc call needs only the topologies
    dev = _Dev()
    for call, (k118, k944, seed) in enumerate(((24, 4, 3500000), (9, 2, 3600000))):
        A = _scen(synth, n118, k118, seed)
        B = _scen(synth, n944, k944, seed + 1000)
        net_of = np.array([0] * k118 + [1] * k944, np.int32)
        net_of = net_of[np.random.default_rng(call).permutation(len(net_of))]
        ia = ib = 0
        P, Q, L = [], [], []
        for k in net_of:
            src, i = (A, ia) if k == 0 else (B, ib)
            P.append(src[0][i]); Q.append(src[1][i]); L.append(src[2][i])
            if k == 0:
                ia += 1
            else:
                ib += 1
        P, Q, L = (np.concatenate(x) for x in (P, Q, L))
        ref = _reference(pf, nets, opts, net_of, P, Q, L)
        n = len(net_of)
        tot = len(P)
        tot_pv = sum(nets[k].nPV for k in net_of)
        shapes = (np.full(tot, -1.0), np.full(tot, -1.0), np.full(n, -1, np.int32), np.full(n, -1, np.int32),
                  np.full(n, -1.0), np.full(tot_pv, 7, np.uint8), np.full(n, -1, np.int32))
        d_in = [dev.put(x) for x in (net_of, P, Q, L)]
        d_out = [dev.put(x) for x in shapes]
        print(f"[noalloc] call {call}: {k118} x 118-bus + {k944} x 944-bus", flush=True)
        st = lib.pf_batch_solve_noalloc(handle.h, n, *d_in, None, None, *d_out, None)
        assert st == 0, (st, pf.last_error())
        got = [dev.get(p, like) for p, like in zip(d_out, shapes)]
        dev.free()
        got[3] = got[3].astype(bool)
        got[5] = got[5].astype(bool)
        print(f"[noalloc] call {call}: converged {int(got[3].sum())}/{n}, iterations {got[2].tolist()}",
              flush=True)
        for name, a, b in zip(("V", "theta", "iterations", "converged", "max_mismatch", "limQ", "flips"),
                              got, ref):
            assert np.array_equal(a, np.asarray(b).astype(a.dtype)), name
        assert got[3].all()
