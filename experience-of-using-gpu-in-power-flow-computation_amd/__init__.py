"""Python binding of libpf (MI355X Newton-Raphson power flow) over its C ABI.

The product is the C ABI in include/pf.h (libpf.so, HIP kernels for gfx950)
and the ``flumen`` CLI that reads data.txt like the reference's executables.
This module only marshals arrays for tests and benchmarks; every number it
returns was computed by libpf.  There is no CPU fallback: if libpf.so is
missing, importing the binding raises.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libpf.so")

PF_OK, PF_NOT_CONVERGED, PF_SINGULAR, PF_PARSE_ERROR, PF_DEVICE_ERROR, PF_INVALID_ARG, PF_IO_ERROR = range(7)
ESPARSO, DENSO, HIBRIDO_C, HIBRIDO_A, HIBRIDO_B = 0, 1, 2, 3, 4
FP64, MIXED = 0, 1
BUS_FIELDS = ["V", "theta", "Pliq", "Qliq", "Pload", "Qload", "Pg", "Qg", "Vbase", "gsh", "bsh"]


class PfError(RuntimeError):
    def __init__(self, status: int, what: str):
        self.status = status
        super().__init__(f"{what}: {status_str(status)} ({status})")


class Options(C.Structure):
    _fields_ = [("metodo", C.c_int), ("lim_inj_reat", C.c_int), ("openmp", C.c_int),
                ("v_inicial", C.c_double), ("theta_inicial", C.c_double),
                ("tolerancia", C.c_double), ("no_max_iter", C.c_int),
                ("arq_entrada", C.c_char * 1024), ("verbose_mode", C.c_int),
                ("laconic_mode", C.c_int), ("output_bench", C.c_int), ("output_ans", C.c_int),
                ("output_iter", C.c_int), ("streams", C.c_int), ("precision", C.c_int),
                ("refine_steps", C.c_int), ("use_graph", C.c_int), ("lu_blocked", C.c_int)]


class NetworkInfo(C.Structure):
    _fields_ = [("nB", C.c_int), ("nL", C.c_int), ("nPV", C.c_int), ("nPQ", C.c_int),
                ("slack", C.c_int), ("n_duplicates", C.c_int), ("baseMVA", C.c_double)]


class Result(C.Structure):
    _fields_ = [("iterations", C.c_int), ("converged", C.c_int), ("singular", C.c_int),
                ("max_mismatch", C.c_double), ("n_switch_events", C.c_int),
                ("solve_ms", C.c_double)]


def build(force: bool = False) -> str:
    """Compile libpf.so in-tree (hipcc, gfx950)."""
    args = ["make", "-s", "-C", HERE, "-j8"]
    if force:
        args.append("-B")
    subprocess.run(args, check=True)
    return LIB_PATH


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is not built (run make in {HERE})")
        L = C.CDLL(LIB_PATH)
        P = C.c_void_p
        L.pf_status_str.restype = C.c_char_p
        L.pf_status_str.argtypes = [C.c_int]
        L.pf_last_error.restype = C.c_char_p
        L.pf_options_default.argtypes = [C.POINTER(Options)]
        L.pf_options_parse.argtypes = [C.c_char_p, C.c_int, C.POINTER(Options)]
        L.pf_options_load.argtypes = [C.c_char_p, C.c_int, C.POINTER(Options)]
        L.pf_read_case.argtypes = [C.c_char_p, C.c_double, C.c_double, C.POINTER(P)]
        L.pf_network_free.argtypes = [P]
        L.pf_network_get_info.argtypes = [P, C.POINTER(NetworkInfo)]
        L.pf_network_get_bus.argtypes = [P, C.c_int, P]
        L.pf_network_get_lists.argtypes = [P, P, P, P, P, P]
        L.pf_network_get_branches.argtypes = [P, P, P, P, P, P, P, P]
        L.pf_solver_create.argtypes = [P, C.POINTER(Options), C.c_int, C.POINTER(P)]
        L.pf_solver_destroy.argtypes = [P]
        L.pf_solve.argtypes = [P, C.POINTER(Result)]
        L.pf_solver_reset.argtypes = [P]
        L.pf_solver_set_state.argtypes = [P, P, P]
        L.pf_solver_set_injections.argtypes = [P, P, P, P]
        L.pf_solver_get_state.argtypes = [P, P, P, P]
        L.pf_solver_lu_info.argtypes = [P, P, P, P, P]
        L.pf_stage_ybus.argtypes = [P, P, P, P, P]
        L.pf_stage_pq.argtypes = [P, P, P, P, P]
        L.pf_stage_flows.argtypes = [P, P, P, P, P, P, P]
        L.pf_solver_flows.argtypes = [P, P, P, P, P]
        L.pf_stage_mismatch.argtypes = [P, P, P, P, P, P, P]
        L.pf_stage_jacobian.argtypes = [P, P, P, P, P, P, P, P, C.c_int, P]
        L.pf_stage_lu_solve.argtypes = [P, P, P, P, P, P]
        L.pf_stage_limits.argtypes = [P, C.c_int, P, P, P, P, P]
        L.pf_probe_create.argtypes = [P, C.POINTER(P), P, P, P, P]
        L.pf_probe_free.argtypes = [P]
        L.pf_probe_get.argtypes = [P, P, P, P, P, P]
        L.pf_probe_factor_solve.argtypes = [P, P, P]
        L.pf_probe_flat_factor_solve.argtypes = [P, P, P]
        L.pf_probe_levels.argtypes = [P, P, P]
        L.pf_probe_chain.argtypes = [P, P, P]
        L.pf_probe_parts.argtypes = [P, P, P, P, P, P]
        if hasattr(L, "pf_warmstart_create"):
            L.pf_warmstart_create.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, C.c_ulonglong, C.c_int,
                                              C.POINTER(P)]
            L.pf_warmstart_destroy.argtypes = [P]
            L.pf_warmstart_info.argtypes = [P, P, P]
            L.pf_warmstart_set_norm.argtypes = [P, P, P, P, P]
            L.pf_warmstart_get_params.argtypes = [P, P]
            L.pf_warmstart_set_params.argtypes = [P, P]
            L.pf_warmstart_grad.argtypes = [P, P, P, C.c_int, P, P]
            L.pf_warmstart_fit.argtypes = [P, P, P, C.c_int, C.c_int, C.c_int, C.c_float, C.c_float,
                                           C.c_ulonglong, P]
            L.pf_warmstart_predict.argtypes = [P, P, C.c_int, P]
            L.pf_warmstart_stage.argtypes = [P, P, C.c_int]
            L.pf_warmstart_predict_into.argtypes = [P, P, C.c_int, C.c_int]
            L.pf_warmstart_gemm_probe.argtypes = [P, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, P, P, P]
        if hasattr(L, "pf_batch_create"):
            L.pf_batch_create.argtypes = [P, C.c_int, C.POINTER(Options), C.c_int, C.POINTER(P)]
            L.pf_batch_destroy.argtypes = [P]
            L.pf_batch_set_scenarios.argtypes = [P, C.c_int, P, P, P, P, P, P]
            L.pf_batch_solve.argtypes = [P, P]
            L.pf_batch_get.argtypes = [P, P, P, P, P, P]
            L.pf_batch_get_limq.argtypes = [P, P, P]
            L.pf_batch_autotune.argtypes = [P, C.c_char_p, C.c_int, P]
            L.pf_batch_get_shape.argtypes = [P, C.c_int, P, P, P, P]
            L.pf_batch_set_shape.argtypes = [P, C.c_int, C.c_int, C.c_int, C.c_int]
        _lib = L
    return _lib


def status_str(s: int) -> str:
    return lib().pf_status_str(s).decode()


def last_error() -> str:
    return lib().pf_last_error().decode()


def _p(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def _chk(s: int, what: str):
    if s != PF_OK:
        raise PfError(s, what + (f" [{last_error()}]" if s in (PF_DEVICE_ERROR, PF_INVALID_ARG) else ""))


def options(text: str | None = None, cpu_layout: int = 0, **kw) -> Options:
    """data.txt options (defaults of CPU/global.cpp when text is None)."""
    o = Options()
    if text is None:
        lib().pf_options_default(C.byref(o))
    else:
        _chk(lib().pf_options_parse(text.encode(), cpu_layout, C.byref(o)), "pf_options_parse")
    for k, v in kw.items():
        if k == "arq_entrada":
            v = v.encode()
        setattr(o, k, v)
    return o


class Network:
    """A CDF/CDFX case read by libpf's reader."""

    def __init__(self, path: str, v0: float = 1.0, theta0: float = 0.0):
        h = C.c_void_p()
        _chk(lib().pf_read_case(str(path).encode(), v0, theta0, C.byref(h)), f"pf_read_case({path})")
        self.h = h
        info = NetworkInfo()
        _chk(lib().pf_network_get_info(h, C.byref(info)), "pf_network_get_info")
        self.nB, self.nL, self.nPV, self.nPQ = info.nB, info.nL, info.nPV, info.nPQ
        self.slack, self.n_duplicates, self.baseMVA = info.slack, info.n_duplicates, info.baseMVA
        for i, name in enumerate(BUS_FIELDS):
            a = np.zeros(self.nB)
            _chk(lib().pf_network_get_bus(h, i, _p(a)), "pf_network_get_bus")
            setattr(self, name, a)
        self.pv = np.zeros(self.nPV, np.int32)
        self.pq = np.zeros(self.nPQ, np.int32)
        self.Qmin, self.Qmax, self.Vset = (np.zeros(self.nPV) for _ in range(3))
        lib().pf_network_get_lists(h, _p(self.pv), _p(self.pq), _p(self.Qmin), _p(self.Qmax), _p(self.Vset))
        self.frm = np.zeros(self.nL, np.int32)
        self.to = np.zeros(self.nL, np.int32)
        self.r, self.x, self.bch, self.tap_re, self.tap_im = (np.zeros(self.nL) for _ in range(5))
        lib().pf_network_get_branches(h, _p(self.frm), _p(self.to), _p(self.r), _p(self.x),
                                      _p(self.bch), _p(self.tap_re), _p(self.tap_im))

    def __del__(self):
        if getattr(self, "h", None) is not None and _lib is not None:
            _lib.pf_network_free(self.h)
            self.h = None


class SymbolicProbe:
    """Host-only view of the ordering / symbolic factor program (no GPU)."""

    def __init__(self, net: Network):
        h = C.c_void_p()
        n, nnzL, nslot, lev = (C.c_int() for _ in range(4))
        _chk(lib().pf_probe_create(net.h, C.byref(h), C.byref(n), C.byref(nnzL), C.byref(nslot),
                                   C.byref(lev)), "pf_probe_create")
        self.h = h
        self.n, self.nnzL, self.nslot, self.levels = n.value, nnzL.value, nslot.value, lev.value
        ynnz = C.c_int()
        lib().pf_probe_get(h, None, None, None, None, C.byref(ynnz))
        self.perm = np.zeros(self.n, np.int32)
        self.slot_e = np.zeros(self.nslot, np.int32)
        self.rowptr = np.zeros(net.nB + 1, np.int32)
        self.colind = np.zeros(ynnz.value, np.int32)
        lib().pf_probe_get(h, _p(self.perm), _p(self.slot_e), _p(self.rowptr), _p(self.colind), None)
        self.lvl_ptr = np.zeros(self.levels + 1, np.int32)
        self.colptr = np.zeros(self.n + 1, np.int32)
        lib().pf_probe_levels(h, _p(self.lvl_ptr), _p(self.colptr))

    def parts(self):
        """(part_ptr reshaped per group: [n_parts, part_levels + 1] boundaries,
        part_nodes, parent) of the subtree groups below the narrow top."""
        npart, nl = C.c_int(), C.c_int()
        _chk(lib().pf_probe_parts(self.h, C.byref(npart), C.byref(nl), None, None, None), "pf_probe_parts")
        ptr = np.zeros(npart.value * nl.value + 1, np.int32)
        nodes = np.zeros(max(int(self.lvl_ptr[nl.value]), 1), np.int32)
        parent = np.zeros(self.n, np.int32)
        _chk(lib().pf_probe_parts(self.h, None, None, _p(ptr), _p(nodes), _p(parent)), "pf_probe_parts")
        return npart.value, nl.value, ptr, nodes[:int(self.lvl_ptr[nl.value])], parent

    @property
    def chain(self):
        """(length, first node) of the trailing dense chain factored as one block."""
        m, c0 = C.c_int(), C.c_int()
        _chk(lib().pf_probe_chain(self.h, C.byref(m), C.byref(c0)), "pf_probe_chain")
        return m.value, c0.value

    def factor_solve(self, blocks: np.ndarray, rhs: np.ndarray) -> np.ndarray:
        b = np.ascontiguousarray(blocks, np.float64).copy()
        x = np.ascontiguousarray(rhs, np.float64).copy()
        _chk(lib().pf_probe_factor_solve(self.h, _p(b), _p(x)), "pf_probe_factor_solve")
        return x

    def flat_factor_solve(self, blocks: np.ndarray, rhs: np.ndarray) -> np.ndarray:
        """The same through the order of the GPU's blocked program (flat pair
        programs, the narrow top's split sweeps), restated on the host."""
        b = np.ascontiguousarray(blocks, np.float64).copy()
        x = np.ascontiguousarray(rhs, np.float64).copy()
        _chk(lib().pf_probe_flat_factor_solve(self.h, _p(b), _p(x)), "pf_probe_flat_factor_solve")
        return x

    def __del__(self):
        if getattr(self, "h", None) is not None and _lib is not None:
            _lib.pf_probe_free(self.h)
            self.h = None


class Solver:
    """One network on one GPU (pf_solver)."""

    def __init__(self, net: Network, opts: Options, device: int = 0):
        h = C.c_void_p()
        _chk(lib().pf_solver_create(net.h, C.byref(opts), device, C.byref(h)), "pf_solver_create")
        self.h, self.net, self.opts = h, net, opts

    def __del__(self):
        if getattr(self, "h", None) is not None and _lib is not None:
            _lib.pf_solver_destroy(self.h)
            self.h = None

    def solve(self, check: bool = False) -> dict:
        r = Result()
        s = lib().pf_solve(self.h, C.byref(r))
        if s == PF_DEVICE_ERROR or (check and s != PF_OK):
            _chk(s, "pf_solve")
        return dict(status=s, iterations=r.iterations, converged=bool(r.converged),
                    singular=bool(r.singular), max_mismatch=r.max_mismatch,
                    n_switch_events=r.n_switch_events, solve_ms=r.solve_ms)

    def reset(self):
        _chk(lib().pf_solver_reset(self.h), "pf_solver_reset")

    def set_state(self, V, theta):
        V = np.ascontiguousarray(V, np.float64)
        th = np.ascontiguousarray(theta, np.float64)
        _chk(lib().pf_solver_set_state(self.h, _p(V), _p(th)), "pf_solver_set_state")

    def set_injections(self, Pliq, Qliq, Qload):
        a = [np.ascontiguousarray(x, np.float64) for x in (Pliq, Qliq, Qload)]
        _chk(lib().pf_solver_set_injections(self.h, *[_p(x) for x in a]), "pf_solver_set_injections")

    def state(self):
        V = np.zeros(self.net.nB)
        th = np.zeros(self.net.nB)
        lq = np.zeros(max(self.net.nPV, 1), np.uint8)
        _chk(lib().pf_solver_get_state(self.h, _p(V), _p(th), _p(lq)), "pf_solver_get_state")
        return V, th, lq[:self.net.nPV].astype(bool)

    def lu_info(self):
        v = [C.c_int() for _ in range(4)]
        _chk(lib().pf_solver_lu_info(self.h, *[C.byref(x) for x in v]), "pf_solver_lu_info")
        return dict(n=v[0].value, nnzL=v[1].value, levels=v[2].value, tail_levels=v[3].value)

    # ---- stages
    def ybus(self):
        nnz = C.c_int()
        _chk(lib().pf_stage_ybus(self.h, None, None, None, C.byref(nnz)), "pf_stage_ybus")
        rp = np.zeros(self.net.nB + 1, np.int32)
        ci = np.zeros(nnz.value, np.int32)
        y = np.zeros(2 * nnz.value)
        _chk(lib().pf_stage_ybus(self.h, _p(rp), _p(ci), _p(y), None), "pf_stage_ybus")
        return rp, ci, y[0::2] + 1j * y[1::2]

    def pq(self, V, theta):
        V = np.ascontiguousarray(V, np.float64)
        th = np.ascontiguousarray(theta, np.float64)
        P = np.zeros(self.net.nB)
        Q = np.zeros(self.net.nB)
        _chk(lib().pf_stage_pq(self.h, _p(V), _p(th), _p(P), _p(Q)), "pf_stage_pq")
        return P, Q

    def flows(self, V=None, theta=None):
        """Branch flows as the reference reports them, (Pdp, Qdp, Ppd, Qpd)
        [nL]: at the given state, or at the solver's current one."""
        out = [np.zeros(self.net.nL) for _ in range(4)]
        if V is None:
            _chk(lib().pf_solver_flows(self.h, *[_p(a) for a in out]), "pf_solver_flows")
        else:
            V = np.ascontiguousarray(V, np.float64)
            th = np.ascontiguousarray(theta, np.float64)
            _chk(lib().pf_stage_flows(self.h, _p(V), _p(th), *[_p(a) for a in out]), "pf_stage_flows")
        return tuple(out)

    def _lq(self, limQ):
        if limQ is None:
            limQ = np.zeros(self.net.nPV, bool)
        return np.ascontiguousarray(np.asarray(limQ, np.uint8).reshape(-1)[: self.net.nPV].copy()
                                    if self.net.nPV else np.zeros(1, np.uint8))

    def mismatch(self, V, theta, limQ, QliqLim):
        lq = self._lq(limQ)
        V = np.ascontiguousarray(V, np.float64)
        th = np.ascontiguousarray(theta, np.float64)
        ql = np.ascontiguousarray(QliqLim, np.float64)
        n = self.net.nB - 1 + self.net.nPQ + int(lq[: self.net.nPV].sum())
        g = np.zeros(n)
        m = C.c_double()
        _chk(lib().pf_stage_mismatch(self.h, _p(V), _p(th), _p(lq), _p(ql), _p(g), C.byref(m)),
             "pf_stage_mismatch")
        return g, m.value

    def jacobian(self, V, theta, limQ=None):
        lq = self._lq(limQ)
        V = np.ascontiguousarray(V, np.float64)
        th = np.ascontiguousarray(theta, np.float64)
        nJ, nnz = C.c_int(), C.c_int()
        cap = 4 * (self.net.nB + 2 * self.net.nL) + 16
        rp = np.zeros(2 * self.net.nB + 1, np.int32)
        cj = np.zeros(cap, np.int32)
        vj = np.zeros(cap)
        _chk(lib().pf_stage_jacobian(self.h, _p(V), _p(th), _p(lq), C.byref(nJ), _p(rp), _p(cj),
                                     _p(vj), cap, C.byref(nnz)), "pf_stage_jacobian")
        return rp[: nJ.value + 1], cj[: nnz.value], vj[: nnz.value]

    def lu_solve(self, V, theta, limQ, rhs):
        lq = self._lq(limQ)
        V = np.ascontiguousarray(V, np.float64)
        th = np.ascontiguousarray(theta, np.float64)
        b = np.ascontiguousarray(rhs, np.float64)
        x = np.zeros_like(b)
        _chk(lib().pf_stage_lu_solve(self.h, _p(V), _p(th), _p(lq), _p(b), _p(x)), "pf_stage_lu_solve")
        return x

    def limits(self, mode, V, Q, limQ, QliqLim):
        lq = self._lq(limQ).copy()
        ql = np.ascontiguousarray(QliqLim, np.float64).copy()
        V = np.ascontiguousarray(V, np.float64)
        Q = np.ascontiguousarray(Q, np.float64)
        f = C.c_int()
        _chk(lib().pf_stage_limits(self.h, mode, _p(V), _p(Q), _p(lq), _p(ql), C.byref(f)),
             "pf_stage_limits")
        return lq[: self.net.nPV].astype(bool), ql, f.value


class Batch:
    """Many scenarios of one or more networks in one persistent launch."""

    def __init__(self, nets, opts: Options, device: int = 0):
        arr = (C.c_void_p * len(nets))(*[n.h for n in nets])
        h = C.c_void_p()
        _chk(lib().pf_batch_create(arr, len(nets), C.byref(opts), device, C.byref(h)), "pf_batch_create")
        self.h, self.nets = h, list(nets)
        self.n_scen = 0

    def __del__(self):
        if getattr(self, "h", None) is not None and _lib is not None:
            _lib.pf_batch_destroy(self.h)
            self.h = None

    def set_scenarios(self, net_of, Pliq, Qliq, Qload, V0=None, theta0=None):
        net_of = np.ascontiguousarray(net_of, np.int32)
        a = [np.ascontiguousarray(x, np.float64) for x in (Pliq, Qliq, Qload)]
        v0 = None if V0 is None else np.ascontiguousarray(V0, np.float64)
        t0 = None if theta0 is None else np.ascontiguousarray(theta0, np.float64)
        _chk(lib().pf_batch_set_scenarios(self.h, len(net_of), _p(net_of), *[_p(x) for x in a],
                                          _p(v0), _p(t0)), "pf_batch_set_scenarios")
        self.n_scen = len(net_of)
        self.net_of = net_of
        self.total = int(sum(self.nets[i].nB for i in net_of))

    def solve(self) -> float:
        ms = C.c_double()
        _chk(lib().pf_batch_solve(self.h, C.byref(ms)), "pf_batch_solve")
        return ms.value

    def autotune(self, cache_path=None, reps: int = 2) -> int:
        """Pick the fastest launch shape per network for the current
        scenarios (or apply a cached one); returns the number of candidate
        shapes measured."""
        n = C.c_int()
        _chk(lib().pf_batch_autotune(self.h, None if cache_path is None else os.fsencode(cache_path),
                                     reps, C.byref(n)), "pf_batch_autotune")
        return n.value

    def set_shape(self, net: int, mode: int, block: int, per_cu: int = 0):
        _chk(lib().pf_batch_set_shape(self.h, net, mode, block, per_cu), "pf_batch_set_shape")

    def shape(self, net: int) -> dict:
        m, b, g = C.c_int(), C.c_int(), C.c_int()
        lds = C.c_longlong()
        _chk(lib().pf_batch_get_shape(self.h, net, C.byref(m), C.byref(b), C.byref(g), C.byref(lds)),
             "pf_batch_get_shape")
        return {"mode": m.value, "block": b.value, "grid": g.value, "lds_bytes": lds.value}

    def results(self):
        V = np.zeros(self.total)
        th = np.zeros(self.total)
        it = np.zeros(self.n_scen, np.int32)
        cv = np.zeros(self.n_scen, np.int32)
        mm = np.zeros(self.n_scen)
        _chk(lib().pf_batch_get(self.h, _p(V), _p(th), _p(it), _p(cv), _p(mm)), "pf_batch_get")
        return V, th, it, cv.astype(bool), mm

    def limq(self):
        """Final PV->PQ state per scenario: (list of bool arrays [nPV of the
        scenario's network], flips per scenario)."""
        npv = [self.nets[i].nPV for i in self.net_of]
        lq = np.zeros(max(1, sum(npv)), np.uint8)
        fl = np.zeros(self.n_scen, np.int32)
        _chk(lib().pf_batch_get_limq(self.h, _p(lq), _p(fl)), "pf_batch_get_limq")
        out, o = [], 0
        for k in npv:
            out.append(lq[o:o + k].astype(bool))
            o += k
        return out, fl


class WarmStart:
    """The learned warm start (MLP) trained and run by libpf's own kernels."""

    def __init__(self, n_in: int, n_out: int, hidden: int = 512, n_hidden_layers: int = 3,
                 seed: int = 1, device: int = 0):
        h = C.c_void_p()
        _chk(lib().pf_warmstart_create(n_in, n_out, hidden, n_hidden_layers, seed, device, C.byref(h)),
             "pf_warmstart_create")
        self.h, self.n_in, self.n_out = h, n_in, n_out
        n = C.c_size_t()
        nl = C.c_int()
        lib().pf_warmstart_info(h, C.byref(n), C.byref(nl))
        self.n_params, self.n_layers = n.value, nl.value
        self.hidden, self.n_hidden_layers = hidden, n_hidden_layers

    def __del__(self):
        if getattr(self, "h", None) is not None and _lib is not None:
            _lib.pf_warmstart_destroy(self.h)
            self.h = None

    def layer_shapes(self):
        K = [self.n_in] + [self.hidden] * self.n_hidden_layers
        N = [self.hidden] * self.n_hidden_layers + [self.n_out]
        return list(zip(K, N))

    def set_norm(self, xm, xs, ym, ys):
        a = [np.ascontiguousarray(v, np.float32) for v in (xm, xs, ym, ys)]
        _chk(lib().pf_warmstart_set_norm(self.h, *[_p(v) for v in a]), "pf_warmstart_set_norm")

    def params(self) -> np.ndarray:
        out = np.zeros(self.n_params, np.float32)
        _chk(lib().pf_warmstart_get_params(self.h, _p(out)), "pf_warmstart_get_params")
        return out

    def set_params(self, p):
        p = np.ascontiguousarray(p, np.float32)
        assert p.size == self.n_params
        _chk(lib().pf_warmstart_set_params(self.h, _p(p)), "pf_warmstart_set_params")

    def grad(self, X, Y):
        X = np.ascontiguousarray(X, np.float32)
        Y = np.ascontiguousarray(Y, np.float32)
        g = np.zeros(self.n_params, np.float32)
        loss = C.c_double()
        _chk(lib().pf_warmstart_grad(self.h, _p(X), _p(Y), X.shape[0], _p(g), C.byref(loss)),
             "pf_warmstart_grad")
        return g, loss.value

    def fit(self, X, Y, epochs=10, batch=256, lr=1e-3, weight_decay=1e-4, seed=7):
        X = np.ascontiguousarray(X, np.float32)
        Y = np.ascontiguousarray(Y, np.float32)
        el = np.zeros(epochs)
        _chk(lib().pf_warmstart_fit(self.h, _p(X), _p(Y), X.shape[0], epochs, batch, lr, weight_decay, seed,
                                    _p(el)), "pf_warmstart_fit")
        return el

    def predict(self, X) -> np.ndarray:
        X = np.ascontiguousarray(X, np.float32)
        Y = np.zeros((X.shape[0], self.n_out), np.float32)
        _chk(lib().pf_warmstart_predict(self.h, _p(X), X.shape[0], _p(Y)), "pf_warmstart_predict")
        return Y

    def stage(self, X):
        """Features of every scenario of a batch to the GPU, once."""
        X = np.ascontiguousarray(X, np.float32)
        _chk(lib().pf_warmstart_stage(self.h, _p(X), X.shape[0]), "pf_warmstart_stage")

    def predict_into(self, batch, v_col: int, th_col: int):
        """Inference on the staged rows, written into `batch` as each scenario's
        starting |V| (row columns v_col..) and theta (th_col..)."""
        _chk(lib().pf_warmstart_predict_into(self.h, batch.h, v_col, th_col), "pf_warmstart_predict_into")

    def gemm_probe(self, A, B, ta=False, tb=False):
        A = np.ascontiguousarray(A, np.float32)
        B = np.ascontiguousarray(B, np.float32)
        M = A.shape[1] if ta else A.shape[0]
        K = A.shape[0] if ta else A.shape[1]
        N = B.shape[0] if tb else B.shape[1]
        Cm = np.zeros((M, N), np.float32)
        _chk(lib().pf_warmstart_gemm_probe(self.h, int(ta), int(tb), M, N, K, _p(A), _p(B), _p(Cm)),
             "pf_warmstart_gemm_probe")
        return Cm
