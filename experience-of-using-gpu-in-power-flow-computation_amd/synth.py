"""Synthetic power-flow cases written in IEEE Common Data Format (CDF / CDFX).

The written file is the canonical input: every value is rounded to the CDF
field it is written into, so the readers (the product's C++ reader and the
oracle's C reader) see exactly the same numbers the file holds.

Column layout follows the reader the reference ships
(CPU/leitura/cdf.cpp:163-327 for CDF, :329-493 for CDFX; 0-based
``substr(start, len)``):

=====================  ============  ============
field                  CDF           CDFX
=====================  ============  ============
header baseMVA         (31, 6)       (31, 6)
bus id                 (0, 4)        (0, 5)
bus type               (24, 2)       (25, 2)
final |V|              (27, 6)       (28, 6)
final angle (deg)      (33, 7)       (34, 7)
load MW / MVAr         (40,9)/(49,10) (41,9)/(50,10)
gen MW / MVAr          (59,8)/(67,8)  (60,8)/(68,8)
base kV                (76, 7)       (77, 7)
Qmax / Qmin MVAr       (90,7)/(98,7)  (91,7)/(99,7)
shunt G / B (pu)       (106,8)/(114,8) (107,8)/(115,8)
branch from / to       (0,4)/(4,5)    (0,5)/(6,5)
R / X / B              (19,10)/(29,11)/(40,10)  (21,10)/(31,11)/(42,10)
tap / phase (deg)      (76,6)/(83,7)  (78,6)/(85,7)
=====================  ============  ============

A line holding the substring ``-999`` ends a section in that reader
(CPU/leitura/cdf.cpp:182, 250), so the writer refuses to emit one.

Randomness comes from SplitMix64 (Steele, Lea & Flood, "Fast splittable
pseudorandom number generators", OOPSLA 2014), so a case is a pure function of
its seed.

Workloads
---------
* ``case14(seed=14)``: 14 buses (1 slack, 4 PV, 9 PQ), 20 branches = random
  spanning tree + 7 chords, 3 tap transformers.
* ``case118(seed=118)``: 118 buses (1 slack, 53 PV, 64 PQ), 186 branches
  (117 tree + 69 chords), 9 taps, total load 42 pu.
* ``multi_copy(n_copies, copy_seed0, n_ties, tie_seed)``: copies of the
  118-bus network joined by tie lines (X ~ U(0.02, 0.1), R = X/10, B = 0).
  Copy 0 keeps the slack; the other copies' slacks become PV buses whose
  active dispatch is that copy's stand-alone slack output, so every copy is
  balanced and the ties carry little power.  128 copies = 15,104 buses (CDFX).
* ``scenario_loads(...)``: per-scenario load factors clip(exp(0.15 N(0,1)),
  0.7, 1.3) per bus, PV dispatch rescaled by the ratio of total loads.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

MASK64 = (1 << 64) - 1


class SplitMix64:
    """SplitMix64 generator (64-bit state, golden-gamma increment)."""

    def __init__(self, seed: int):
        self.state = seed & MASK64

    def next_u64(self) -> int:
        self.state = (self.state + 0x9E3779B97F4A7C15) & MASK64
        z = self.state
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & MASK64
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & MASK64
        return z ^ (z >> 31)

    def uniform(self, lo: float = 0.0, hi: float = 1.0) -> float:
        u = (self.next_u64() >> 11) * (1.0 / 9007199254740992.0)
        return lo + (hi - lo) * u

    def randint(self, n: int) -> int:
        """Uniform integer in [0, n)."""
        return int(self.uniform() * n) % n

    def normal(self) -> float:
        """Box-Muller standard normal deviate."""
        u1 = self.uniform()
        u2 = self.uniform()
        return math.sqrt(-2.0 * math.log(1.0 - u1)) * math.cos(2.0 * math.pi * u2)

    def permutation(self, n: int) -> list:
        p = list(range(n))
        for i in range(n - 1, 0, -1):
            j = self.randint(i + 1)
            p[i], p[j] = p[j], p[i]
        return p


# --------------------------------------------------------------------------
# Network description (values in pu, already rounded to the CDF fields)
# --------------------------------------------------------------------------

PQ, PV, SLACK = 0, 2, 3


@dataclass
class Bus:
    id: int
    type: int
    v: float = 1.0          # |V| final/setpoint (pu)
    ang: float = 0.0        # degrees
    pload: float = 0.0      # MW
    qload: float = 0.0      # MVAr
    pg: float = 0.0         # MW
    qg: float = 0.0         # MVAr
    qmax: float = 0.0       # MVAr
    qmin: float = 0.0       # MVAr
    gsh: float = 0.0        # pu
    bsh: float = 0.0        # pu
    basekv: float = 138.0


@dataclass
class Branch:
    f: int                  # from bus id
    t: int                  # to bus id
    r: float
    x: float
    b: float = 0.0
    tap: float = 0.0        # 0 = no transformer
    phase: float = 0.0      # degrees


@dataclass
class Case:
    buses: list = field(default_factory=list)
    branches: list = field(default_factory=list)
    base_mva: float = 100.0
    title: str = "SYNTHETIC"


def _r(x: float, nd: int) -> float:
    """Round to what a fixed-point CDF field of ``nd`` decimals holds."""
    return float(f"{x:.{nd}f}")


# --------------------------------------------------------------------------
# Single-area generator (the 14- and 118-bus workloads)
# --------------------------------------------------------------------------

def _area(seed: int, n_bus: int, n_pv: int, n_chords: int, n_taps: int,
          total_load_pu: float | None, id0: int = 0, pv_share: float = 0.84,
          load_on_pv: bool = False, qlim_scale: float = 0.0,
          root_at_slack: bool = False) -> Case:
    rng = SplitMix64(seed)
    base = 100.0
    # bus 1 (local index 0) is the slack; PV buses drawn from the rest
    others = list(range(1, n_bus))
    perm = rng.permutation(len(others))
    pv_set = set(others[perm[i]] for i in range(n_pv))
    buses = []
    for i in range(n_bus):
        typ = SLACK if i == 0 else (PV if i in pv_set else PQ)
        buses.append(Bus(id=id0 + i + 1, type=typ))
    buses[0].v = 1.04
    # loads on PQ buses (and on PV buses when load_on_pv)
    for b in buses:
        if b.type == PQ or (load_on_pv and b.type == PV):
            p = rng.uniform(0.05, 0.5)
            pf = rng.uniform(0.85, 0.98)
            b.pload = p
            b.qload = p * math.tan(math.acos(pf))
    ptot = sum(b.pload for b in buses)
    scale = (total_load_pu / ptot) if total_load_pu else 1.0
    for b in buses:
        b.pload = _r(b.pload * scale * base, 2)
        b.qload = _r(b.qload * scale * base, 2)
    ptot_mw = sum(b.pload for b in buses)
    # PV setpoints, limits, dispatch
    pvs = [b for b in buses if b.type == PV]
    w = [rng.uniform(0.5, 1.5) for _ in pvs]
    wsum = sum(w)
    pv_total = pv_share * ptot_mw
    qs = scale * qlim_scale if qlim_scale else 1.0
    for b, wi in zip(pvs, w):
        b.v = _r(rng.uniform(0.98, 1.06), 4)
        b.qmax = _r(rng.uniform(0.2, 0.5) * qs * base, 2)
        b.qmin = _r(-rng.uniform(0.2, 0.5) * qs * base, 2)
        b.pg = _r(pv_total * wi / wsum, 2)
    buses[0].pg = _r(1.05 * ptot_mw - sum(b.pg for b in pvs), 2)
    # topology: random recursive spanning tree (rooted at the slack, bus 1
    # of the area, when root_at_slack: the slack is then the best-connected
    # bus) + distinct chords
    if root_at_slack:
        order = [0] + [1 + p for p in rng.permutation(n_bus - 1)]
    else:
        order = rng.permutation(n_bus)
    pairs = []
    seen = set()
    for k in range(1, n_bus):
        a = order[k]
        b = order[rng.randint(k)]
        key = (min(a, b), max(a, b))
        seen.add(key)
        pairs.append(key)
    while len(pairs) < (n_bus - 1) + n_chords:
        a = rng.randint(n_bus)
        b = rng.randint(n_bus)
        if a == b:
            continue
        key = (min(a, b), max(a, b))
        if key in seen:
            continue
        seen.add(key)
        pairs.append(key)
    branches = []
    for (a, b) in pairs:
        r = rng.uniform(0.01, 0.08)
        x = r * rng.uniform(2.0, 5.0)
        bc = rng.uniform(0.0, 0.05)
        branches.append(Branch(f=id0 + a + 1, t=id0 + b + 1, r=_r(r, 5), x=_r(x, 5), b=_r(bc, 5)))
    tperm = rng.permutation(len(branches))
    for k in range(n_taps):
        br = branches[tperm[k]]
        br.tap = _r(rng.uniform(0.95, 1.05), 4)
        br.b = 0.0          # B = 0 on transformers: MATPOWER and "eficiente" Ybus forms agree
    return Case(buses=buses, branches=branches, base_mva=base)


def case14(seed: int = 14) -> Case:
    c = _area(seed, 14, 4, 7, 3, None)
    c.title = f"SYNTHETIC 14 BUS seed {seed}"
    return c


def case118(seed: int = 118) -> Case:
    c = area118(seed)
    c.title = f"SYNTHETIC 118 BUS seed {seed}"
    return c


# --------------------------------------------------------------------------
# Multi-copy networks (15,104-bus single solve; 1,888- and 944-bus variants)
# --------------------------------------------------------------------------

def _ybus_dense(case: Case) -> np.ndarray:
    ids = {b.id: i for i, b in enumerate(case.buses)}
    n = len(case.buses)
    Y = np.zeros((n, n), dtype=complex)
    for br in case.branches:
        f, t = ids[br.f], ids[br.t]
        y = 1.0 / complex(br.r, br.x)
        a = br.tap if br.tap != 0.0 else 1.0
        tap = a * complex(math.cos(math.radians(br.phase)), math.sin(math.radians(br.phase)))
        Y[f, f] += (y + 0.5j * br.b) / abs(tap) ** 2
        Y[f, t] += -y / tap.conjugate()
        Y[t, f] += -y / tap
        Y[t, t] += y + 0.5j * br.b
    for i, b in enumerate(case.buses):
        Y[i, i] += complex(b.gsh, b.bsh)
    return Y


def _screen_solve(case: Case, q_limits: bool, load_scale: float = 1.0):
    """Generator-side screening solve: dense polar NR (MATPOWER complex
    derivatives) with, optionally, the textbook PV->PQ outer loop (a PV bus
    whose Q leaves its limits is fixed at the limit and the case re-solved).

    Returns (converged, Vm, Va, S) -- used only to accept/reject a candidate
    network and to balance copies before they are tied together.
    """
    Y = _ybus_dense(case)
    n = len(case.buses)
    base = case.base_mva
    types = np.array([b.type for b in case.buses])
    sl = int(np.nonzero(types == SLACK)[0][0])
    Vm = np.ones(n)
    Va = np.zeros(n)
    vset = np.array([b.v for b in case.buses])
    Vm[types != PQ] = vset[types != PQ]
    pl = np.array([b.pload for b in case.buses]) * load_scale
    ql = np.array([b.qload for b in case.buses]) * load_scale
    pg = np.array([b.pg for b in case.buses]) * load_scale
    qg = np.array([b.qg for b in case.buses])
    qmax = np.array([b.qmax for b in case.buses])
    qmin = np.array([b.qmin for b in case.buses])
    Sspec = ((pg - pl) + 1j * (qg - ql)) / base
    is_pq = types == PQ
    for _outer in range(12):
        pv = [i for i in range(n) if i != sl and not is_pq[i]]
        pq = [i for i in range(n) if is_pq[i]]
        pvpq = sorted(pv + pq)
        ok = False
        for _ in range(20):
            V = Vm * np.exp(1j * Va)
            I = Y @ V
            S = V * np.conj(I)
            mis = S - Sspec
            F = np.concatenate([mis[pvpq].real, mis[pq].imag])
            if not np.all(np.isfinite(F)):
                break
            if np.max(np.abs(F)) < 1e-9:
                ok = True
                break
            dS_dVa = 1j * V[:, None] * np.conj(np.diag(I) - Y * V[None, :])
            dS_dVm = V[:, None] * np.conj(Y * (V / Vm)[None, :]) + np.diag(np.conj(I) * V / Vm)
            J = np.block([[dS_dVa[np.ix_(pvpq, pvpq)].real, dS_dVm[np.ix_(pvpq, pq)].real],
                          [dS_dVa[np.ix_(pq, pvpq)].imag, dS_dVm[np.ix_(pq, pq)].imag]])
            try:
                dx = np.linalg.solve(J, -F)
            except np.linalg.LinAlgError:
                break
            Va[pvpq] += dx[:len(pvpq)]
            Vm[pq] += dx[len(pvpq):]
        if not ok:
            return False, Vm, Va, None
        if not q_limits:
            return True, Vm, Va, S
        qgen = S.imag * base + ql
        viol = [i for i in pv if qgen[i] > qmax[i] + 1e-2 or qgen[i] < qmin[i] - 1e-2]
        if not viol:
            return True, Vm, Va, S
        for i in viol:
            is_pq[i] = True
            qfix = qmax[i] if qgen[i] > qmax[i] else qmin[i]
            Sspec[i] = Sspec[i].real + 1j * (qfix - ql[i]) / base
    return False, Vm, Va, None


def _healthy(case: Case) -> bool:
    """Acceptance test of a candidate 118-bus area: solvable with margin.

    * no Q limits: converges with min |V| >= 0.85 and max |angle| <= 60 deg;
    * with Q limits (textbook switching): converges with min |V| >= 0.80;
    * every load and dispatch x1.3: converges with min |V| >= 0.70 (room for
      the +30 % scenario perturbations of the batched workloads).
    """
    ok, Vm, Va, _ = _screen_solve(case, False)
    if not ok or Vm.min() < 0.85 or np.degrees(np.abs(Va)).max() > 60.0:
        return False
    ok, Vm, _, _ = _screen_solve(case, True)
    if not ok or Vm.min() < 0.80:
        return False
    ok, Vm, _, _ = _screen_solve(case, False, load_scale=1.3)
    return bool(ok and Vm.min() >= 0.70)


def _slack_dispatch_mw(case: Case) -> float:
    """Active output (MW) of the slack after a stand-alone screening solve."""
    ok, _, _, S = _screen_solve(case, False)
    if not ok:
        raise RuntimeError("stand-alone copy did not converge")
    sl = [i for i, b in enumerate(case.buses) if b.type == SLACK][0]
    return float(S[sl].real * case.base_mva + case.buses[sl].pload)


def area118(seed: int, id0: int = 0) -> Case:
    """118-bus area for ``seed``: candidates drawn from seeds
    seed, seed + 1_000_003, seed + 2 * 1_000_003, ... until one passes
    :func:`_healthy` (rejection sampling keeps the stated distributions)."""
    for attempt in range(64):
        c = _area(seed + attempt * 1_000_003, 118, 53, 69, 9, 42.0, id0=id0,
                  pv_share=1.0, load_on_pv=True, qlim_scale=2.0, root_at_slack=True)
        if _healthy(c):
            return c
    raise RuntimeError(f"no healthy 118-bus area found for seed {seed}")


def multi_copy(n_copies: int, copy_seed0: int, n_ties: int, tie_seed: int,
               title: str | None = None, loss_share: float = 0.0,
               slack_ties: bool = True) -> Case:
    """Copies of the 118-bus area tied together.

    Copy c is ``area118(copy_seed0 + c)`` with bus ids offset by 118 c.  Copy
    0's slack is the global slack; every other copy's slack becomes a PV bus
    dispatched at its stand-alone slack output minus ``loss_share`` times that
    copy's losses (the global slack supplies that share).  Ties: a random
    recursive spanning tree over the copies rooted at copy 0, then random
    extra pairs of distinct copies; each tie joins a random bus of each copy
    (with ``slack_ties``, a tie touching copy 0 lands on the global slack).

    Why ``slack_ties``: from a flat start the first Newton step has to carry
    every copy's losses (about 1 pu each, since each copy is balanced at its
    stand-alone solution) into the single slack bus; through copy 0's own
    branches that step is several radians wide and NR diverges at 32 copies
    and up.  Landing copy 0's ties on the slack bus lets 128 copies converge
    from flat start (6 iterations without Q limits, 13 with them).
    """
    out = Case(title=title or f"SYNTHETIC {118 * n_copies} BUS ({n_copies} x 118)")
    for c in range(n_copies):
        area = area118(copy_seed0 + c, id0=118 * c)
        if c > 0:
            pg = _slack_dispatch_mw(area)
            losses = pg + sum(b.pg for b in area.buses[1:]) - sum(b.pload for b in area.buses)
            rng_c = SplitMix64((copy_seed0 + c) ^ 0x5A5A5A5A)
            sl = area.buses[0]
            sl.type = PV
            sl.pg = _r(pg - loss_share * losses, 2)
            sl.qmax = _r(rng_c.uniform(0.2, 0.5) * 2.0 * 100.0, 2)
            sl.qmin = _r(-rng_c.uniform(0.2, 0.5) * 2.0 * 100.0, 2)
        out.buses.extend(area.buses)
        out.branches.extend(area.branches)
    rng = SplitMix64(tie_seed)
    seen = set()
    ties = []

    def end(cc):
        if slack_ties and cc == 0:
            return 1
        return 118 * cc + rng.randint(118) + 1

    def add_tie(ca, cb):
        for _ in range(1000):
            a, b = end(ca), end(cb)
            key = (min(a, b), max(a, b))
            if key not in seen:
                seen.add(key)
                return key
        raise RuntimeError("could not place a distinct tie")

    corder = [0] + [1 + p for p in rng.permutation(n_copies - 1)]
    for k in range(1, n_copies):
        ties.append(add_tie(corder[k], corder[rng.randint(k)]))
    while len(ties) < n_ties:
        ca = rng.randint(n_copies)
        cb = rng.randint(n_copies)
        if ca == cb:
            continue
        ties.append(add_tie(ca, cb))
    for (a, b) in ties:
        x = rng.uniform(0.02, 0.1)
        out.branches.append(Branch(f=a, t=b, r=_r(x / 10.0, 5), x=_r(x, 5), b=0.0))
    return out


def case15104() -> Case:
    return multi_copy(128, 118000, 256, 15104, "SYNTHETIC 15104 BUS (128 x 118)")


def case1888() -> Case:
    return multi_copy(16, 1888000, 32, 1888, "SYNTHETIC 1888 BUS (16 x 118)")


def case944() -> Case:
    return multi_copy(8, 944000, 16, 944, "SYNTHETIC 944 BUS (8 x 118)")


# --------------------------------------------------------------------------
# Writer
# --------------------------------------------------------------------------

def _check(line: str) -> str:
    if "-999" in line:
        raise ValueError(f"field would end a CDF section: {line!r}")
    return line


def _bus_line(b: Bus, wide: bool) -> str:
    idw = 5 if wide else 4
    s = f"{b.id:>{idw}d} {('B' + str(b.id))[:12]:<12}  1  1 {b.type:>2d} "
    assert len(s) == (28 if wide else 27), s      # |V| field starts here
    vdes = b.v if b.type != PQ else 0.0
    s += f"{b.v:6.4f}{b.ang:7.2f}{b.pload:9.2f}{b.qload:10.2f}{b.pg:8.2f}{b.qg:8.2f} "
    s += f"{b.basekv:7.1f} {vdes:6.4f}{b.qmax:7.2f} {b.qmin:7.2f} "
    s += f"{b.gsh:8.4f}{b.bsh:8.4f} {0:>4d}"
    return _check(s)


def _branch_line(br: Branch, wide: bool) -> str:
    idw = 5 if wide else 4
    s = f"{br.f:>{idw}d} {br.t:>{idw}d}  1  1 1 {1 if br.tap == 0 else 2}"
    assert len(s) == (21 if wide else 19), s      # R field starts here
    s += f"{br.r:10.5f}{br.x:11.5f}{br.b:10.5f}"
    s += "{:>5d}{:>6d}{:>6d}{:>5d}{:>2d}  ".format(0, 0, 0, 0, 0)
    s += f"{br.tap:6.4f} {br.phase:7.2f}"
    s += " 0.0    0.0     0.0    0.0    0.0"
    return _check(s)


def write_cdf(case: Case, path: str, wide: bool | None = None) -> None:
    """Write ``case`` in CDF (ids < 10000) or CDFX (``wide``) layout."""
    if wide is None:
        wide = max(b.id for b in case.buses) > 9999
    lines = []
    hdr = f"{' 01/01/26 SYNTHETIC':<31}{case.base_mva:<6.1f} 2026 W {case.title}"
    assert hdr[31:37].strip().startswith(str(int(case.base_mva))), hdr
    lines.append(hdr)
    lines.append(f"BUS DATA FOLLOWS                            {len(case.buses)} ITEMS")
    lines += [_bus_line(b, wide) for b in case.buses]
    lines.append("-999")
    lines.append(f"BRANCH DATA FOLLOWS                         {len(case.branches)} ITEMS")
    lines += [_branch_line(br, wide) for br in case.branches]
    lines.append("-999")
    lines.append("LOSS ZONES FOLLOWS                     1 ITEMS")
    lines.append("  1 SYNTHETIC")
    lines.append("-99")
    lines.append("END OF DATA")
    with open(path, "w") as fh:
        fh.write("\n".join(lines) + "\n")


# --------------------------------------------------------------------------
# Scenario perturbations (batched workloads, warm-start training data)
# --------------------------------------------------------------------------

_GAMMA = np.uint64(0x9E3779B97F4A7C15)


def _mix64(z: np.ndarray) -> np.ndarray:
    z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
    z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
    return z ^ (z >> np.uint64(31))


def load_factors_batch(n_bus: int, seeds) -> np.ndarray:
    """Per-bus load factors of many scenarios at once: row s is the factor
    vector of SplitMix64 stream ``seeds[s]`` (k-th output = mix(seed + k*gamma)),
    two uniforms per bus through Box-Muller, clip(exp(0.15 N(0,1)), 0.7, 1.3)."""
    seeds = np.asarray(seeds, dtype=np.uint64).reshape(-1, 1)
    k = np.arange(1, 2 * n_bus + 1, dtype=np.uint64).reshape(1, -1)
    with np.errstate(over="ignore"):
        z = _mix64(seeds + k * _GAMMA)
    u = (z >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)
    u1, u2 = u[:, 0::2], u[:, 1::2]
    nrm = np.sqrt(-2.0 * np.log1p(-u1)) * np.cos(2.0 * np.pi * u2)
    return np.clip(np.exp(0.15 * nrm), 0.7, 1.3)


def load_factors(n_bus: int, seed: int) -> np.ndarray:
    """clip(exp(0.15 * N(0,1)), 0.7, 1.3) per bus for one scenario."""
    return load_factors_batch(n_bus, [seed])[0]


def scenario_injections(pg: np.ndarray, pload: np.ndarray, qg: np.ndarray, qload: np.ndarray,
                        is_pv: np.ndarray, seed: int):
    """Net injections (pu) of one perturbed scenario.

    Every bus load (P and Q) is scaled by its factor; PV active dispatch is
    scaled by the ratio of the new to the old total active load.
    Returns (Pliq, Qliq, Qload') arrays over buses.
    """
    P, Q, L = scenario_batch(pg, pload, qg, qload, is_pv, [seed])
    return P[0], Q[0], L[0]


def scenario_batch(pg, pload, qg, qload, is_pv, seeds):
    """Vectorised :func:`scenario_injections` over many seeds -> (Pliq, Qliq, Qload'), each [n, nB]."""
    f = load_factors_batch(len(pload), seeds)
    pl = pload[None, :] * f
    ql = qload[None, :] * f
    tot = pload.sum()
    ratio = pl.sum(1, keepdims=True) / tot if tot != 0 else np.ones((len(f), 1))
    pg2 = np.where(is_pv[None, :], pg[None, :] * ratio, pg[None, :])
    return pg2 - pl, qg[None, :] - ql, ql
