"""Host-side checks of libpf that need no GPU: the C ABI exports, the reader
(bit-exact against the oracle's), the data.txt parser and its quirks, the
ordering / symbolic factor program, and the case generator."""
import os
import re
import subprocess

import numpy as np
import pytest

from conftest import GOLDEN, IEEE, PKG_DIR, ROOT

HEADER = os.path.join(ROOT, "include", "pf.h")


def test_exports_every_declared_symbol(pf):
    text = open(HEADER).read()
    names = set(re.findall(r"^\s*(?:const\s+char\s*\*|pf_status|void)\s*\*?\s*(pf_\w+)\s*\(", text, re.M))
    assert len(names) >= 30
    out = subprocess.run(["nm", "-D", "--defined-only", os.path.join(PKG_DIR, "libpf.so")],
                         capture_output=True, text=True, check=True).stdout
    exported = set(re.findall(r"\sT\s(pf_\w+)", out))
    missing = names - exported
    assert not missing, missing
    lib = pf.lib()
    for n in names:
        assert hasattr(lib, n)


def _paths(cases):
    return [os.path.join(IEEE, f"ieee{n}.cdf") for n in (14, 30, 57, 118)] + \
           [os.path.join(GOLDEN, "c14.cdf"), os.path.join(GOLDEN, "c118.cdf"), cases.path("c944")]


def test_reader_parity_with_oracle(pf, oracle, cases):
    for p in _paths(cases):
        a = pf.Network(p)
        b = oracle.Case(p)
        assert (a.nB, a.nL, a.nPV, a.nPQ, a.slack, a.n_duplicates) == \
               (b.nB, b.nL, b.nPV, b.nPQ, b.slack, b.n_duplicates)
        for f in ["V", "theta", "Pliq", "Qliq", "Pload", "Qload", "Pg", "Qg", "Vbase", "gsh", "bsh",
                  "Qmin", "Qmax", "Vset", "r", "x", "bch", "tap_re", "tap_im"]:
            assert np.array_equal(getattr(a, f), getattr(b, f)), (p, f)
        for f in ["pv", "pq", "frm", "to"]:
            assert np.array_equal(getattr(a, f), getattr(b, f)), (p, f)


def test_cdfx_layout(pf, oracle, synth, tmp_path):
    c = synth.case14()
    p1, p2 = str(tmp_path / "a.cdf"), str(tmp_path / "a.cdfx")
    synth.write_cdf(c, p1, wide=False)
    synth.write_cdf(c, p2, wide=True)
    a, b = pf.Network(p1), pf.Network(p2)
    for f in ["V", "Pliq", "Qliq", "Qmax", "r", "x", "bch", "tap_re"]:
        assert np.array_equal(getattr(a, f), getattr(b, f))
    assert np.array_equal(oracle.Case(p2).x, b.x)


@pytest.mark.parametrize("name", ["ieee14", "ieee118"])
def test_reader_cdfx_recolumned(pf, name):
    """libpf reads the same network from an IEEE CDF file and from its CDFX
    re-columning (tests/golden/make_cdfx.py)."""
    a = pf.Network(os.path.join(IEEE, name + ".cdf"))
    b = pf.Network(os.path.join(IEEE, name + ".cdfx"))
    assert (a.nB, a.nL, a.nPV, a.nPQ, a.slack) == (b.nB, b.nL, b.nPV, b.nPQ, b.slack)
    for attr in ("V", "theta", "Pliq", "Qliq", "Pload", "Qload", "pv", "pq", "Qmin", "Qmax", "Vset",
                 "frm", "to", "r", "x", "bch", "tap_re", "tap_im"):
        assert np.array_equal(getattr(a, attr), getattr(b, attr)), attr


def test_reader_errors(pf, tmp_path):
    src = open(os.path.join(GOLDEN, "c14.cdf")).read().split("\n")
    def write(name, lines):
        p = str(tmp_path / name)
        open(p, "w").write("\n".join(lines))
        return p
    two_slack = write("a.cdf", src[:3] + [src[3][:24] + " 3" + src[3][26:]] + src[4:])
    bad_type = write("b.cdf", src[:3] + [src[3][:24] + " 7" + src[3][26:]] + src[4:])
    j = next(i for i, l in enumerate(src) if l.startswith("BRANCH")) + 1
    unknown = write("c.cdf", src[:j] + ["9999" + src[j][4:]] + src[j + 1:])
    for p in (two_slack, bad_type, unknown):
        with pytest.raises(pf.PfError) as e:
            pf.Network(p)
        assert e.value.status == pf.PF_PARSE_ERROR
    with pytest.raises(pf.PfError) as e:
        pf.Network(str(tmp_path / "missing.cdf"))
    assert e.value.status == pf.PF_IO_ERROR
    with pytest.raises(pf.PfError) as e:   # .pwf and .m are read, other extensions are not
        pf.Network(str(tmp_path / "case.pwf"))
    assert e.value.status == pf.PF_IO_ERROR
    with pytest.raises(pf.PfError) as e:
        pf.Network(write("case.raw", src))
    assert e.value.status == pf.PF_INVALID_ARG


DATA_TXT = """metodo         = {metodo};        // esparso, denso or hibridoC
openmp         = OFF;
v_inicial      = 1.02;            // flat start
theta_inicial  = 0.0;
tolerancia     = 1e-8;
no_max_iter    = 37;
arq_entrada    = ieee118.cdf;     // input file name
laconic_mode   = ON;
output_bench   = SCREEN;
output_ans     = OFF;
{extra}"""


def test_options_parse(pf):
    o = pf.options(DATA_TXT.format(metodo="hibridoC", extra="lim_inj_reat = ON;\n"), cpu_layout=0)
    assert o.metodo == pf.HIBRIDO_C and o.lim_inj_reat == 1 and o.openmp == 0
    assert o.v_inicial == 1.02 and o.tolerancia == 1e-8 and o.no_max_iter == 37
    assert o.arq_entrada == b"ieee118.cdf" and o.laconic_mode == 1 and o.output_ans == 0
    assert o.output_bench == 1
    o = pf.options(DATA_TXT.format(metodo="esparso", extra=""), cpu_layout=1)
    # CPU build prefixes the systems folder (CPU/leitura/config.cpp:108)
    assert o.arq_entrada == b"sistemas/ieee118.cdf" and o.metodo == pf.ESPARSO
    # a key missing from the file reads the value after the file's first '='
    # ("esparso"), which is neither ON nor OFF: lim_inj_reat keeps its default
    assert o.lim_inj_reat == 0
    o = pf.options(DATA_TXT.format(metodo="denso", extra="precisao = MISTA;\ngrafo = ON;\n"))
    assert o.metodo == pf.DENSO and o.precision == pf.MIXED and o.use_graph == 1
    # the GPU build's other hybrid names (GPU/sistInfo.h:103-110)
    assert pf.options(DATA_TXT.format(metodo="hibridoA", extra="")).metodo == pf.HIBRIDO_A
    assert pf.options(DATA_TXT.format(metodo="hibridoB", extra="")).metodo == pf.HIBRIDO_B


def test_options_errors(pf, tmp_path):
    with pytest.raises(pf.PfError) as e:
        pf.options(DATA_TXT.format(metodo="esparsoSimples", extra=""))
    assert e.value.status == pf.PF_INVALID_ARG
    with pytest.raises(pf.PfError) as e:
        pf.options(DATA_TXT.format(metodo="esparso", extra="").replace("= 37", "= abc"))
    assert e.value.status == pf.PF_PARSE_ERROR
    o = pf.Options()
    assert pf.lib().pf_options_load(str(tmp_path / "none.txt").encode(), 0, pf.C.byref(o)) == pf.PF_IO_ERROR


def _dense_from_slots(net, pr, blocks):
    n, sl = pr.n, net.slack
    iperm = np.empty(n, int)
    iperm[pr.perm] = np.arange(n)
    coorow = np.repeat(np.arange(net.nB), np.diff(pr.rowptr))
    A = np.zeros((2 * n, 2 * n))
    for s in range(pr.nslot):
        e = pr.slot_e[s]
        if e < 0:
            continue
        k, m = coorow[e], pr.colind[e]
        a, c = iperm[k - (k > sl)], iperm[m - (m > sl)]
        A[2 * a:2 * a + 2, 2 * c:2 * c + 2] = blocks[s].reshape(2, 2)
    return A


@pytest.mark.parametrize("which", ["ieee14", "ieee118", "c118", "c944", "c1888"])
def test_symbolic_program_solves(pf, cases, which):
    """Factor + solve over the symbolic program (nodes numbered level by level,
    the trailing dense chain as one Gauss-Jordan-inverted Schur complement)
    against numpy's dense solve of the same matrix."""
    path = os.path.join(IEEE, which + ".cdf") if which.startswith("ieee") else cases.path(which)
    net = pf.Network(path)
    pr = pf.SymbolicProbe(net)
    assert sorted(pr.perm) == list(range(pr.n))
    # every Ybus entry off the slack lands in exactly one slot
    used = pr.slot_e[pr.slot_e >= 0]
    coorow = np.repeat(np.arange(net.nB), np.diff(pr.rowptr))
    off_slack = np.sum((coorow != net.slack) & (pr.colind != net.slack))
    assert len(used) == len(set(used)) == off_slack
    rng = np.random.default_rng(7)
    blocks = np.zeros((pr.nslot, 4))
    for s in range(pr.nslot):
        if pr.slot_e[s] >= 0:
            blocks[s] = rng.standard_normal(4) + (np.array([8, 0, 0, 8]) if s < pr.n else 0)
    b = rng.standard_normal(2 * pr.n)
    x = pr.factor_solve(blocks, b)
    xr = np.linalg.solve(_dense_from_slots(net, pr, blocks), b)
    assert np.abs(x - xr).max() <= 1e-12 * np.abs(xr).max()
    # the GPU's blocked order: the subtree groups' and the narrow top's flat
    # pair programs (the top's pairs from below it applied first) and the
    # top's split sweeps give the same solution
    xf = pr.flat_factor_solve(blocks, b)
    assert np.abs(xf - xr).max() <= 1e-12 * np.abs(xr).max()
    # levels are contiguous node ranges, and the chain is the last levels
    lp = pr.lvl_ptr
    assert lp[0] == 0 and lp[-1] == pr.n and np.all(np.diff(lp) > 0)
    # subtree groups: every node below the top exactly once, on its own level,
    # and each group closed under the elimination tree below the top
    npart, nl, ptr, nodes, parent = pr.parts()
    nw = lp[nl]
    assert sorted(nodes.tolist()) == list(range(nw))
    group = np.full(pr.n, -1)
    for g in range(npart):
        for l in range(nl):
            seg = nodes[ptr[g * nl + l]:ptr[g * nl + l + 1]]
            assert np.all((seg >= lp[l]) & (seg < lp[l + 1]))
            group[seg] = g
    for j in range(nw):
        if 0 <= parent[j] < nw:
            assert group[parent[j]] == group[j]
    m, c0 = pr.chain
    if which.startswith("c"):   # the generated meshed networks end in a dense separator chain
        assert m >= 4
    if m:
        # the dense top block is the last levels (its columns need not be full:
        # the solve above checks the zero entries of S)
        assert c0 == pr.n - m and c0 in lp.tolist()
        assert np.all(pr.colptr[c0 + 1:] - pr.colptr[c0:-1] <= np.arange(m - 1, -1, -1))


def test_splitmix64_known_answers(synth):
    # published outputs of SplitMix64 (Steele, Lea & Flood 2014) for seeds 0 and 1234567
    r = synth.SplitMix64(0)
    assert [r.next_u64() for _ in range(3)] == [0xE220A8397B1DCDAF, 0x6E789E6AA1B965F4, 0x06C45D188009454F]
    r = synth.SplitMix64(1234567)
    assert [r.next_u64() for _ in range(3)] == [6457827717110365317, 3203168211198807973, 9817491932198370423]


def test_generator_workloads(pf, synth, cases, tmp_path):
    c = synth.case118()
    assert len(c.buses) == 118 and len(c.branches) == 186
    assert sum(b.type == 2 for b in c.buses) == 53 and sum(b.type == 0 for b in c.buses) == 64
    assert sum(br.tap != 0 for br in c.branches) == 9
    assert all(br.b == 0 for br in c.branches if br.tap != 0)
    assert abs(sum(b.pload for b in c.buses) / 100 - 42.0) < 0.05
    # the written file is a pure function of the seed
    p = str(tmp_path / "again.cdf")
    synth.write_cdf(synth.case118(), p)
    assert open(p).read() == open(os.path.join(GOLDEN, "c118.cdf")).read()
    n = pf.Network(cases.path("c944"))
    assert (n.nB, n.nPV, n.nPQ, n.nL) == (944, 8 * 53 + 7, 8 * 64, 8 * 186 + 16)
    c14 = synth.case14()
    assert (sum(b.type == 2 for b in c14.buses), sum(b.type == 0 for b in c14.buses), len(c14.branches)) == (4, 9, 20)
    f = synth.load_factors(118, 3000000)
    assert f.min() >= 0.7 and f.max() <= 1.3 and np.array_equal(f, synth.load_factors(118, 3000000))


# ------------------------------------------------ MATPOWER and ANAREDE readers

def _net_arrays(net):
    return {a: np.asarray(getattr(net, a)) for a in ("V", "theta", "Pliq", "Qliq", "Pload", "Qload", "Pg", "Qg",
                                                      "Vbase", "gsh", "bsh", "pv", "pq", "Qmin", "Qmax", "Vset",
                                                      "frm", "to", "r", "x", "bch", "tap_re", "tap_im")}


def test_reader_matpower_known_answers(pf):
    """tests/golden/mp5.m read by hand (CPU/leitura/m.cpp:246-536): bus rows
    in file order, the first generator row of a PV / slack bus, none for a PQ
    bus or a PV bus without one, shunts over baseMVA, every branch kept."""
    n = pf.Network(os.path.join(GOLDEN, "mp5.m"), v0=1.0, theta0=0.0)
    a = _net_arrays(n)
    assert (n.nB, n.nL, n.nPV, n.nPQ, n.slack) == (5, 6, 2, 2, 0)
    assert list(a["pv"]) == [1, 3] and list(a["pq"]) == [2, 4]
    assert np.array_equal(a["V"], [1.06, 1.045, 1.0, 1.01, 1.0]) and np.all(a["theta"] == 0)
    assert np.array_equal(a["Pload"], [0, 21.7 / 100, 47.8 / 100, 7.6 / 100, 94.2 / 100])
    assert np.array_equal(a["Qload"], [0, 12.7 / 100, -3.9 / 100, 1.6 / 100, 19 / 100])
    assert np.array_equal(a["Pg"], [232.4 / 100, 40 / 100, 0, 0, 0])      # bus 5 is PQ: its generator is not read
    assert np.array_equal(a["Qg"], [-16.9 / 100, 42.4 / 100, 0, 0, 0])
    assert np.array_equal(a["Pliq"], a["Pg"] - a["Pload"]) and np.array_equal(a["Qliq"], a["Qg"] - a["Qload"])
    assert np.array_equal(a["Qmax"], [50 / 100, 0]) and np.array_equal(a["Qmin"], [-40 / 100, 0])
    assert np.array_equal(a["Vset"], [1.045, 1.01])
    assert np.array_equal(a["gsh"], [0, 0, 0, 0, 5 / 100]) and np.array_equal(a["bsh"], [0, 0, 19 / 100, 0, 0])
    assert np.array_equal(a["Vbase"], [230, 230, 138, 138, 138])
    assert list(a["frm"]) == [0, 0, 1, 1, 2, 3] and list(a["to"]) == [1, 4, 2, 2, 3, 4]
    assert np.array_equal(a["r"], [0.01938, 0.05403, 0.04699, 0.04699, 0.06701, 0.01335])
    phi = 5 * np.pi / 180
    assert np.array_equal(a["tap_re"], [1, 1, 0.978 * np.cos(phi), 1, 1, 1])
    assert np.array_equal(a["tap_im"], [0, 0, 0.978 * np.sin(phi), 0, 0, 0])


def test_reader_matpower_ieee14(pf):
    """ieee14.m (tests/golden/make_reader_fixtures.py restates the CDF text in
    MATPOWER form) reads as the same network as ieee14.cdf."""
    a = _net_arrays(pf.Network(os.path.join(IEEE, "ieee14.cdf")))
    b = _net_arrays(pf.Network(os.path.join(IEEE, "ieee14.m")))
    for k in a:
        if k in ("gsh", "bsh"):   # pu in the CDF, MVAr at 1 pu over baseMVA in MATPOWER
            assert np.allclose(a[k], b[k], rtol=1e-15, atol=1e-17), k
        elif k != "Vbase":
            assert np.array_equal(a[k], b[k]), k


def test_reader_pwf_known_answers(pf):
    """tests/golden/pwf5.pwf read by hand (CPU/leitura/pwf.cpp:3-516): fixed
    columns, implied decimal points, defaults of blank fields, 32-bit record
    values, BASE 50, the voltage group lookup, a merged repeated branch with
    the reference's 1 / (B_raw + b) charging, and a DCSC series capacitor."""
    f32 = np.float32
    n = pf.Network(os.path.join(GOLDEN, "pwf5.pwf"), v0=1.0, theta0=0.0)
    a = _net_arrays(n)
    assert (n.nB, n.nL, n.nPV, n.nPQ, n.slack, n.n_duplicates) == (4, 5, 1, 2, 0, 1)
    assert list(a["pv"]) == [1] and list(a["pq"]) == [2, 3]
    assert np.array_equal(a["V"], [float(f32(1.06)), float(f32(1.045)), 1.0, 1.0])
    assert np.array_equal(a["Pg"], [float(f32(232.4)) / 50, float(f32(40.0)) / 50, 0, 0])
    assert np.array_equal(a["Qg"], [float(f32(-16.9)) / 50, float(f32(42.4)) / 50, 0, 0])
    assert np.array_equal(a["Qmin"], [float(f32(-40.0)) / 50]) and np.array_equal(a["Qmax"], [float(f32(50.0)) / 50])
    assert np.array_equal(a["Pload"], [0, float(f32(21.7)) / 50, float(f32(94.2)) / 50, float(f32(7.6)) / 50])
    assert np.array_equal(a["Qload"], [0, float(f32(12.7)) / 50, float(f32(19.0)) / 50, float(f32(1.6)) / 50])
    assert np.array_equal(a["Vbase"], [230, 0, 138, 138])   # group " B" is not the DGBT key "B"
    assert np.all(a["gsh"] == 0) and np.all(a["bsh"] == 0)
    assert list(a["frm"]) == [0, 0, 1, 2, 3] and list(a["to"]) == [1, 2, 2, 3, 0]

    def pct(v):   # float record value over an int 100, in float
        return float(f32(v) / f32(100))
    z1 = complex(0.0, pct(4.2))
    z2 = complex(pct(1.0), pct(4.2))
    zm = z1 * z2 / (z1 + z2)
    assert np.allclose(a["r"], [pct(1.938), pct(5.4), zm.real, pct(1.335), 0.0], rtol=1e-15, atol=0)
    assert np.allclose(a["x"], [pct(5.917), pct(22.3), zm.imag, pct(4.211), pct(-2.5)], rtol=1e-15, atol=0)
    assert np.array_equal(a["bch"], [float(f32(5.28)) / 50, float(f32(4.92)) / 50, 1.0, float(f32(1.28)) / 50, 0])
    ph = float(f32(-3.0)) * np.pi / 180
    assert np.array_equal(a["tap_re"], [1, 1, float(f32(0.978)) * np.cos(ph), 1, 1])
    assert np.array_equal(a["tap_im"], [0, 0, float(f32(0.978)) * np.sin(ph), 0, 0])
