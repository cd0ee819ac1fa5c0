"""The flumen CLI as a drop-in for the reference executables: data.txt in the
working directory, the case it names, and the reference's outputs -- the
iteration lines and PV<->PQ switch messages, the flow report (impressao2,
CPU/impressao.cpp:3-36) checked number by number against the oracle, and the
CSV of phase timings (benchmarksPrintFile, CPU/impressao.cpp:258-419)."""
import os
import re
import shutil
import subprocess

import numpy as np
import pytest

from conftest import GOLDEN, PKG_DIR

pytestmark = pytest.mark.gpu

DATA = """metodo         = {metodo};
lim_inj_reat   = ON;
openmp         = OFF;
v_inicial      = 1.0;
theta_inicial  = 0.0;
tolerancia     = 1e-8;
no_max_iter    = 100;
arq_entrada    = {case};
laconic_mode   = OFF;
verbose_mode   = OFF;
output_bench   = ALL;
output_ans     = ON;
output_iter    = ON;
streams        = OFF;
"""


def _run(tmp, case, metodo="esparso"):
    with open(os.path.join(tmp, "data.txt"), "w") as f:
        f.write(DATA.format(metodo=metodo, case=case))
    exe = os.path.join(PKG_DIR, "flumen")
    print(f"[flumen] {case} {metodo}", flush=True)
    r = subprocess.run([exe], cwd=tmp, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    return r.stdout


def _report(out):
    """Bus lines and branch lines of the flow report, by the reference's
    fixed widths ('%4d%11.5f x6' then '%4d  %11.5f  %11.5f')."""
    body = out.split("FLUXOS DE POTENCIA", 1)[1].split("Relatorio de execucao", 1)[0]
    buses, flows = [], []
    for line in body.splitlines()[2:]:
        if not line.strip():
            continue
        if line[:10].strip():
            head = line[6:]
            nums = [int(head[0:4])] + [float(head[4 + 11 * k:15 + 11 * k]) for k in range(6)]
            buses.append(nums)
            tail = line[6 + 4 + 66 + 2:]
        else:
            tail = line[78:]
        if tail.strip():
            flows.append((len(buses) - 1, int(tail[0:4]), float(tail[6:17]), float(tail[19:30])))
    return buses, flows


def test_flumen_report_and_csv(pf, oracle, tmp_path):
    tmp = str(tmp_path)
    shutil.copy(os.path.join(GOLDEN, "c14.cdf"), tmp)
    out = _run(tmp, "c14.cdf")
    assert "[ESPARSO]" in out and "Execucao concluida." in out
    oc = oracle.Case(os.path.join(tmp, "c14.cdf"))
    o = oc.solve(tol=1e-8, lim=True, reuse_symbolic=True)
    rp, ci, y = oc.ybus()
    P, Q = oracle.pq(rp, ci, y, o["V"], o["theta"])
    Pdp, Qdp, Ppd, Qpd = oc.flows(o["V"], o["theta"])
    buses, flows = _report(out)
    assert len(buses) == oc.nB
    half = 0.5e-5 + 1e-9   # printed with 5 decimals
    for i, (bid, V, ang, pli, qli, pge, qge) in enumerate(buses):
        assert abs(V - o["V"][i]) <= half and abs(ang - o["theta"][i]) <= half
        assert abs(pli - P[i]) <= half and abs(qli - Q[i]) <= half
        assert abs(pge - (P[i] + oc.Pload[i])) <= half and abs(qge - (Q[i] + oc.Qload[i])) <= half
    # every branch appears at both of its buses, in branch order
    expect = []
    for i in range(oc.nB):
        for l in range(oc.nL):
            if oc.frm[l] == i:
                expect.append((i, l, True))
            elif oc.to[l] == i:
                expect.append((i, l, False))
    assert len(flows) == len(expect)
    for (bi, other, p, q), (i, l, fwd) in zip(flows, expect):
        assert bi == i
        assert abs(p - (Pdp[l] if fwd else Ppd[l])) <= half and abs(q - (Qdp[l] if fwd else Qpd[l])) <= half
    # iteration lines: one per NR pass
    lines = re.findall(r"max\(gLim\) = (\S+); sum\(gLim\) = (\S+)", out)
    assert len(lines) == o["iterations"] + 1
    assert float(lines[-1][0]) <= 1e-8
    # the CSV: a two-line header, then one numbered row per run
    _run(tmp, "c14.cdf")
    csv = open(os.path.join(tmp, "out_double_esparso-c14.csv")).read().splitlines()
    assert csv[0].startswith("Execucao;ms total;no iteracoes;ms Ybus;ms calcPQ;;ms Jacobiano;;")
    assert csv[1].startswith(";;;; total; medio por iteracao;")
    assert len(csv) == 4
    for k, row in enumerate(csv[2:], start=1):
        f = row.split(";")
        assert f[0] == str(k) and f[2] == str(o["iterations"]) and f[-1] == ""
        assert len(f) == 13
        assert all(re.fullmatch(r"-?\d+,\d+", x) for x in f[1:2] + f[3:12])


@pytest.mark.parametrize("metodo", ["esparso", "denso", "hibridoC"])
def test_flumen_switch_messages(pf, tmp_path, metodo):
    """tests/golden/c3_limits.cdf switches its PV bus to PQ at Qmin on pass 1
    and back on pass 2 (traced in test_oracle.py): flumen prints the
    reference's messages before the iteration lines of those passes."""
    tmp = str(tmp_path)
    shutil.copy(os.path.join(GOLDEN, "c3_limits.cdf"), tmp)
    out = _run(tmp, "c3_limits.cdf", metodo)
    body = [l for l in out.splitlines() if l.startswith("max(gLim)") or l.startswith("Barra ")]
    kinds = [("msg" if l.startswith("Barra") else "max") for l in body]
    assert kinds == ["max", "msg", "max", "msg", "max", "max", "max"]
    assert body[1] == "Barra 2 [0]: PV -> PQ (limite inferior)"
    assert body[3] == "Barra 2 [0]: PQ (limite inferior) -> PV"
    assert os.path.exists(os.path.join(tmp, f"out_double_{metodo}-c3_limits.csv"))
