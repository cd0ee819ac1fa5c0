"""Writes the hand-made MATPOWER (.m) and ANAREDE (.pwf) fixtures of the
reader tests, and the MATPOWER restatement of the bundled IEEE 14-bus case.

    python tests/golden/make_reader_fixtures.py

mp5.m: five buses in file order 1, 2, 7, 4, 5 (bus 7 breaks the 1..n
numbering), tab-separated as MATPOWER writes its cases, with comments, a
trailing comment after a row, "Inf", a PV bus with two generator rows (the
first one counts), a PV bus without one (no limits, no generation), a PQ bus
with one (its generation is not read), a phase-shifting transformer and a
repeated branch (not merged).

pwf5.pwf: the ANAREDE records of a four-bus case in fixed columns: implied
decimal points, blank fields taking their defaults, a comment line, a bus
of type 3, voltage groups, a repeated branch (merged), a series capacitor
(DCSC) and BASE 50 in DCTE.

ieee14.m: tests/golden/ieee/ieee14.cdf restated in MATPOWER form (bus
shunts in MW / MVAr at 1 pu, generation on the generator rows of its PV and
slack buses) by this script's own column reading of the CDF text.
"""
from __future__ import annotations

import os

HERE = os.path.dirname(os.path.abspath(__file__))

MP5 = [
    "function mpc = mp5",
    "%% five buses for the reader tests",
    "mpc.version = '2';",
    "mpc.baseMVA = 100;",
    "",
    "%% bus data",
    "%\tbus_i\ttype\tPd\tQd\tGs\tBs\tarea\tVm\tVa\tbaseKV\tzone\tVmax\tVmin",
    "mpc.bus = [",
    "\t1\t3\t0\t0\t0\t0\t1\t1.06\t0\t230\t1\t1.1\t0.9;",
    "\t2\t2\t21.7\t12.7\t0\t0\t1\t1.045\t-4.98\t230\t1\t1.1\t0.9;",
    "\t7\t1\t47.8\t-3.9\t0\t19\t1\t1\t-10\t138\t1\t1.1\t0.9;\t% a comment after a row",
    "\t4\t2\t7.6\t1.6\t0\t0\t1\t1.01\t0\t138\t1\t1.1\t0.9;",
    "\t5\t1\t94.2\t19\t5\t0\t1\t1\t0\t138\t1\tInf\t0.9;",
    "];",
    "",
    "%% generator data",
    "mpc.gen = [",
    "\t1\t232.4\t-16.9\t10\t0\t1.06\t100\t1\t332.4\t0\t0\t0\t0\t0\t0\t0\t0\t0\t0\t0\t0;",
    "\t2\t40\t42.4\t50\t-40\t1.045\t100\t1\t140\t0\t0\t0\t0\t0\t0\t0\t0\t0\t0\t0\t0;",
    "\t2\t10\t5\t99\t-99\t1.045\t100\t1\t140\t0\t0\t0\t0\t0\t0\t0\t0\t0\t0\t0\t0;",
    "\t5\t30\t10\t20\t-20\t1\t100\t1\t100\t0\t0\t0\t0\t0\t0\t0\t0\t0\t0\t0\t0;",
    "];",
    "",
    "%% branch data",
    "mpc.branch = [",
    "\t1\t2\t0.01938\t0.05917\t0.0528\t0\t0\t0\t0\t0\t1\t-360\t360;",
    "\t1\t5\t0.05403\t0.22304\t0.0492\t0\t0\t0\t0\t0\t1\t-360\t360;",
    "\t2\t7\t0.04699\t0.19797\t0.0438\t0\t0\t0\t0.978\t5\t1\t-360\t360;",
    "\t2\t7\t0.04699\t0.19797\t0.0438\t0\t0\t0\t0\t0\t0\t-360\t360;",
    "\t7\t4\t0.06701\t0.17103\t0.0346\t0\t0\t0\t0\t0\t1\t-360\t360;",
    "\t4\t5\t0.01335\t0.04211\t0.0128\t0\t0\t0\t0\t0\t1\t-360\t360;",
    "];",
]


def put(line: str, col: int, text: str) -> str:
    """Writes text at 0-based column col of a fixed-column record."""
    if len(line) < col:
        line += " " * (col - len(line))
    return line[:col] + text + line[col + len(text):]


def dbar(num, tipo, group, v, ang, pg, qg, qmin, qmax, pl, ql):
    l = put("", 0, f"{num:5d}")
    l = put(l, 7, tipo)
    l = put(l, 8, group)
    l = put(l, 10, f"BUS{num:<8d}")
    for c, t in ((24, v), (28, ang), (32, pg), (37, qg), (42, qmin), (47, qmax), (58, pl), (63, ql)):
        l = put(l, c, t)
    return l.rstrip()


def dlin(a, b, r, x, bs, tap, ph):
    l = put("", 0, f"{a:5d}")
    l = put(l, 10, f"{b:5d}")
    for c, t in ((20, r), (26, x), (32, bs), (38, tap), (53, ph)):
        l = put(l, c, t)
    return l.rstrip()


def pwf5():
    out = ["TITU", " four buses for the reader tests", "DCTE", "(Mn) (Val) (Mn) (Val)", "TEPA   .1 BASE  50.0",
           "99999", "DBAR", "(Num)OETGb(   nome   )Gl( V)( A)( Pg)( Qg)( Qn)( Qm)(Bc  )( Pl)( Ql)"]
    out += [dbar(10, "2", "AB", "1060", "   0", "232.4", "-16.9", "", "", "", ""),
            dbar(20, "1", " B", "1045", "  -5", "  40.", " 42.4", "-40.0", " 50.0", " 21.7", " 12.7"),
            "(a comment inside the section)",
            dbar(30, " ", "  ", "    ", "    ", "", "", "", "", " 94.2", " 19.0"),
            dbar(40, "3", "  ", " 990", "    ", "", "", "", "", "  7.6", "  1.6"),
            "99999", "DLIN", "(De )d O d(Pa )NcEP ( R% )( X% )(Mvar)(Tap)(Tmn)(Tmx)(Phs)"]
    out += [dlin(10, 20, " 1.938", " 5.917", "  5.28", "", ""),
            dlin(10, 30, "   540", "  2230", "  4920", "", ""),
            dlin(20, 30, "      ", "   4.2", "      ", "  978", "  -3."),
            dlin(20, 30, "   1.0", "   4.2", "   1.0", "", ""),
            dlin(30, 40, " 1.335", " 4.211", "  1.28", "", ""),
            "99999", "DCSC"]
    l = put("", 0, f"{40:5d}")
    l = put(l, 9, f"{10:5d}")
    l = put(l, 45, "  -2.5")
    out += [l, "99999", "DGBT", "(G ( kV)", "AB 230.0", " 0 138.0", "99999", "FIM"]
    return out


def cdf_to_matpower(path: str) -> list:
    """ieee14.cdf in MATPOWER form: bus rows from the CDF bus records
    (IEEE CDF columns: number 1-4, type 25-26, final V 28-33, angle 34-40,
    load MW/MVAr 41-49/50-59, generation 60-67/68-75, base kV 77-83,
    Qmax/Qmin 91-97/99-105, G/B shunt 107-114/115-122), generator rows for
    the PV and slack buses, branch rows from the branch records (from 1-4,
    to 6-9, R 20-29, X 30-40, B 41-50, tap 77-82, angle 84-90)."""
    lines = open(path, newline="").read().replace("\r", "").split("\n")
    base = float(lines[0][31:37])
    i = 2
    bus = []
    while "-999" not in lines[i]:
        bus.append(lines[i])
        i += 1
    i += 2
    br = []
    while "-999" not in lines[i]:
        br.append(lines[i])
        i += 1
    tmap = {0: 1, 1: 1, 2: 2, 3: 3}
    out = ["function mpc = ieee14", "mpc.version = '2';", f"mpc.baseMVA = {base:g};", "mpc.bus = ["]
    gens = []
    for l in bus:
        num, typ = int(l[0:4]), int(l[24:26])
        v, ang = float(l[27:33]), float(l[33:40])
        pd, qd, pg, qg = float(l[40:49]), float(l[49:59]), float(l[59:67]), float(l[67:75])
        kv, qmax, qmin = float(l[76:83]), float(l[90:97]), float(l[98:105])
        gs, bs = float(l[106:114]), float(l[114:122])
        vals = [num, tmap[typ], pd, qd, repr(gs * base), repr(bs * base), 1, v, ang, kv, 1, 1.06, 0.94]
        out.append("\t" + "\t".join(str(x) for x in vals) + ";")
        if typ in (2, 3):
            gens.append([num, pg, qg, qmax, qmin, v, base, 1, 999, 0] + [0] * 11)
    out += ["];", "mpc.gen = ["] + ["\t" + "\t".join(str(x) for x in g) + ";" for g in gens] + ["];", "mpc.branch = ["]
    for l in br:
        vals = [int(l[0:4]), int(l[4:9]), l[19:29].strip(), l[29:40].strip(), l[40:50].strip(), 0, 0, 0,
                l[76:82].strip(), l[83:90].strip(), 1, -360, 360]
        out.append("\t" + "\t".join(str(x) for x in vals) + ";")
    out.append("];")
    return out


def main() -> None:
    with open(os.path.join(HERE, "mp5.m"), "w") as f:
        f.write("\n".join(MP5) + "\n")
    with open(os.path.join(HERE, "pwf5.pwf"), "w") as f:
        f.write("\n".join(pwf5()) + "\n")
    with open(os.path.join(HERE, "ieee", "ieee14.m"), "w") as f:
        f.write("\n".join(cdf_to_matpower(os.path.join(HERE, "ieee", "ieee14.cdf"))) + "\n")


if __name__ == "__main__":
    main()
