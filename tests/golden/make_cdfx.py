"""Re-columns bundled IEEE CDF files into the CDFX layout, independently of
the package's case generator.

CDFX (CPU/leitura/cdf.cpp:329-493) widens the bus number to 5 characters:
every bus-record field moves one column right (id (0,4) -> (0,5), type
(24,2) -> (25,2), ...), and in branch records the from-bus field widens by
one and the to-bus field by one more (from (0,4) -> (0,5), to (4,5) ->
(6,5), R (19,10) -> (21,10), ...).  So a bus record gets one blank in
front, a branch record a blank in front and another after the from field.
Header, section and terminator lines are copied unchanged.

    python tests/golden/make_cdfx.py    # writes tests/golden/ieee/*.cdfx
"""
from __future__ import annotations

import os

HERE = os.path.dirname(os.path.abspath(__file__))
IEEE = os.path.join(HERE, "ieee")


def recolumn(text: str) -> str:
    lines = text.split("\n")
    out = lines[:2]                       # title card and the bus section header
    i = 2
    while "-999" not in lines[i]:         # bus records
        out.append(" " + lines[i])
        i += 1
    out += lines[i:i + 2]                 # terminator and the branch section header
    i += 2
    while "-999" not in lines[i]:         # branch records
        l = lines[i]
        out.append(" " + l[:4] + " " + l[4:])
        i += 1
    out += lines[i:]
    return "\n".join(out)


def main() -> None:
    for name in ("ieee14", "ieee118"):
        src = os.path.join(IEEE, name + ".cdf")
        with open(src, newline="") as f:
            text = f.read()
        with open(os.path.join(IEEE, name + ".cdfx"), "w", newline="") as f:
            f.write(recolumn(text))


if __name__ == "__main__":
    main()
