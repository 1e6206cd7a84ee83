"""Writes plus_potentials.json: the reference's Ba+ - He fits V_Ba_plus_gs,
_2D0, _2D1 and _2D2 (V_impur.f90:1218-1396) evaluated exactly at a list of
radii, for the tests of the X+ potential of the uplus energy term.

Run in a container that holds the reference sources:
    python tests/golden/make_plus_vectors.py [/root/reference]
It reads each fit's Block Data (REAL*16 coefficients Pg, REAL*8 rcutoff,
fcutoff, r0, aa, bb, cc, integer k0) and evaluates, in exact rational
arithmetic (fractions.Fraction, no rounding until the final double),
    V = fcutoff                          r <= rcutoff
        |aa r^2 + bb r + cc|             r <= r0
        sum_k Pg(k) r^-(k+k0), k=1..13   otherwise,
the branches of the V_Ba_plus_* functions.  The file holds only radii and
values; Select_Pot's own r < r_cutoff -> umax rule is not part of it.
"""
import json
import os
import re
import sys
from fractions import Fraction

NAMES = ["Ba_plus_gs", "Ba_plus_2D0", "Ba_plus_2D1", "Ba_plus_2D2"]
RADII = [0.25, 0.9, 1.0, 1.0000000000000002, 1.05, 1.5, 1.75, 1.9999999999999998, 2.0, 2.0000000000000004,
         2.01, 2.1, 2.25, 2.5, 2.75, 3.0, 3.25, 3.5, 4.0, 4.5, 5.0, 6.0, 7.0, 8.0, 10.0, 12.5, 15.0, 20.0,
         30.0, 50.0]


def fits(src):
    out = {}
    for name in NAMES:
        m = re.search(r"Block Data Inicio_LJ_V_%s\n(.*?)\nEnd" % name, src, re.S)
        block = m.group(1)
        pg = [Fraction(x[:-2]) for x in re.findall(r"[-+]?\d+\.\d+Q0", block)]
        assert len(pg) == 13, name
        par = {k: Fraction(float(re.search(r"Data %s/\s*([-+0-9.D]+)/" % k, block).group(1).replace("D", "e")))
               for k in ("rcutoff", "fcutoff", "r0", "aa", "bb", "cc")}
        par["k0"] = int(re.search(r"Data k0/\s*(\d+)/", block).group(1))
        out[name] = (pg, par)
    return out


def v_exact(pg, par, r):
    x = Fraction(r)
    if x <= par["rcutoff"]:
        return float(par["fcutoff"])
    if x <= par["r0"]:
        return float(abs(par["aa"] * x * x + par["bb"] * x + par["cc"]))
    return float(sum(c / x ** (k + 1 + par["k0"]) for k, c in enumerate(pg)))


def main():
    ref = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
    with open(os.path.join(ref, "src", "V_impur.f90"), encoding="latin-1") as f:
        src = f.read()
    data = {"radii": RADII}
    for name, (pg, par) in fits(src).items():
        data[name] = [v_exact(pg, par, r) for r in RADII]
    out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "plus_potentials.json")
    with open(out, "w") as f:
        json.dump(data, f, indent=1)
    print("wrote", out)


if __name__ == "__main__":
    main()
