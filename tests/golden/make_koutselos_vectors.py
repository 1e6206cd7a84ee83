"""Writes plus_koutselos.json: the reference's Koutselos alkali-ion - He
potentials V_Koutselos_{Na,K,Rb,Cs}_plus_gs (V_impur.f90:1704-1883, selected
as '<Ion>_plus_Koutselos' at :232-292; prep.sh:5-7 selects the Rb+ one)
evaluated in 50-digit arithmetic at a list of radii, for the tests of the X+
potential of the uplus energy term.

Run in a container that holds the reference sources:
    python tests/golden/make_koutselos_vectors.py [/root/reference]
It reads each function's parameters from its declarations (v_0, rho, R_m,
C4, C6, C8, A, a_a, B, b_b, a_d, a_q, a_o, a_bohr, Hartree, eV_K, all REAL*8
decimals) and evaluates with mpmath, x = r / a_bohr,
    V = [v_0 (A e^{-a_a x/rho} - B e^{-b_b x/rho})
         - (CC4/x^4 + CC6/x^6 + CC8/x^8) h(x)] Hartree eV_K,
    CC_n = C_n + a_{d,q,o}/2,  h = 1 (x >= 1.28 R_m), else exp(-(1.28 R_m/x - 1)^2).
For each radius it also stores the sum of the magnitudes of the terms (the
two exchange exponentials and the dispersion, in K), the scale on which an
FP64 evaluation is compared where they cancel.  Select_Pot's own
r < r_cutoff -> umax rule is not part of the file.
"""
import json
import os
import re
import sys

import mpmath as mp

IONS = ["Na", "K", "Rb", "Cs"]
KEYS = ["v_0", "rho", "R_m", "C4", "C6", "C8", "A", "a_a", "B", "b_b", "a_d", "a_q", "a_o", "a_bohr", "Hartree",
        "eV_K"]
RADII = [0.5, 1.0, 1.5, 2.0, 2.5, 2.75, 3.0, 3.25, 3.5, 3.75, 4.0, 4.25, 4.5, 5.0, 5.5, 6.0, 7.0, 8.0, 10.0, 12.5,
         15.0, 20.0, 30.0]


def params(src, ion):
    m = re.search(r"Double Precision Function V_Koutselos_%s_plus_gs\(xx\)\n(.*?)End Function" % ion, src, re.S)
    body = m.group(1)
    par = {}
    for k in KEYS:
        v = re.search(r"\b%s\s*=\s*([-+0-9.]+)d0" % k, body)
        par[k] = mp.mpf(v.group(1))
    return par


def v_exact(par, r):
    x = mp.mpf(r) / par["a_bohr"]
    rs = x / par["rho"]
    ea, eb = par["A"] * mp.exp(-par["a_a"] * rs), par["B"] * mp.exp(-par["b_b"] * rs)
    vex = par["v_0"] * (ea - eb)
    aux = par["R_m"] * mp.mpf("1.28")
    h = mp.mpf(1) if x >= aux else mp.exp(-(aux / x - 1) ** 2)
    cc4 = par["C4"] + par["a_d"] / 2
    cc6 = par["C6"] + par["a_q"] / 2
    cc8 = par["C8"] + par["a_o"] / 2
    disp = -(cc4 / x ** 4 + cc6 / x ** 6 + cc8 / x ** 8) * h
    eh = par["Hartree"] * par["eV_K"]
    return float((vex + disp) * eh), float((par["v_0"] * (ea + eb) + abs(disp)) * eh)


def main():
    mp.mp.dps = 50
    ref = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
    with open(os.path.join(ref, "src", "V_impur.f90"), encoding="latin-1") as f:
        src = f.read()
    data = {"radii": RADII}
    for ion in IONS:
        par = params(src, ion)
        # both sides of the switch of the damping function, x = 1.28 R_m
        rs = float(par["R_m"] * mp.mpf("1.28") * par["a_bohr"])
        radii = RADII + [rs * (1 - 1e-9), rs * (1 + 1e-9)]
        vals = [v_exact(par, r) for r in radii]
        data[f"{ion}_plus_Koutselos"] = {"radii": radii, "V": [v for v, _ in vals], "scale": [s for _, s in vals]}
    out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "plus_koutselos.json")
    with open(out, "w") as f:
        json.dump(data, f, indent=1)
    print("wrote", out)


if __name__ == "__main__":
    main()
