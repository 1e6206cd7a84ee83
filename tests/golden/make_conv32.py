"""Writes conv32_oracle.json: the FP64 oracle relaxed to convergence
(|dE/E| per step < 1e-9, energies every 50 steps) on a 32^3 grid, spacing
0.75 A, N=100 He atoms, OTC with h_core=2.377 A, Morse D=30 K re=4 A a=1/A
at (0, 0, 1.5 A), seeded Fermi start."""
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

import eff_pot_amd.abi as abi  # noqa: E402
from oracle.oracle import Oracle  # noqa: E402

CFG = dict(n=32, xmax=12.0, n4=100.0, h_core=2.377, seed=99, rho0=0.0218, afermi=0.6, noise=1e-6,
           rimp=[0.0, 0.0, 1.5], tol=1e-9, pener=50, max_steps=40000)


def params():
    return abi.make_params(CFG["n"], xmax=CFG["xmax"], n4=CFG["n4"], core=abi.CORE_OTC, h_core=CFG["h_core"])


def dt_of(p):
    hx = 2 * p.xmax / p.nx
    return 0.25 / (p.h2o2m4 * 7.0729 * 3 / hx ** 2 + 1000.0 + 50.0)


def main(out):
    p = params()
    o = Oracle(p)
    o.init_fermi(CFG["rho0"], 2.22 * CFG["n4"] ** (1 / 3), CFG["afermi"], CFG["noise"], CFG["seed"])
    o.set_impurity(abi.make_morse(rimp=tuple(CFG["rimp"])))
    dt = dt_of(p)
    t0 = time.time()
    r = o.solve(abi.SolveOpts(dt=dt, tol=CFG["tol"], n_rk=3, pener=CFG["pener"], max_steps=CFG["max_steps"],
                              polish_steps=0, tol_fp32=0.0, check_finite=1))
    res = {"config": CFG, "dt": dt, "steps": r.steps, "converged": r.converged, "rel_de": r.rel_de,
           "energies": {k: getattr(r.e, k) for k, _ in abi.Energies._fields_},
           "oracle_seconds": time.time() - t0}
    with open(out, "w") as f:
        json.dump(res, f, indent=1)


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else os.path.join(os.path.dirname(__file__), "conv32_oracle.json"))
