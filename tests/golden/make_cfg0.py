"""Writes cfg0_oracle.json: the FP64 oracle on the small droplet workload
(N=300 He atoms, 64^3 grid, spacing 0.75 A, OTC with h_core=2.377 A, Morse
D=30 K re=4 A a=1/A at the droplet centre, seeded Fermi start), 3 RK-Gill
+ 1997 predictor-corrector imaginary-time steps with energies every 100 steps.
The energies are the fixture the GPU path is checked against."""
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

import eff_pot_amd.abi as abi  # noqa: E402
from oracle.oracle import Oracle  # noqa: E402

CFG = dict(n=64, xmax=24.0, n4=300.0, h_core=2.377, seed=20261203, rho0=0.0218, afermi=0.6, noise=1e-6,
           steps=2000, pener=100)


def params():
    return abi.make_params(CFG["n"], xmax=CFG["xmax"], n4=CFG["n4"], core=abi.CORE_OTC, h_core=CFG["h_core"])


def dt_of(p):
    hx = 2 * p.xmax / p.nx
    lam = p.h2o2m4 * 7.0729 * 3 / hx ** 2 + 1000.0 + 50.0
    return 0.25 / lam


def main(out):
    p = params()
    o = Oracle(p)
    R = 2.22 * CFG["n4"] ** (1.0 / 3.0)
    o.init_fermi(CFG["rho0"], R, CFG["afermi"], CFG["noise"], CFG["seed"])
    o.set_impurity(abi.make_morse())
    dt = dt_of(p)
    t0 = time.time()
    r = o.solve(abi.SolveOpts(dt=dt, tol=0.0, n_rk=3, pener=CFG["pener"], max_steps=CFG["steps"],
                              polish_steps=0, tol_fp32=0.0, check_finite=1))
    res = {"config": CFG, "dt": dt, "steps": r.steps, "rel_de": r.rel_de,
           "energies": {k: getattr(r.e, k) for k, _ in abi.Energies._fields_},
           "oracle_seconds": time.time() - t0}
    with open(out, "w") as f:
        json.dump(res, f, indent=1)


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else os.path.join(os.path.dirname(__file__), "cfg0_oracle.json"))
