"""Effective-potential scan: E_imp(z0) for many impurity positions at once.

All positions are slots of one batched context (one launch per kernel for the
whole scan).  Optional warm start: the pure droplet (no impurity) is relaxed
once and every slot starts from it instead of the Fermi profile.  Positions
whose impurity lies outside the box are computed (the Morse field there is
only its tail inside the grid) and flagged.
"""
import time

import numpy as np

from . import Context, SolveOpts, make_params, make_morse, default_dt, CORE_OTC, PREC_MIXED, POT_NULL
from . import SLOT_START, SLOT_DONE, SLOT_FAILED


def scan_positions(natoms, npos, zmax_frac=1.6):
    R = 2.22 * natoms ** (1.0 / 3.0)
    return [zmax_frac * R * k / (npos - 1) for k in range(npos)]


def run_scan(n=96, spacing=0.6, natoms=1000.0, npos=48, precision=PREC_MIXED, tol=1e-9, tol_fp32=1e-7,
             max_steps=60000, warm=False, seed=20261203, h_core=2.377, progress=None, graph_steps=24):
    """Relax all positions in one batched context (ticked until every slot is
    done; up to `graph_steps` predictor-corrector steps per graph launch
    between energy checks); `progress(msg)` is called about every 30 s."""
    xmax = spacing * n / 2
    p = make_params(n, xmax=xmax, n4=natoms, core=CORE_OTC, h_core=h_core, precision=precision, batch=npos,
                    psi_complex=0)
    zs = scan_positions(natoms, npos)
    R = 2.22 * natoms ** (1.0 / 3.0)
    dt = default_dt(p)
    opts = SolveOpts(dt=dt, tol=tol, n_rk=3, pener=50, max_steps=max_steps, polish_steps=max_steps,
                     tol_fp32=tol_fp32, check_finite=0, graph_steps=graph_steps)
    warm_psi = None
    if warm:
        pw = make_params(n, xmax=xmax, n4=natoms, core=CORE_OTC, h_core=h_core, precision=precision, batch=1,
                         psi_complex=0)
        with Context(pw) as c:
            c.init_fermi(0.0218, R, 0.6, 1e-6, seed)
            imp = make_morse(D=0.0)
            imp.kind = POT_NULL
            c.set_impurity(imp)
            c.solve(opts)
            warm_psi = c.psi()
    ctx = Context(p)
    for b, z in enumerate(zs):
        if warm_psi is not None:
            ctx.upload_psi(warm_psi, slot=b)
        else:
            ctx.init_fermi(0.0218, R, 0.6, 1e-6, seed, slot=b)
        ctx.set_impurity(make_morse(rimp=(0.0, 0.0, z)), slot=b)
    slots = ctx.new_slots()
    for b in range(npos):
        slots[b].phase = SLOT_START
        slots[b].fp32 = 1 if precision == PREC_MIXED else 0
    t0 = time.perf_counter()
    t_report = t0
    while any(slots[b].phase not in (SLOT_DONE, SLOT_FAILED) for b in range(npos)):
        ctx.tick(opts, slots)
        if progress and slots[0].since == 0 and time.perf_counter() - t_report > 30.0:
            t_report = time.perf_counter()
            done = sum(slots[b].phase in (SLOT_DONE, SLOT_FAILED) for b in range(npos))
            progress(f"{t_report - t0:.0f} s: {done}/{npos} positions converged, "
                     f"steps {min(slots[b].steps for b in range(npos))}..{max(slots[b].steps for b in range(npos))}")
    res = [slots[b] for b in range(npos)]
    ctx.close()
    out = []
    for z, r in zip(zs, res):
        out.append(dict(z0=z, inside=abs(z) <= xmax, eimp=r.e.eimpu, etot=r.e.etot, steps=r.steps,
                        converged=bool(r.converged)))
    return out, dict(n=n, spacing=spacing, natoms=natoms, dt=dt, warm=warm)


def veff_curve(results):
    return np.array([[r["z0"], r["eimp"]] for r in results])
