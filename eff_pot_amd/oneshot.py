"""The eff_pot one-shot workflow as one batched GPU job.

The reference computes the impurity energy of a stored density one file per
process: loop.sh:9-13 and get_impurity_energy.sh:7-16 run the program once per
snapshot, which reads it (readenc mode 3, readenc.f90:151-168), builds the
impurity field at the snapshot's rimp (potenimp), calls poten and energy
(DFT4He3d.f90:437-439) and prints the energy balance, U* (eHeX), U+ (uplus
with Select_Pot(selec_plus)) and E_SO among it.  Here the snapshots are
read with effpot_snapshot_read and evaluated `batch` at a time as the slots of
one context: one potenimp per slot, one poten and one energy launch sequence
for the whole batch.  Snapshots on different grids go to different contexts.

    python -m eff_pot_amd.oneshot --morse 30 4 1 den-*.dat

prints, per file, the reference's "FIRST ENERGY BALANCE", impurity-energy
labels and X-He CoM distance (formats 6050 / 6060, DFT4He3d.f90:460-463), so
get_impurity_energy.sh's parsing applies unchanged, followed by a one-line
summary.  The X+ potential defaults to the reference's selec_plus
('Ba_plus_2D2', r_cutoff_plus 1.05, umax_plus 6.260015e4).
"""
import argparse
import sys
from collections import OrderedDict

import math

from . import (Context, make_params, make_morse, make_plus, make_state_imp, read_snapshot, StateImp, STATE_P,
               STATE_D, CORE_OTC, CORE_OT, PREC_FP64, ICON_ZERO, H4OT, PLUS_DEFAULT, make_select, exciplex_moments,
               SELECT_PAIR, SELECT_JZ, SELECT_FIX, SELECT_AVERAGE, SELECT_MIXTURE)

SELECT_MODES = {"pair": SELECT_PAIR, "jz": SELECT_JZ, "fix": SELECT_FIX, "average": SELECT_AVERAGE,
                "mixture": SELECT_MIXTURE}


def _grid_key(h):
    return (h.nx, h.ny, h.nz, round(h.xmax, 12), round(h.ymax, 12), round(h.zmax, 12))


def evaluate(paths, impurity, core=CORE_OTC, h_core=2.377, icon=ICON_ZERO, batch=8, rimp=None, plus="default",
             state=None, exclusion=0.0):
    """E_imp and the energy balance of each snapshot in `paths`.

    impurity: an effpot Impurity (its rimp is replaced by the snapshot's own
    rimp unless `rimp` is given).  plus: the X+ potential of uplus (an effpot
    Plus, None for none, "default" for the reference's selec_plus).  state: a
    StateImp template for a P/D internal-state impurity instead of `impurity`
    (instates runs on each snapshot's density, as readenc does for mode 3
    with instate >= 0; its invar, eigenvalues and eso are returned).
    exclusion: z_exciplex_exclusion, the radius about the impurity whose
    points are left out of eHeX and uplus (energy.f90:155).  Returns
    one dict per path, in order, with the energies (effpot_energies fields),
    the impurity position used and the X-He centre-of-mass distance.
    """
    if isinstance(plus, str) and plus == "default":
        plus = make_plus(*PLUS_DEFAULT)
    headers = [read_snapshot(p, header_only=True)[0] for p in paths]
    groups = OrderedDict()
    for i, h in enumerate(headers):
        groups.setdefault(_grid_key(h), []).append(i)
    out = [None] * len(paths)
    extra = {}
    for key, idx in groups.items():
        h0 = headers[idx[0]]
        B = max(1, min(batch, len(idx)))
        p = make_params(h0.nx, xmax=h0.xmax, ymax=h0.ymax, zmax=h0.zmax, ny=h0.ny, nz=h0.nz, n4=1.0,
                        core=core, h_core=h_core, icon=icon, precision=PREC_FP64, batch=B, psi_complex=1)
        with Context(p) as ctx:
            for start in range(0, len(idx), B):
                chunk = idx[start:start + B]
                for s in range(B):
                    ctx.set_active(s, s < len(chunk))
                for s, i in enumerate(chunk):
                    h, psi = read_snapshot(paths[i])
                    ctx.upload_psi(psi, slot=s)
                    pos = tuple(rimp) if rimp is not None else tuple(h.rimp)
                    if state is not None:
                        st = StateImp.from_buffer_copy(state)
                        st.rimp[:] = pos
                        extra[i] = ctx.set_impurity_state(st, slot=s)
                    else:
                        imp = make_morse(rimp=pos, D=impurity.D, re=impurity.re, a=impurity.a,
                                         r_cutoff=impurity.r_cutoff, umax=impurity.umax, cap=impurity.cap)
                        imp.kind = impurity.kind
                        ctx.set_impurity(imp, slot=s)
                    ctx.set_plus(plus, slot=s)
                    ctx.set_exclusion(exclusion, slot=s)
                ctx.poten()
                es = ctx.energy()
                for s, i in enumerate(chunk):
                    d = dict(es[s])
                    d["path"] = paths[i]
                    d["rimp"] = tuple(rimp) if rimp is not None else tuple(headers[i].rimp)
                    # r_cm divides by the atom number (the context's n4 is 1 here)
                    for c in ("xcm", "ycm", "zcm"):
                        d[c] /= d["natoms"]
                    if i in extra:
                        n = 6 if extra[i].state == STATE_P else 10
                        d["invar"] = extra[i].invar_c()
                        d["eigen"] = list(extra[i].eigen)[:n]
                    d["dist"] = math.sqrt((d["rimp"][0] - d["xcm"]) ** 2 + (d["rimp"][1] - d["ycm"]) ** 2 +
                                          (d["rimp"][2] - d["zcm"]) ** 2)
                    out[i] = d
    return out


def print_balance(d, f=sys.stdout):
    """The driver's stdout labels (formats 6050 and 6060) for one result."""
    f.write(f"Number of He4 atoms{d['natoms']:15.6E}\n")
    f.write("\n\n    FIRST ENERGY BALANCE: \n\n")
    f.write(f"    Number of particles ..........: {d['natoms']:18.6f}  \n")
    f.write(f"    TOTAL   energy (He) ..........: {d['etot4']:18.6f} K\n")
    f.write(f"    Energy per particle (He) .....: {d['etot4'] / d['natoms']:18.6f} K\n")
    f.write(f"    Kinetic energy (He) ..........: {d['ekin4']:18.6f} K\n")
    f.write(f"    Lennard-Jones energy (He) ....: {d['elj4']:18.6f} K\n")
    f.write(f"    Alpha_s term  energy (He) ....: {d['ealphas']:18.6f} K\n")
    f.write(f"    Solid energy (He)  ...........: {d['esolid']:18.6f} K\n")
    f.write(f"    Correlation energy   (He) ....: {d['ecor4']:18.6f} K\n")
    f.write(f"    Impurity energy (X) ..........: {d['eimpu']:18.6f} K\n")
    f.write(f"    Kinetic energy (X) ...........: {d['ekinx']:18.6f} K\n")
    f.write(f"    Interaction energy (X*-He) ...: {d['eHeX']:18.6f} K\n")
    f.write(f"    Interaction energy (X+-He) ...: {d['uplus']:18.6f} K\n")
    f.write(f"    Spin-Orbit energy (X) ........: {d['eso']:18.6f} K\n")
    f.write(f"    TOTAL ENERGY (He+X) ..........: {d['etot']:18.6f} K\n\n")
    f.write(f" X-HE CoM distance:    {d['dist']:.15g}     \n")


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("files", nargs="+")
    ap.add_argument("--morse", type=float, nargs=3, metavar=("D", "RE", "A"), default=(30.0, 4.0, 1.0))
    ap.add_argument("--rimp", type=float, nargs=3, default=None, help="impurity position (default: each file's)")
    ap.add_argument("--selec-plus", default=PLUS_DEFAULT[0], help="X+-He potential of uplus (Select_Pot name)")
    ap.add_argument("--r-cutoff-plus", type=float, default=PLUS_DEFAULT[1])
    ap.add_argument("--umax-plus", type=float, default=PLUS_DEFAULT[2])
    ap.add_argument("--lstate", choices=["P", "D"], default=None,
                    help="P/D internal-state impurity (Ba+ fits of selec_pi/sigma/delta) instead of --morse")
    ap.add_argument("--instate", type=int, default=0)
    ap.add_argument("--als", type=float, default=0.0, help="spin-orbit splitting Als_P / Als_D (K)")
    ap.add_argument("--select", choices=list(SELECT_MODES), default="pair",
                    help="instates' choice of invar: pair (default), jz (Ldiag_jz), fix (Lexcite_state_fix), "
                         "average (Laverage_P_value), mixture (Lmixture)")
    ap.add_argument("--ljz", default="", help="Ljz of --select jz ('-3/2', '+1/2', ...; '' = default)")
    ap.add_argument("--a-pi-3o2", type=float, default=0.94)
    ap.add_argument("--a-sig-1o2", type=float, default=0.34117444)
    ap.add_argument("--exciplex", choices=["Ring", "Linear"], default=None,
                    help="Lexciplex_state_fix: P moments of a fixed exciplex instead of the density")
    ap.add_argument("--exciplex-axis", choices=["X", "Y", "Z"], default="Z")
    ap.add_argument("--r-exc", type=float, default=3.9)
    ap.add_argument("--z-exciplex-exclusion", type=float, default=0.0,
                    help="radius about the impurity left out of eHeX and uplus (energy.f90:155)")
    ap.add_argument("--core", choices=["OT", "OTC"], default="OTC")
    ap.add_argument("--h-core", type=float, default=2.377)
    ap.add_argument("--batch", type=int, default=8)
    ap.add_argument("--quiet", action="store_true", help="only the summary lines")
    a = ap.parse_args(argv)
    imp = make_morse(D=a.morse[0], re=a.morse[1], a=a.morse[2])
    plus = make_plus(a.selec_plus, a.r_cutoff_plus, a.umax_plus)
    state = None
    if a.lstate:
        sel = make_select(SELECT_MODES[a.select], a.ljz, a.a_pi_3o2, a.a_sig_1o2, fixed_moments=bool(a.exciplex))
        state = make_state_imp(STATE_P if a.lstate == "P" else STATE_D, instate=a.instate, als=a.als, select=sel)
        if a.exciplex:
            if a.lstate != "P":
                ap.error("--exciplex needs --lstate P")
            m = exciplex_moments(state.v_pi, state.v_sigma, a.exciplex, a.exciplex_axis, a.r_exc)
            for i in range(7):
                state.moments[i] = m[i]
    res = evaluate(a.files, imp, core=CORE_OTC if a.core == "OTC" else CORE_OT,
                   h_core=a.h_core if a.h_core else H4OT, batch=a.batch, rimp=a.rimp, plus=plus, state=state,
                   exclusion=a.z_exciplex_exclusion)
    for d in res:
        if not a.quiet:
            print(f"# {d['path']}")
            print_balance(d)
        print(f"{d['path']}  E_imp {d['eimpu']:.10f} K  U+ {d['uplus']:.10f} K  E_tot {d['etot']:.10f} K  "
              f"N {d['natoms']:.6f}  d_CoM {d['dist']:.6f} A")
    return 0


if __name__ == "__main__":
    sys.exit(main())
