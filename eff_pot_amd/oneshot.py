"""The eff_pot one-shot workflow as one batched GPU job.

The reference computes the impurity energy of a stored density one file per
process: loop.sh:9-13 and get_impurity_energy.sh:7-16 run the program once per
snapshot, which reads it (readenc mode 3, readenc.f90:151-168), builds the
impurity field at the snapshot's rimp (potenimp), calls poten and energy
(DFT4He3d.f90:437-439) and prints the energy balance.  Here the snapshots are
read with effpot_snapshot_read and evaluated `batch` at a time as the slots of
one context: one potenimp per slot, one poten and one energy launch sequence
for the whole batch.  Snapshots on different grids go to different contexts.

    python -m eff_pot_amd.oneshot --morse 30 4 1 den-*.dat

prints, per file, the reference's "FIRST ENERGY BALANCE" and impurity-energy
labels (formats 6050 / 6060 of the driver), so get_impurity_energy.sh's parsing
applies unchanged, followed by a one-line summary.
"""
import argparse
import sys
from collections import OrderedDict

from . import Context, make_params, make_morse, read_snapshot, CORE_OTC, CORE_OT, PREC_FP64, ICON_ZERO, H4OT


def _grid_key(h):
    return (h.nx, h.ny, h.nz, round(h.xmax, 12), round(h.ymax, 12), round(h.zmax, 12))


def evaluate(paths, impurity, core=CORE_OTC, h_core=2.377, icon=ICON_ZERO, batch=8, rimp=None):
    """E_imp and the energy balance of each snapshot in `paths`.

    impurity: an effpot Impurity (its rimp is replaced by the snapshot's own
    rimp unless `rimp` is given).  Returns one dict per path, in order, with
    the energies (effpot_energies fields) and the grid/impurity position used.
    """
    headers = [read_snapshot(p, header_only=True)[0] for p in paths]
    groups = OrderedDict()
    for i, h in enumerate(headers):
        groups.setdefault(_grid_key(h), []).append(i)
    out = [None] * len(paths)
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
                    imp = make_morse(rimp=pos, D=impurity.D, re=impurity.re, a=impurity.a,
                                     r_cutoff=impurity.r_cutoff, umax=impurity.umax, cap=impurity.cap)
                    imp.kind = impurity.kind
                    ctx.set_impurity(imp, slot=s)
                ctx.poten()
                es = ctx.energy()
                for s, i in enumerate(chunk):
                    d = dict(es[s])
                    d["path"] = paths[i]
                    d["rimp"] = tuple(rimp) if rimp is not None else tuple(headers[i].rimp)
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


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("files", nargs="+")
    ap.add_argument("--morse", type=float, nargs=3, metavar=("D", "RE", "A"), default=(30.0, 4.0, 1.0))
    ap.add_argument("--rimp", type=float, nargs=3, default=None, help="impurity position (default: each file's)")
    ap.add_argument("--core", choices=["OT", "OTC"], default="OTC")
    ap.add_argument("--h-core", type=float, default=2.377)
    ap.add_argument("--batch", type=int, default=8)
    ap.add_argument("--quiet", action="store_true", help="only the summary lines")
    a = ap.parse_args(argv)
    imp = make_morse(D=a.morse[0], re=a.morse[1], a=a.morse[2])
    res = evaluate(a.files, imp, core=CORE_OTC if a.core == "OTC" else CORE_OT,
                   h_core=a.h_core if a.h_core else H4OT, batch=a.batch, rimp=a.rimp)
    for d in res:
        if not a.quiet:
            print(f"# {d['path']}")
            print_balance(d)
        print(f"{d['path']}  E_imp {d['eimpu']:.10f} K  E_tot {d['etot']:.10f} K  N {d['natoms']:.6f}")
    return 0


if __name__ == "__main__":
    sys.exit(main())
