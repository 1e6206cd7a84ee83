"""GPU relaxations against committed oracle fixtures, mixed precision against
FP64, and continuous batching (a request solved alone and in a mix).

Tolerances: FP64 trajectories 1e-8 relative after 2000 steps; converged
E_imp within 1e-3 K of the FP64 oracle (the accuracy the mixed form must keep) for both the FP64
and the mixed FP32/FP64 path; a slot's result is independent of the others
(1e-12 relative psi, 1e-9 K energies)."""
import json
import os
import re
import subprocess

import numpy as np
import pytest

import eff_pot_amd as ep
from eff_pot_amd.abi import make_params, make_morse, CORE_OTC, PREC_MIXED, SolveOpts

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(__file__), "golden")
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def load(name):
    return json.load(open(os.path.join(GOLD, name)))


def test_cfg0_trajectory_matches_oracle(gpu_available):
    g = load("cfg0_oracle.json")
    c = g["config"]
    print("[cfg0] 64^3 N=300 2000 steps", flush=True)
    p = make_params(c["n"], xmax=c["xmax"], n4=c["n4"], core=CORE_OTC, h_core=c["h_core"])
    with ep.Context(p) as ctx:
        ctx.init_fermi(c["rho0"], 2.22 * c["n4"] ** (1 / 3), c["afermi"], c["noise"], c["seed"])
        ctx.set_impurity(make_morse())
        r = ctx.solve(SolveOpts(dt=g["dt"], tol=0.0, n_rk=3, pener=c["pener"], max_steps=c["steps"],
                                polish_steps=0, tol_fp32=0.0, check_finite=1))[0]
    assert r.steps == g["steps"]
    for k in ("etot", "eimpu", "ekin4", "elj4", "ealphas", "ecor4", "natoms", "mu"):
        ref = g["energies"][k]
        assert abs(getattr(r.e, k) - ref) <= 1e-8 * abs(ref), (k, getattr(r.e, k), ref)


def _conv32(precision, tol_fp32=1e-7):
    g = load("conv32_oracle.json")
    c = g["config"]
    p = make_params(c["n"], xmax=c["xmax"], n4=c["n4"], core=CORE_OTC, h_core=c["h_core"], precision=precision)
    ctx = ep.Context(p)
    ctx.init_fermi(c["rho0"], 2.22 * c["n4"] ** (1 / 3), c["afermi"], c["noise"], c["seed"])
    ctx.set_impurity(make_morse(rimp=tuple(c["rimp"])))
    r = ctx.solve(SolveOpts(dt=g["dt"], tol=c["tol"], n_rk=3, pener=c["pener"], max_steps=c["max_steps"],
                            polish_steps=c["max_steps"], tol_fp32=tol_fp32, check_finite=1))[0]
    return g, ctx, r


def test_converged_fp64_vs_oracle(gpu_available):
    print("[conv32] fp64", flush=True)
    g, ctx, r = _conv32(0)
    ctx.close()
    assert r.converged
    assert abs(r.e.eimpu - g["energies"]["eimpu"]) <= 1e-3
    assert abs(r.e.etot - g["energies"]["etot"]) <= 1e-3
    assert abs(r.steps - g["steps"]) <= 2 * g["config"]["pener"]


def test_converged_mixed_vs_oracle_and_fp64(gpu_available):
    print("[conv32] mixed", flush=True)
    g, c64, r64 = _conv32(0)
    d64 = c64.den()
    c64.close()
    g, c32, r32 = _conv32(PREC_MIXED)
    d32 = c32.den()
    c32.close()
    assert r32.converged
    assert abs(r32.e.eimpu - g["energies"]["eimpu"]) <= 1e-3, (r32.e.eimpu, g["energies"]["eimpu"])
    assert abs(r32.e.etot - g["energies"]["etot"]) <= 1e-3
    l2 = np.linalg.norm(d32 - d64) / np.linalg.norm(d64)
    print(f"mixed: steps={r32.steps} (fp64 {r64.steps}) dE_imp={r32.e.eimpu - r64.e.eimpu:.2e} K L2(rho)={l2:.2e}")
    # measured 8.6e-10 and 2.2e-9 K since a solve starts from the defect correction of its starting state
    assert l2 < 1e-7
    assert abs(r32.e.eimpu - r64.e.eimpu) < 1e-6


def test_mixed_stops_where_fp64_stops(gpu_available):
    """With tol_fp32 = tol (the benchmark's setting) the FP32 phase hands over
    to FP64 fields on the |dE/E| trend, so the mixed solve stops at the same
    energy check as the FP64 one and E_imp agrees far below 1e-3 K."""
    print("[conv32] mixed, tol_fp32 = tol", flush=True)
    g, c64, r64 = _conv32(0)
    c64.close()
    g, c32, r32 = _conv32(PREC_MIXED, tol_fp32=g["config"]["tol"])
    c32.close()
    print(f"steps mixed {r32.steps} fp64 {r64.steps}, dE_imp {r32.e.eimpu - r64.e.eimpu:.2e} K")
    assert r32.converged and r64.converged
    assert r32.steps == r64.steps
    assert abs(r32.e.eimpu - r64.e.eimpu) < 1e-6  # measured 1.9e-7 K


# requests (z0, seed, Morse D) and the slot each takes in the mix: request 3
# reuses slot 0 as soon as request 0 is done, while requests 1 and 2 run
REQS = [(0.5, 1, 25.0), (1.5, 2, 30.0), (-2.0, 3, 35.0), (1.0, 4, 28.0)]
SLOT_OF = [0, 1, 2, 0]


def _load(ctx, slot, z, seed, D=30.0):
    ctx.init_fermi(0.0218, 2.22 * 100 ** (1 / 3), 0.6, 1e-6, seed, slot=slot)
    ctx.set_impurity(make_morse(rimp=(0.0, 0.0, z), D=D), slot=slot)


def _req(ctx, slot, k):
    z, seed, D = REQS[k]
    _load(ctx, slot, z, seed, D)


@pytest.mark.parametrize("precision,cplx", [(0, 1), (PREC_MIXED, 1), (0, 0), (PREC_MIXED, 0)])
def test_same_result_alone_and_in_the_mix(precision, cplx, gpu_available):
    print(f"[mix] precision={precision} complex={cplx}", flush=True)
    opts = SolveOpts(dt=1.9e-4, tol=1e-8, n_rk=3, pener=50, max_steps=40000, polish_steps=40000, tol_fp32=3e-8,
                     check_finite=1)

    def params():
        return make_params(32, xmax=12.0, n4=100.0, core=CORE_OTC, h_core=2.377, precision=precision, batch=3,
                           psi_complex=cplx)

    alone = []
    for k in range(4):
        sk = SLOT_OF[k]
        with ep.Context(params()) as ctx:
            for s in range(3):
                if s != sk:
                    ctx.set_active(s, False)
            _req(ctx, sk, k)
            slots = ctx.new_slots()
            slots[sk].phase = ep.SLOT_START
            slots[sk].fp32 = 1 if precision else 0
            while slots[sk].phase not in (ep.SLOT_DONE, ep.SLOT_FAILED):
                ctx.tick(opts, slots)
            alone.append((slots[sk].e.eimpu, slots[sk].e.etot, slots[sk].steps, ctx.psi(sk)))
    # the mix: requests admitted at different times into one context, and a
    # slot re-used (its state is reset while the other slots keep stepping)
    with ep.Context(params()) as ctx:
        slots = ctx.new_slots()
        admit = {0: 0, 1: 37, 2: 180}
        running = {}  # slot -> request
        results = {}
        t = 0
        while len(results) < 4:
            for k, at in admit.items():
                if at == t:
                    _req(ctx, SLOT_OF[k], k)
                    slots[SLOT_OF[k]].phase = ep.SLOT_START
                    slots[SLOT_OF[k]].fp32 = 1 if precision else 0
                    running[SLOT_OF[k]] = k
            ctx.tick(opts, slots)
            for s, k in list(running.items()):
                if slots[s].phase in (ep.SLOT_DONE, ep.SLOT_FAILED):
                    results[k] = (slots[s].phase, slots[s].converged, slots[s].e.eimpu, slots[s].e.etot,
                                  slots[s].steps, ctx.psi(s))
                    del running[s]
                    if k == 0:
                        admit[3] = t + 1
            t += 1
        for k in range(4):
            ea, et, st, pa = alone[k]
            ph, conv, e_imp, e_tot, steps, psi = results[k]
            assert ph == ep.SLOT_DONE and conv
            assert steps == st
            assert abs(e_imp - ea) <= 1e-9 and abs(e_tot - et) <= 1e-9, (k, e_imp - ea, e_tot - et)
            assert np.linalg.norm(psi - pa) <= 1e-12 * np.linalg.norm(pa)


def test_nonfinite_density_is_reported(gpu_available):
    p = make_params(32, xmax=12.0, n4=100.0)
    with ep.Context(p) as ctx:
        psi = np.full(ctx.shape, 0.1, dtype=complex)
        psi[3, 4, 5] = np.nan
        ctx.upload_psi(psi)
        ctx.set_impurity(make_morse())
        with pytest.raises(ep.EffpotError) as ei:
            ctx.solve(SolveOpts(dt=1.9e-4, tol=1e-9, n_rk=3, pener=10, max_steps=100, polish_steps=0,
                                tol_fp32=0.0, check_finite=1))
        assert ei.value.code == ep.abi.ERR_NONFINITE


def test_fortran_driver_labels(gpu_available, tmp_path):
    exe = os.path.join(ROOT, "driver", "effpot_driver")
    if not os.path.exists(exe):
        pytest.skip("Fortran driver not built")
    nml = ("&input\n nx=32, ny=32, nz=32, xmax=12.d0, ymax=12.d0, zmax=12.d0, n4=100, mode=0,\n"
           " core4='OTC', h_core=2.377d0, selec_pi='Morse', zimp=1.5d0, niter=0, seed=99\n/\n")
    out = subprocess.run([exe], input=nml, capture_output=True, text=True, cwd=tmp_path, timeout=120)
    assert out.returncode == 0, out.stderr
    txt = out.stdout
    eint = float(re.search(r"Interaction energy \(X\*-He\) \.*: +(\S+)", txt).group(1))
    rcm = float(re.search(r"X-HE CoM distance:\s+(\S+)", txt).group(1))
    p = make_params(32, xmax=12.0, n4=100.0, core=CORE_OTC, h_core=2.377)
    with ep.Context(p) as ctx:
        ctx.init_fermi(0.0218, 2.22 * 100 ** (1 / 3), 0.6, 1e-6, 99)
        ctx.set_impurity(make_morse(rimp=(0.0, 0.0, 1.5)))
        ctx.poten()
        e = ctx.energy()[0]
    assert abs(eint - e["eHeX"]) < 1e-5
    assert abs(rcm - np.sqrt(e["xcm"] ** 2 + e["ycm"] ** 2 + (1.5 - e["zcm"]) ** 2)) < 1e-5
    # awk field positions used by get_impurity_energy.sh
    line = [l for l in txt.splitlines() if "Interaction energy (X*" in l][0]
    assert float(line.split()[4]) == eint


def test_fortran_driver_reads_snapshots(gpu_available, tmp_path):
    """readenc mode 3 in the Fortran driver: the library's reader (fast_io,
    formatted and unformatted files) and the driver's own list-directed READ
    give the energies of the same psi that the Python binding computes,
    including uplus with the default X+ potential (Ba_plus_2D2)."""
    exe = os.path.join(ROOT, "driver", "effpot_driver")
    if not os.path.exists(exe):
        pytest.skip("Fortran driver not built")
    print("[driver] mode 3", flush=True)
    p = make_params(32, xmax=12.0, n4=100.0, core=CORE_OTC, h_core=2.377)
    rimp = (0.25, -0.5, 1.75)
    with ep.Context(p) as ctx:
        ctx.init_fermi(0.0218, 2.22 * 100 ** (1 / 3), 0.6, 1e-6, 4242)
        psi = ctx.psi()
        ctx.set_impurity(make_morse(rimp=rimp))
        ctx.set_plus(ep.make_plus())  # the namelist default selec_plus = 'Ba_plus_2D2'
        ctx.poten()
        e = ctx.energy()[0]
    h = ep.Snapshot()
    h.nx = h.ny = h.nz = 32
    h.xmax = h.ymax = h.zmax = 12.0
    h.hx = h.hy = h.hz = 0.75
    h.rimp[:] = rimp
    h.ninvar = 1
    got = []
    for binary, fast in ((False, True), (True, True), (False, False)):
        name = f"den_{int(binary)}.dat"   # filedenin holds 60 characters: a name in the run's cwd
        ep.write_snapshot(str(tmp_path / name), h, psi, binary=binary)
        nml = (f"&input\n nx=32, ny=32, nz=32, xmax=12.d0, ymax=12.d0, zmax=12.d0, n4=100, mode=3,\n"
               f" filedenin='{name}', fast_io={'.true.' if fast else '.false.'},\n"
               " core4='OTC', h_core=2.377d0, selec_pi='Morse', niter=0\n/\n")
        out = subprocess.run([exe], input=nml, capture_output=True, text=True, cwd=tmp_path, timeout=120)
        assert out.returncode == 0, out.stderr
        eint = float(re.search(r"Interaction energy \(X\*-He\) \.*: +(\S+)", out.stdout).group(1))
        eplus = float(re.search(r"Interaction energy \(X\+-He\) \.*: +(\S+)", out.stdout).group(1))
        got.append((eint, eplus))
        assert abs(eint - e["eHeX"]) < 1e-5 and abs(eplus - e["uplus"]) < 1e-5 and e["uplus"] != 0.0
    assert got[0] == got[1] == got[2]


# An input.dat as written for the reference's eff_pot run: keys of its /input/
# namelist (DFT4He3d.f90:104-134) that this driver reads and does not use,
# plus the Morse impurity and core radius keys of this driver.
REF_INPUT = """&input
 title = 'He4 + Rb+ snapshots',
 fftwplan = 'FFTW_PATIENT',
 nthread = 20,
 filedenin = 'he4.input.density',
 fileout = 'DFT4He3d.res',
 mode = 0,
 n4 = 100,
 nx = 32, ny = 32, nz = 32,
 xmax = 12.d0, ymax = 12.d0, zmax = 12.d0,
 core4 = 'OTC', h_core = 2.377d0,
 selec_pi = 'Morse',
 npd = 13, icon = 13, niter = 0, pener = 50,
 lrkpc = .true., lrk = .false., deltat = 1.d-4, mimpur = 85.47d0,
 vximp = 0.d0, vyimp = 0.d0, vzimp = 0.d0,
 lsurf = .false., irespar = 0, pdenpar = 50, iter0 = 0, time0 = 0.d0,
 selec_gs = 'Ba_plus_gs', r_cutoff_gs = 1.05d0, umax_gs = 6.836177d3,
 L_v_aditional = .true.,
 Lfix_lambda = .false., Ev0 = 0.d0,
 z_exciplex_position = -3.5d0
/
"""


def _prep(text, x_plus, r_cutoff_plus, umax_plus):
    """The edits prep.sh:9-15 makes to input.dat: the filedenin line and the
    mode line replaced (djogger.dat, mode 3), the X+ potential keys added
    after &input, and the L_v_aditional line replaced (.false.)."""
    out = []
    for line in text.splitlines():
        if "filedenin" in line:
            line = " filedenin\t\t= 'djogger.dat'"
        elif "mode" in line:
            line = " mode\t\t\t= 3"
        elif "L_v_aditional" in line:
            line = " L_v_aditional\t\t= .false."
        out.append(line)
        if "&input" in line:
            out += [f" selec_plus\t\t= '{x_plus}'", f" r_cutoff_plus\t\t= {r_cutoff_plus}",
                    f" umax_plus\t\t= {umax_plus}"]
    return "\n".join(out) + "\n"


def _awk(text, pattern, field):
    """awk '/pattern/{print $field}' on text (one matching line expected)."""
    hits = [ln.split()[field - 1] for ln in text.splitlines() if re.search(pattern, ln)]
    assert len(hits) == 1, (pattern, hits)
    return hits[0]


def test_reference_snapshot_workflow(gpu_available, tmp_path):
    """prep.sh -> loop.sh -> get_impurity_energy.sh on this driver: an
    input.dat with the reference's keys, edited as prep.sh edits it
    (Rb_plus_Koutselos, r_cutoff_plus 2.d0, umax_plus 6.1105922888E+03), runs
    on two formatted snapshots linked to djogger.dat, and every field the
    scripts' awk lines pick (time, X-He CoM distance, Ekin, U*, U+, E_SO)
    parses as a number equal to what the Python binding computes."""
    exe = os.path.join(ROOT, "driver", "effpot_driver")
    if not os.path.exists(exe):
        pytest.skip("Fortran driver not built")
    print("[driver] reference workflow, Rb_plus_Koutselos", flush=True)
    (tmp_path / "input.dat").write_text(_prep(REF_INPUT, "Rb_plus_Koutselos", "2.d0", "6.1105922888E+03"))
    p = make_params(32, xmax=12.0, n4=100.0, core=CORE_OTC, h_core=2.377)
    plus = ep.make_plus("Rb_plus_Koutselos", 2.0, 6.1105922888e3)
    for k, (rimp, seed, t) in enumerate((((0.25, -0.5, 1.75), 31, 1.5), ((0.0, 0.75, -2.25), 32, 3.0))):
        with ep.Context(p) as ctx:
            ctx.init_fermi(0.0218, 2.22 * 100 ** (1 / 3), 0.6, 1e-6, seed)
            psi = ctx.psi()
            ctx.set_impurity(make_morse(rimp=rimp))
            ctx.set_plus(plus)
            ctx.poten()
            e = ctx.energy()[0]
        h = ep.Snapshot()
        h.nx = h.ny = h.nz = 32
        h.xmax = h.ymax = h.zmax = 12.0
        h.hx = h.hy = h.hz = 0.75
        h.rimp[:] = rimp
        h.ninvar = 1
        snap = tmp_path / f"density.{k:03d}.dat"
        ep.write_snapshot(str(snap), h, psi, binary=False)
        # the comment lines printoutc writes first (printoutc.f90:92-94)
        body = snap.read_text()
        snap.write_text(f"# Iteration {k}\n#  Total evolution time(ps)...:{t:20.12E}\n"
                        f"#  xcm,ycm,zcm:{0.0:16.8E}{0.0:16.8E}{0.0:16.8E}\n" + body)
        link = tmp_path / "djogger.dat"
        if link.exists() or link.is_symlink():
            link.unlink()
        link.symlink_to(snap.name)
        res = subprocess.run([exe], stdin=open(tmp_path / "input.dat"), capture_output=True, text=True,
                             cwd=tmp_path, timeout=120)
        assert res.returncode == 0, res.stderr
        head = "\n".join(link.read_text().splitlines()[:3])
        time_ps = float(_awk(head, r"Total evolution time", 5))
        rcm = float(_awk(res.stdout, r"X-HE CoM distance", 4))
        ekin = float(_awk(res.stdout, r"Kinetic energy \(X", 5))
        eint = float(_awk(res.stdout, r"Interaction energy \(X\*", 5))
        eplus = float(_awk(res.stdout, r"Interaction energy \(X\+", 5))
        eso = float(_awk(res.stdout, r"Spin-Orbit energy", 5))
        assert time_ps == t
        d = np.sqrt((rimp[0] - e["xcm"]) ** 2 + (rimp[1] - e["ycm"]) ** 2 + (rimp[2] - e["zcm"]) ** 2)
        assert abs(rcm - d) < 1e-5
        assert ekin == 0.0 and eso == 0.0
        assert abs(eint - e["eHeX"]) < 1e-5
        assert e["uplus"] != 0.0 and abs(eplus - e["uplus"]) < 1e-5, (eplus, e["uplus"])


@pytest.mark.parametrize("precision,cplx", [(0, 1), (PREC_MIXED, 0)])
def test_graph_replay_matches_eager(precision, cplx, gpu_available, monkeypatch):
    """Every replay of the captured predictor-corrector step, with eager energy
    work and synchronisation in between, equals the eager launch sequence
    (mixed: FP32 phase with its defect corrections, then the FP64 polish)."""
    print(f"[graph] replay vs eager precision={precision} complex={cplx}", flush=True)
    opts = SolveOpts(dt=1.9e-4, tol=0.0, n_rk=3, pener=1 if precision == 0 else 5, max_steps=60, polish_steps=30,
                     tol_fp32=0.0, check_finite=0)
    traj = {}
    for graphs in ("1", "0"):
        monkeypatch.setenv("EFFPOT_GRAPHS", graphs)
        p = make_params(32, xmax=12.0, n4=100.0, core=CORE_OTC, h_core=2.377, batch=2, precision=precision,
                        psi_complex=cplx)
        with ep.Context(p) as ctx:
            for s in range(2):
                _load(ctx, s, [0.5, -1.0][s], 11 + s, 25.0 + 5 * s)
            slots = ctx.new_slots()
            for s in range(2):
                slots[s].phase = ep.SLOT_START
                slots[s].fp32 = 1 if precision else 0
            es = []
            while any(slots[s].phase not in (ep.SLOT_DONE, ep.SLOT_FAILED) for s in range(2)):
                ctx.tick(opts, slots)
                es.append([(slots[s].e.etot, slots[s].e.eimpu) for s in range(2)])
            traj[graphs] = (es, [ctx.psi(s) for s in range(2)])
    assert len(traj["1"][0]) == len(traj["0"][0])
    for a, b in zip(traj["1"][0], traj["0"][0]):
        assert a == b
    for a, b in zip(traj["1"][1], traj["0"][1]):
        assert np.array_equal(a, b)


@pytest.mark.parametrize("precision,batch", [(0, 1), (PREC_MIXED, 3)])
def test_multi_step_graph_matches_single_steps(precision, batch, gpu_available):
    """graph_steps = 7 replays 7 predictor-corrector steps per graph launch
    whenever no slot is due an energy check or a phase change: the same
    kernels in the same order, so every slot's energies, step count and final
    wavefunction equal the one-step-per-tick run bit for bit.  7 does not
    divide pener = 25, and one slot is admitted 4 ticks late, so the
    single-step fallback runs too.  The multi-step run alternates 7- and
    5-step graphs tick by tick, so the graph is re-captured (after the
    stream drains) while slots are in flight."""
    print(f"[graph] {batch} slots precision={precision}: 7/5-step graphs vs single steps", flush=True)
    out = {}
    for gs in (0, 7):
        def opts_of(tick):
            k = gs if gs == 0 or tick % 2 == 0 else 5
            return SolveOpts(dt=1.9e-4, tol=0.0, n_rk=3, pener=25, max_steps=160, polish_steps=60, tol_fp32=0.0,
                             check_finite=0, graph_steps=k)
        p = make_params(32, xmax=12.0, n4=100.0, core=CORE_OTC, h_core=2.377, batch=batch, precision=precision,
                        psi_complex=0)
        with ep.Context(p) as ctx:
            for s in range(batch):
                _load(ctx, s, [0.5, -1.0, 1.5][s], 21 + s, 25.0 + 5 * s)
            slots = ctx.new_slots()
            ticks = 0
            while True:
                for s in range(batch):
                    if slots[s].phase == ep.SLOT_IDLE and ticks >= 4 * s:
                        slots[s].phase = ep.SLOT_START
                        slots[s].fp32 = 1 if precision else 0
                if all(slots[s].phase in (ep.SLOT_DONE, ep.SLOT_FAILED) for s in range(batch)):
                    break
                ctx.tick(opts_of(ticks), slots)
                ticks += 1
            out[gs] = ([(slots[s].steps, slots[s].e.etot, slots[s].e.eimpu, slots[s].rel_de) for s in range(batch)],
                       [ctx.psi(s) for s in range(batch)], ticks)
    assert out[7][2] < out[0][2]
    assert out[7][0] == out[0][0]
    for a, b in zip(out[7][1], out[0][1]):
        assert np.array_equal(a, b)


@pytest.mark.parametrize("precision", [0, PREC_MIXED])
def test_real_psi_path_is_bit_identical(precision, gpu_available):
    """Imaginary time from a real start keeps Im psi = 0 exactly, so the real
    storage path does the same arithmetic on half the bytes: equal up to the
    compiler's FMA contraction choices (a few ulp)."""
    print(f"[real psi] precision={precision}", flush=True)
    opts = SolveOpts(dt=1.9e-4, tol=0.0, n_rk=3, pener=25, max_steps=200, polish_steps=60, tol_fp32=0.0,
                     check_finite=1)
    out = {}
    for cplx in (1, 0):
        p = make_params(32, xmax=12.0, n4=100.0, core=CORE_OTC, h_core=2.377, precision=precision,
                        psi_complex=cplx)
        with ep.Context(p) as ctx:
            _load(ctx, 0, 1.0, 5, 25.0)
            r = ctx.solve(opts)[0]
            out[cplx] = (r.e.etot, r.e.eimpu, r.e.ekin4, r.e.mu, r.steps, ctx.psi())
    a, b = out[1], out[0]
    assert a[4] == b[4]
    for x, y in zip(a[:4], b[:4]):
        assert abs(x - y) <= 1e-13 * abs(x)
    assert np.abs(a[5].imag).max() == 0.0
    assert np.abs(a[5] - b[5]).max() <= 1e-14 * np.abs(a[5]).max()
    with ep.Context(make_params(32, xmax=12.0, psi_complex=0)) as ctx:
        with pytest.raises(ep.EffpotError):
            ctx.upload_psi(np.full(ctx.shape, 0.1 + 0.01j))
