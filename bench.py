#!/usr/bin/env python3
"""Headline benchmark: converged impurity-energy solves per second at 128^3.

Workload (the production single-GPU configuration): He-4 droplet of N=1000
atoms on a 128^3 grid with spacing 0.5 A (xmax = 32 A), OT functional with
the alpha_s term (core 'OTC'), LJ core h = 2.377 A, isotropic Morse impurity
D=30 K, re=4 A, a=1/A near z0 = 0.5 R (R = 2.22 N^(1/3) = 22.2 A).  Each solve
starts from the seeded Fermi droplet, runs 3 RK-Gill steps and Hamming
predictor-corrector steps in imaginary time with renormalisation to N until
|dE/E| per step < 1e-9 (FP64 state at the end in every mode).

A "step" of this benchmark is one batch of B such solves run to convergence
(B droplets that differ in the seed of their initial noise, all with the
impurity at z0 = 0.5 R); value = solves / second.  One converged batch takes
about 100 s (~83,000 predictor-corrector steps per solve), so the timed region
is `--batches` batches (default 1) whatever `--steps` asks, and the line
reports the batches actually timed as `steps` (`steps_requested` keeps the
request).  `--warmup W` runs W x 100 untimed ticks of a loaded batch (graph
capture, caches, clocks) before the timed batch.
Modes: --mode mixed (default): FP64 wavefunction and integrator, FP32 mean
field (all FFTs and density stencils) with a defect correction refreshed from
the FP64 field at every energy check, FP64 reductions and energy checks, FP64
final polish; --mode fp64: the plain FP64 path, one solve per context.  The
mixed run also solves slot 0's input with the plain FP64 path (untimed) and
reports |dE_imp| and the speed-up against it.

Also reported: the live roofline of the dominant kernel (H psi + corrector),
timed with HIP events on the library's stream, and the CPU baseline (the
FP64 oracle, timed per step on a bounded sample on this host).
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1)
    ap.add_argument("--warmup", type=int, default=0)
    ap.add_argument("--mode", choices=["mixed", "fp64"], default="mixed")
    ap.add_argument("--batch", type=int, default=4)
    ap.add_argument("--n", type=int, default=128)
    ap.add_argument("--max-steps", type=int, default=100000)
    ap.add_argument("--tol", type=float, default=1e-9)
    ap.add_argument("--tol-fp32", type=float, default=None,
                    help="|dE/E| per step that ends the FP32 phase of a mixed solve (default tol)")
    ap.add_argument("--batches", type=int, default=1,
                    help="converged batches timed by the headline workload (each takes ~100 s)")
    ap.add_argument("--cpu-seconds", type=float, default=12.0,
                    help="wall-clock budget of the timed CPU-oracle sample (at least 2 PC steps)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-roofline", action="store_true")
    ap.add_argument("--no-tune", action="store_true")
    ap.add_argument("--no-ref", action="store_true",
                    help="skip the plain FP64 solve of slot 0's input that gives |dE_imp| and the speed-up")
    ap.add_argument("--tile", type=int, nargs=2, default=None, help="stencil tile (TX, z chunk) instead of autotuning")
    ap.add_argument("--workload", choices=["solve128", "cfg0", "scan", "serve", "stress256", "corrector", "poten",
                                           "oneshot", "realtime"],
                    default="solve128",
                    help="solve128 (default, the headline); the others are the batched scan, the serving stream, "
                         "the 256^3 stress test and the small CPU-runnable case")
    ap.add_argument("--npos", type=int, default=48)
    ap.add_argument("--nreq", type=int, default=32)
    ap.add_argument("--rate", type=float, default=0.5)
    ap.add_argument("--serve-budget-gb", type=float, default=96.0,
                    help="HBM the serving workload may reserve for solver state (admission by footprint)")
    ap.add_argument("--stress-steps", type=int, default=300)
    ap.add_argument("--seed", type=int, default=20261203, help="noise seed of the initial droplets")
    ap.add_argument("--graph-steps", type=int, default=24,
                    help="predictor-corrector steps per graph launch between energy checks (1: one per tick)")
    a = ap.parse_args()
    if a.tol_fp32 is None:
        a.tol_fp32 = a.tol
    return a


def workload(n, mode, batch):
    import eff_pot_amd as ep
    N4 = 1000.0
    h = 0.5 * 128 / n if n != 128 else 0.5
    xmax = h * n / 2
    prec = ep.PREC_MIXED if mode == "mixed" else ep.PREC_FP64
    p = ep.make_params(n, xmax=xmax, n4=N4, core=ep.CORE_OTC, h_core=2.377, precision=prec,
                       batch=batch if mode == "mixed" else 1, psi_complex=0)
    R = 2.22 * N4 ** (1.0 / 3.0)
    return p, R


def positions(R, B):
    """Every slot solves the headline workload (impurity at z0 = 0.5 R); the
    slots differ in the seed of their initial noise."""
    return [0.5 * R] * B


def load_batch(ctx, p, R, zs, seed):
    """Seeded Fermi droplets with the Morse impurity at z0 in every slot, phase START."""
    import eff_pot_amd as ep
    for b, z0 in enumerate(zs):
        ctx.init_fermi(0.0218, R, 0.6, 1e-6, seed + b, slot=b)
        ctx.set_impurity(ep.make_morse(rimp=(0.0, 0.0, z0)), slot=b)
    slots = ctx.new_slots()
    for b in range(len(zs)):
        slots[b].phase = ep.SLOT_START
        slots[b].fp32 = 1 if p.precision == ep.PREC_MIXED else 0
    return slots


def warm_ticks(ctx, p, R, zs, opts, nticks, seed):
    """Untimed ticks of a loaded batch: captures the PC-step graph of the slots'
    field precision and warms caches and clocks."""
    slots = load_batch(ctx, p, R, zs, seed)
    for _ in range(nticks):
        ctx.tick(opts, slots)
    ctx.sync()


def solve_batch(ctx, p, R, zs, opts, seed=20261203, count_ticks=False):
    """Load the B solves and tick them to convergence; returns the slots and the
    seconds from the start of the batch to each slot's convergence (and the
    number of ticks with count_ticks)."""
    import eff_pot_amd as ep
    slots = load_batch(ctx, p, R, zs, seed)
    lat = [None] * len(zs)
    t0 = time.perf_counter()
    t_report = t0
    nt = 0
    while any(x is None for x in lat):
        ctx.tick(opts, slots)
        nt += 1
        for b in range(len(zs)):
            if lat[b] is None and slots[b].phase in (ep.SLOT_DONE, ep.SLOT_FAILED):
                lat[b] = time.perf_counter() - t0
        if slots[0].since == 0 and time.perf_counter() - t_report > 30.0:  # after an energy check
            t_report = time.perf_counter()
            print(f"[bench] {t_report - t0:.0f} s: steps {[slots[b].steps for b in range(len(zs))]}, "
                  f"converged {sum(x is not None for x in lat)}/{len(zs)}", file=sys.stderr, flush=True)
    if count_ticks:
        return [slots[b] for b in range(len(zs))], lat, nt
    return [slots[b] for b in range(len(zs))], lat


def percentile(xs, q):
    xs = sorted(xs)
    if not xs:
        return None
    k = (len(xs) - 1) * q
    f = int(k)
    c = min(f + 1, len(xs) - 1)
    return xs[f] + (xs[c] - xs[f]) * (k - f)


def traffic_of(kernel_key):
    """HBM bytes per launch of a kernel, measured with rocprofv3 FETCH_SIZE and
    WRITE_SIZE (calibrated against a copy kernel) and recorded in
    profiles/traffic.json; None when it has not been measured."""
    path = os.path.join(ROOT, "profiles", "traffic.json")
    try:
        with open(path) as f:
            t = json.load(f)
        return t["kernels"][kernel_key]["bytes_per_launch"]
    except (OSError, KeyError, ValueError):
        return None


def common_line(a, metric, value, unit, elapsed, dtype, workload, **extra):
    d = {"metric": metric, "value": value, "unit": unit, "n_gpus": 1, "steps": a.steps, "warmup": a.warmup,
         "ms_per_step": 1000.0 * elapsed / max(a.steps, 1), "higher_is_better": True, "scaling": "weak",
         "vs_baseline": None, "dtype": dtype, "data": "synthetic (seeded Fermi droplets + 1e-6 noise)",
         "config": {"workload": workload}}
    d["config"].update(extra)
    return d


def run_cfg0(a):
    """3 RK + 1997 PC FP64 steps on the N=300, 64^3, h=0.75 A droplet."""
    import eff_pot_amd as ep
    p = ep.make_params(64, xmax=24.0, n4=300.0, core=ep.CORE_OTC, h_core=2.377, psi_complex=0)
    opts = ep.SolveOpts(dt=ep.default_dt(p), tol=0.0, n_rk=3, pener=100, max_steps=2000, polish_steps=0,
                        tol_fp32=0.0, check_finite=0, graph_steps=a.graph_steps)
    with ep.Context(p) as c:
        t0 = time.perf_counter()
        for _ in range(a.steps):
            c.init_fermi(0.0218, 2.22 * 300 ** (1 / 3), 0.6, 1e-6, 20261203)
            c.set_impurity(ep.make_morse())
            r = c.solve(opts)[0]
        el = time.perf_counter() - t0
    return common_line(a, "2000-step relaxations/s (N=300, 64^3)", a.steps / el, "runs/s", el, "f64",
                       "He4 N=300, 64^3 h=0.75 A, OTC h_core=2.377, Morse at the centre, 3 RK + 1997 PC steps",
                       E_tot_K=r.e.etot, E_imp_K=r.e.eimpu, step_ms=1000 * el / a.steps / 2000)


def host_info():
    """The host the CPU baseline ran on: CPU model, logical CPUs, OpenMP
    threads, and whether an FFTW3 library is installed (the oracle does not
    link one: it runs its own FFT)."""
    import ctypes.util
    model = None
    try:
        for ln in open("/proc/cpuinfo"):
            if ln.startswith("model name"):
                model = ln.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return {"cpu_model": model, "nproc": os.cpu_count(), "omp_num_threads": os.environ.get("OMP_NUM_THREADS"),
            "fftw3_installed": ctypes.util.find_library("fftw3") is not None, "oracle_fft": "own (no FFTW)"}


def run_scan(a):
    from eff_pot_amd import scan
    prec = 1 if a.mode == "mixed" else 0
    t0 = time.perf_counter()
    res, meta = scan.run_scan(n=96, spacing=0.6, natoms=1000.0, npos=a.npos, precision=prec, tol=a.tol,
                              tol_fp32=a.tol_fp32, max_steps=a.max_steps,
                              progress=lambda m: print("[scan] " + m, file=sys.stderr, flush=True))
    el = time.perf_counter() - t0
    return common_line(a, "V_eff(z0) scan: converged positions/s (N=1000, 96^3)", len(res) / el, "solves/s", el,
                       "fp32 fields + f64" if prec else "f64",
                       f"He4 N=1000, 96^3 h=0.6 A, {a.npos} impurity positions z0 = 0..1.6R in one batched context",
                       mode=a.mode, curve=[[r["z0"], r["eimp"], r["steps"], r["converged"], r["inside"]] for r in res])


def run_serve(a):
    from eff_pot_amd import serve
    reqs = serve.make_stream(a.nreq, rate=a.rate, seed=7)
    budget = int(a.serve_budget_gb * 2 ** 30)
    srv = serve.Server(budget_bytes=budget, tol=a.tol, tol_fp32=a.tol_fp32, max_steps=a.max_steps,
                       precision=1 if a.mode == "mixed" else 0)
    t0 = time.perf_counter()
    try:
        srv.run(reqs, progress=lambda r: print(f"[serve] {time.perf_counter() - t0:.0f} s: request {r.rid} "
                                               f"grid {r.grid} steps {r.steps} failed {r.failed}",
                                               file=sys.stderr, flush=True))
        stats = srv.stats()
    finally:
        srv.close()
    el = time.perf_counter() - t0
    lat = [r.finish - r.arrival for r in reqs if not r.failed]
    ok = [r for r in reqs if r.converged]
    return common_line(a, "served impurity-energy solves/s (continuous batching)", len(ok) / el, "solves/s", el,
                       "fp32 fields + f64" if a.mode == "mixed" else "f64",
                       f"{a.nreq} requests, Poisson rate {a.rate}/s, N log-uniform [100,3000], grids 64/96/128 at 0.75 A",
                       p50_latency_s=serve.percentile(lat, 0.5), p99_latency_s=serve.percentile(lat, 0.99),
                       admission=stats, converged=len(ok),
                       grids={g: sum(1 for r in reqs if r.grid == g) for g in (64, 96, 128)})


def run_corrector(a):
    """The headline context with B loaded slots, 3 RK steps to fill the
    history, then the corrector timed inside 20 whole predictor-corrector
    steps: the workload the traffic counters are read on."""
    import eff_pot_amd as ep
    B = a.batch if a.mode == "mixed" else 1
    p, R = workload(a.n, a.mode, B)
    dt = ep.default_dt(p)
    with ep.Context(p) as c:
        tile = tuple(a.tile) if a.tile else c.autotune(reps=5)
        c.set_tiles(*tile)
        for b, z0 in enumerate(positions(R, B)):
            c.init_fermi(0.0218, R, 0.6, 1e-6, a.seed + b, slot=b)
            c.set_impurity(ep.make_morse(rimp=(0.0, 0.0, z0)), slot=b)
        c.poten()
        for _ in range(3):
            c.step_rk(dt)
            c.renormalize()
        ms = c.kernel_time(3, reps=20)
    per_pt = 72 if a.mode == "mixed" else 80
    byt = per_pt * a.n ** 3 * B
    return common_line(a, "H psi + corrector GB/s (algorithmic bytes)", byt / (ms * 1e-3) / 1e9, "GB/s", ms * 1e-3,
                       "f64 state, fp32 fields" if a.mode == "mixed" else "f64",
                       f"corrector of the {a.n}^3 N=1000 workload, {B} slots, inside whole PC steps",
                       mode=a.mode, batch=B, stencil_tile=list(tile), ms_per_launch=ms, bytes_per_launch=byt)


POTEN_LAUNCHES = [  # name, algorithmic bytes per slot in units of (N s, M s), OTC, plane transforms
    ("A x r2c (1 field)", 1, 2), ("A y fwd (1)", 0, 4), ("A z conv (1 in, 3 out)", 0, 2 + 6 + 3),
    ("A y inv (3)", 0, 12), ("A x c2r (3)", 3, 6), ("k_grad", 10, 0),
    ("B x r2c (4)", 4, 8), ("B y fwd (4)", 0, 16), ("B z conv (4 in, 4 out)", 0, 20), ("B y inv (4)", 0, 16),
    ("B x c2r (4)", 4, 8), ("k_div", 9, 0), ("C x r2c (1)", 1, 2), ("C y fwd (1)", 0, 4),
    ("C z conv (1 in, 1 out)", 0, 5), ("C y inv (1)", 0, 4), ("C x c2r (1)", 1, 2), ("k_pot4", 8, 0)]


def poten_launch_bytes(k, n, mode, B, plane, iplane=0):
    """Algorithmic bytes of launch k of poten (OTC) for B slots.  k_pot4 reads
    delj4, dencg, sto1, ua, div, den and writes pot4 and ualphas; with FP32
    fields it also reads the defect correction (4 B).  The forward y passes
    are absent when fused into the plane launch, the inverse y passes when
    fused into the inverse plane launch (which moves what the x c2r pass
    alone moves: the spectra read, the real fields written)."""
    s = 4 if mode == "mixed" else 8
    N, M = n ** 3, (n // 2 + 1) * n * n
    if plane and k in (1, 7, 13):
        return 0
    if iplane and k in (3, 9, 15):
        return 0
    _, cn, cm = POTEN_LAUNCHES[k]
    if plane and mode == "mixed" and k in (5, 11):
        # FP32 fields with the plane transforms: k_div forms the density
        # gradient from den (reads falfs, int x/y/z, den; writes 2), so k_grad
        # does not store dx, dy, dz (writes 4)
        cn = 7
    if not plane and k in (5, 6):
        # separate x pass: stage B's x transform reads dx, dy, dz and falfs
        # (once, as an array) and forms falfs * d{x,y,z}den as it loads, so
        # k_grad stores 4 fields instead of 7
        cn = 7 if k == 5 else 5
    b = s * (cn * N + cm * M)
    if k == 17 and mode == "mixed":
        b += 4 * N
    return B * b


def pc_step_bytes(n, mode, B, plane, iplane=0):
    """Algorithmic bytes of one predictor-corrector step of B slots (real psi,
    OTC): every array a launch reads or writes counted once.  Per point and
    slot, with s the field size (4 B mixed, 8 B FP64) and psi and the
    integrator history FP64: H psi + predictor/modifier reads psi, o3, h1, h2,
    pc and pot4, uimp and writes o3, pc, h2 and the density (5*8 + 2s + 3*8 +
    s); H psi + corrector reads psi, o3, o2, h1, h2, pc, pot4, uimp and writes
    pc and psi (6*8 + 2s + 2*8); the renormalisation reads and writes psi
    and, for FP64 fields, writes the density (2*8 [+ s]); poten is the sum of
    its launches (poten_launch_bytes; the predicted density is read through
    its factor to n4 atoms, no pass of its own)."""
    s = 4 if mode == "mixed" else 8
    N = n ** 3
    # the renormalisation stores the density only for FP64 fields
    psi_side = (40 + 2 * s + 24 + s) + (48 + 2 * s + 16) + (16 + (0 if mode == "mixed" else s))
    poten = sum(poten_launch_bytes(k, n, mode, 1, plane, iplane) for k in range(len(POTEN_LAUNCHES)))
    return B * (psi_side * N + poten)


def pc_step_fused_bytes(n, mode, B):
    """Bytes of one PC step of B slots under the fused model: every logical
    array touched once per fused kernel, a 3-D FFT one read plus one write of
    its half spectrum, k-space kernels read once.  poten with the OTC terms
    (derden + term_alfa) is 37 s N + 34 s M per slot (OT part 11 s N + 13 s M,
    OTC extra 26 s N + 21 s M); the psi side is that of pc_step_bytes (psi and
    the integrator FP64 in both modes).  Lower than pc_step_bytes, which counts
    what each of this implementation's launches moves."""
    s = 4 if mode == "mixed" else 8
    N, M = n ** 3, (n // 2 + 1) * n * n
    psi_side = (40 + 2 * s + 24 + s) + (48 + 2 * s + 16) + (16 + (0 if mode == "mixed" else s))
    return B * (psi_side * N + 37 * s * N + 34 * s * M)


def run_poten(a):
    """Each launch of the mean-field sequence of the headline context timed
    alone (20 back-to-back launches, HIP events) with its algorithmic bytes."""
    import eff_pot_amd as ep
    B = a.batch if a.mode == "mixed" else 1
    p, R = workload(a.n, a.mode, B)
    out = []
    with ep.Context(p) as c:
        tile = tuple(a.tile) if a.tile else c.autotune(reps=5)
        c.set_tiles(*tile)
        for b, z0 in enumerate(positions(R, B)):
            c.init_fermi(0.0218, R, 0.6, 1e-6, a.seed + b, slot=b)
            c.set_impurity(ep.make_morse(rimp=(0.0, 0.0, z0)), slot=b)
        c.kernel_time(2, reps=2)
        tot = 0.0
        plane = c.plane_fused(a.mode == "mixed")
        iplane = c.plane_inverse_fused(a.mode == "mixed")
        packed = c.packed_layout(a.mode == "mixed")
        zth = c.zconv_threads()
        for k, (name, _, _) in enumerate(POTEN_LAUNCHES):
            ms = c.kernel_time(100 + k, reps=20)
            byt = poten_launch_bytes(k, a.n, a.mode, B, plane, iplane)
            out.append({"launch": k, "name": name, "ms": ms, "GB_s": byt / (ms * 1e-3) / 1e9 if byt else 0.0})
            tot += ms
        whole = c.kernel_time(2, reps=10)
    return common_line(a, "mean-field (poten) launches: ms and algorithmic GB/s", whole, "ms", whole * 1e-3,
                       "fp32 fields" if a.mode == "mixed" else "f64",
                       f"poten of the {a.n}^3 N=1000 workload, {B} slots", mode=a.mode, batch=B,
                       stencil_tile=list(tile), plane=plane, inverse_plane=iplane, packed=packed,
                       zconv_threads=zth, sum_of_launches_ms=tot,
                       launches=out)


def run_stress(a):
    """256^3, h=0.4 A, N=5000: FP64 single solve and a batch of 8 mixed, timed over a bounded number of steps."""
    import eff_pot_amd as ep
    out = {}
    R = 2.22 * 5000 ** (1 / 3)
    for label, prec, B in (("fp64_single", 0, 1), ("mixed_batch8", 1, 8)):
        p = ep.make_params(256, xmax=51.2, n4=5000.0, core=ep.CORE_OTC, h_core=2.377, precision=prec, batch=B,
                           psi_complex=0)
        opts = ep.SolveOpts(dt=ep.default_dt(p), tol=0.0, n_rk=3, pener=100, max_steps=a.stress_steps,
                            polish_steps=0, tol_fp32=0.0, check_finite=0, graph_steps=a.graph_steps)
        with ep.Context(p) as c:
            for b in range(B):
                c.init_fermi(0.0218, R, 0.6, 1e-6, 11 + b, slot=b)
                c.set_impurity(ep.make_morse(rimp=(0.0, 0.0, 0.5 * R * b / max(B - 1, 1))), slot=b)
            t0 = time.perf_counter()
            r = c.solve(opts)
            el = time.perf_counter() - t0
            s = 4 if prec else 8
            ms_pred, ms_corr, ms_pot = c.kernel_time(0, 5), c.kernel_time(1, 5), c.kernel_time(2, 3)
        n = 256 ** 3
        per_pt = 76 if prec else 88
        out[label] = {"ms_per_step": 1000 * el / a.stress_steps, "slot_steps_per_s": B * a.stress_steps / el,
                      "E_imp_K": [x.e.eimpu for x in r],
                      "pred_GBs": per_pt * n * B / (ms_pred * 1e-3) / 1e9,
                      "corr_GBs": per_pt * n * B / (ms_corr * 1e-3) / 1e9,
                      "poten_ms": ms_pot}
    return common_line(a, "256^3 stress: slot-steps/s", out["mixed_batch8"]["slot_steps_per_s"], "slot-steps/s",
                       1.0, "f64 / fp32 fields", "He4 N=5000, 256^3 h=0.4 A", **out)


def run_oneshot(a):
    """Impurity energies of stored snapshots (the eff_pot one-shot job): 32
    snapshots of the 128^3 N=1000 droplet, evaluated 8 at a time; the rate with
    the unformatted file reads, the formatted (text) read rate, the GPU part
    alone, and the oracle's poten + energy of one snapshot for the CPU."""
    import tempfile
    import eff_pot_amd as ep
    from eff_pot_amd import oneshot
    K, B = 32, 8
    p, R = workload(128, "fp64", 1)
    d = tempfile.mkdtemp(prefix="effpot_oneshot_")
    paths, txt = [], []
    with ep.Context(p) as c:
        for k in range(K):
            c.init_fermi(0.0218, R, 0.6, 1e-6, 100 + k)
            h = ep.Snapshot()
            h.nx = h.ny = h.nz = 128
            h.xmax = h.ymax = h.zmax = p.xmax
            h.hx = h.hy = h.hz = 2 * p.xmax / 128
            h.rimp[:] = (0.0, 0.0, 0.5 * R * k / K)
            path = os.path.join(d, f"den{k}.bin")
            ep.write_snapshot(path, h, c.psi(), binary=True)
            paths.append(path)
            if k < 2:
                t = os.path.join(d, f"den{k}.txt")
                ep.write_snapshot(t, h, c.psi(), binary=False)
                txt.append(t)
    imp = ep.make_morse()
    oneshot.evaluate(paths[:B], imp, batch=B)  # warm-up
    t0 = time.perf_counter()
    res = oneshot.evaluate(paths, imp, batch=B)
    el = time.perf_counter() - t0
    t1 = time.perf_counter()
    for t in txt:
        ep.read_snapshot(t)
    el_txt = (time.perf_counter() - t1) / len(txt)
    # the GPU part alone: B snapshots already resident
    pb = ep.make_params(128, xmax=p.xmax, n4=1.0, core=ep.CORE_OTC, h_core=2.377, batch=B, psi_complex=1)
    with ep.Context(pb) as c:
        for s_ in range(B):
            c.upload_psi(ep.read_snapshot(paths[s_])[1], slot=s_)
        plus = ep.make_plus()
        c.sync()
        t2 = time.perf_counter()
        reps = 5
        for _ in range(reps):
            for s_ in range(B):
                c.set_impurity(ep.make_morse(rimp=(0.0, 0.0, 0.1 * s_)), slot=s_)
                c.set_plus(plus, slot=s_)
            c.poten()
            c.energy()
        c.sync()
        el_gpu = (time.perf_counter() - t2) / (reps * B)
    cpu = None
    if not a.no_cpu:
        from oracle.oracle import Oracle
        o = Oracle(p)
        o.set_psi(ep.read_snapshot(paths[0])[1])
        t3 = time.perf_counter()
        o.set_impurity(imp)
        o.set_plus(ep.make_plus())
        o.poten()
        o.energy()
        el_cpu = time.perf_counter() - t3
        o.close()
        cpu = {"value": 1.0 / el_cpu, "unit": "snapshots/s", "cores": int(os.environ.get("OMP_NUM_THREADS", "1")),
               "kind": "port", "sample": "potenimp + poten + energy of one 128^3 snapshot (resident)",
               "host": host_info()}
    for f_ in paths + txt:
        os.remove(f_)
    os.rmdir(d)
    line = common_line(a, "snapshot impurity energies/s (128^3)", K / el, "snapshots/s", el, "f64",
                       "He4 N=1000 snapshots, 128^3 h=0.5 A, Morse, Ba_plus_2D2 uplus; 32 unformatted files, "
                       "batch 8", text_read_s_per_file=el_txt, gpu_only_ms_per_snapshot=1000 * el_gpu,
                       E_imp_K_first=res[0]["eimpu"])
    line["cpu_baseline"] = cpu
    return line


def run_realtime(a):
    """Real-time PC steps of the 128^3 N=1000 droplet (complex psi, FP64) with a
    moving 40 u Morse impurity, against imaginary-time steps of the same state;
    plus the set-up of a D-state impurity (instates + uimp) at 128^3."""
    import eff_pot_amd as ep
    p, R = workload(128, "fp64", 1)
    p.psi_complex = 1
    nstep = max(a.steps, 50)
    with ep.Context(p) as c:
        c.init_fermi(0.0218, R, 0.6, 1e-6, 7)
        c.set_impurity(ep.make_morse(rimp=(0.0, 0.0, 0.5 * R)))
        c.poten()
        dt = ep.default_dt(p)
        for _ in range(3):
            c.step_rk(dt)
            c.renormalize()
        for _ in range(5):
            c.step_pc(dt)
            c.renormalize()
        c.sync()
        t0 = time.perf_counter()
        for _ in range(nstep):
            c.step_pc(dt)
        c.sync()
        el_it = (time.perf_counter() - t0) / nstep
        c.set_real_time(True)
        c.set_impurity_motion([0.0, 0.0, -2.0], 40.0)
        for _ in range(3):
            c.step_rk(0.5 * dt)
        c.sync()
        t1 = time.perf_counter()
        for _ in range(nstep):
            c.step_pc(0.5 * dt)
        c.sync()
        el_rt = (time.perf_counter() - t1) / nstep
        r, v, acc = c.impurity()
        c.set_real_time(False)
        t2 = time.perf_counter()
        st = c.set_impurity_state(ep.make_state_imp(ep.STATE_D, rimp=(0.0, 0.0, 0.5 * R), instate=0, als=2.0))
        c.sync()
        el_state = time.perf_counter() - t2
    return common_line(a, "real-time PC steps/s (128^3, moving impurity)", 1.0 / el_rt, "steps/s", el_rt * nstep,
                       "f64 complex", "He4 N=1000, 128^3 h=0.5 A, OTC, complex psi, 40 u Morse impurity moving",
                       ms_per_rt_step=1000 * el_rt, ms_per_imag_step_complex=1000 * el_it,
                       d_state_setup_ms=1000 * el_state, impurity_r=list(r), impurity_v=list(v),
                       d_state_eigen=list(st.eigen))


def main():
    a = parse()
    if a.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; multi-GPU runs are not part of it "
                                   "(independent solves would be sharded one context per GPU)",
                          "n_gpus": a.gpus}))
        sys.exit(2)
    import eff_pot_amd as ep
    if a.workload != "solve128":
        fn = {"cfg0": run_cfg0, "scan": run_scan, "serve": run_serve, "stress256": run_stress,
              "corrector": run_corrector, "poten": run_poten, "oneshot": run_oneshot,
              "realtime": run_realtime}[a.workload]
        print(json.dumps(fn(a)), flush=True)
        return

    B = a.batch if a.mode == "mixed" else 1
    p, R = workload(a.n, a.mode, B)
    dt = ep.default_dt(p)
    opts = ep.SolveOpts(dt=dt, tol=a.tol, n_rk=3, pener=50, max_steps=a.max_steps, polish_steps=a.max_steps,
                        tol_fp32=a.tol_fp32, check_finite=0, graph_steps=a.graph_steps)
    zs = positions(R, B)
    ctx = ep.Context(p)
    if a.tile:
        tile = tuple(a.tile)
        ctx.set_tiles(*tile)
    else:
        tile = ctx.autotune(reps=5) if not a.no_tune else None   # setup, before any state is loaded
    # warm-up: W x 100 ticks of a loaded batch (graph capture of both precisions'
    # PC steps, caches, clocks); the timed batch reloads every slot from scratch
    if a.warmup > 0:
        warm_ticks(ctx, p, R, zs, opts, 100 * a.warmup, seed=a.seed + 1000)
    ctx.sync()
    t0 = time.perf_counter()
    results, lats, ticks = [], [], 0
    for _ in range(max(a.batches, 1)):
        results, lat, nt = solve_batch(ctx, p, R, zs, opts, seed=a.seed, count_ticks=True)
        lats += lat
        ticks += nt
    ctx.sync()
    t1 = time.perf_counter()
    elapsed = t1 - t0
    # slot 0's density, before the roofline timings overwrite the states
    rho0 = np.abs(ctx.psi(0)) ** 2 if (a.mode == "mixed" and not a.no_ref) else None
    nb = max(a.batches, 1)
    nsolves = B * nb
    conv = all(r.converged for r in results)
    steps_per_solve = sum(r.steps for r in results) / len(results)
    eimp = [r.e.eimpu for r in results]

    line = {
        "metric": "converged E_imp solves/s @128^3 (|dE|<=1e-3 K); p50/p99 latency; HBM GB/s",
        "value": nsolves / elapsed,
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": nb,
        "steps_requested": a.steps,
        "warmup": a.warmup,
        "ms_per_step": 1000.0 * elapsed / nb,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64 state, fp32 fields" if a.mode == "mixed" else "f64",
        "data": "synthetic (seeded Fermi droplet + 1e-6 noise)",
        "p50_latency_s": percentile(lats, 0.5),
        "p99_latency_s": percentile(lats, 0.99),
        "config": {"workload": f"He4 N=1000 droplet, {a.n}^3 grid h={2 * p.xmax / a.n:.3f} A, OTC h_core=2.377, "
                               f"Morse D=30 re=4 a=1 at z0~0.5R, to |dE/E|<1e-9/step",
                   "mode": a.mode, "batch": B, "dt": dt, "converged": conv, "stencil_tile": tile,
                   "steps_per_solve": steps_per_solve, "steps": [r.steps for r in results], "E_imp_K": eimp,
                   "E_tot_K": [r.e.etot for r in results],
                   # |H psi - mu psi| / (|mu| |psi|) at the stop: how far from an eigenstate
                   "residual": [r.e.resid for r in results]},
        # every tick of the timed batches is one graph-replayed PC step of all
        # running slots (plus the energy checks every 50 steps)
        "ticks": ticks,
        "ms_per_tick": 1000.0 * elapsed / max(ticks, 1),
    }
    # ---- roofline of the dominant kernel: H psi + corrector ------------------
    if not a.no_roofline:
        Npts = a.n ** 3
        # corrector launches timed with HIP events on the library's stream inside
        # whole predictor-corrector steps of the converged states (the caches
        # hold what a relaxation leaves in them); back-to-back launches of the
        # corrector alone are reported too (their inputs partly stay in the
        # 256 MB last-level cache, so they read faster than a relaxation sees)
        ms = ctx.kernel_time(3, reps=20)
        ms_b2b = ctx.kernel_time(1, reps=20)
        # H psi + corrector on a real wavefunction, per point and slot: reads psi, psi_n,
        # psi_{n-2}, Hpsi_n, Hpsi_{n-1}, pc (FP64) and pot4, uimp; writes psi_{n+1} and pc
        # (FP64); the density is stored by the renormalisation that follows.  pot4/uimp are
        # FP64 (80 B) or FP32 in the mixed mode (72 B).
        per_pt = 72 if a.mode == "mixed" else 80
        bytes_alg = per_pt * Npts * B
        ach = bytes_alg / (ms * 1e-3) / 1e9
        tx = tile[0] if tile else 64
        line["roofline"] = {"bound": "hbm", "kernel": f"k_step<double,double,{'float' if a.mode == 'mixed' else 'double'},"
                                                      f"{tx},1> (H psi + Hamming corrector)",
                            "achieved": ach, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                            "frac": ach / HBM_PEAK_GBS,
                            "traffic": traffic_of(f"corrector_{a.mode}_{a.n}_b{B}"), "ms_per_launch": ms,
                            "ms_per_launch_back_to_back": ms_b2b,
                            "bytes_per_launch": bytes_alg}
        line["poten_ms"] = ctx.kernel_time(2, reps=5)
        # the whole PC step (predictor, poten, corrector, norm, renormalisation)
        # against its algorithmic bytes: the fraction of HBM peak a step reaches
        ms_step = ctx.kernel_time(4, reps=10)
        step_b = pc_step_bytes(a.n, a.mode, B, ctx.plane_fused(a.mode == "mixed"),
                               ctx.plane_inverse_fused(a.mode == "mixed"))
        line["pc_step"] = {"ms": ms_step, "bytes": step_b, "achieved": step_b / (ms_step * 1e-3) / 1e9,
                           "peak": HBM_PEAK_GBS, "unit": "GB/s",
                           "frac": step_b / (ms_step * 1e-3) / 1e9 / HBM_PEAK_GBS, "slots": B}
        fused_b = pc_step_fused_bytes(a.n, a.mode, B)
        line["pc_step"]["fused_model"] = {"bytes": fused_b, "achieved": fused_b / (ms_step * 1e-3) / 1e9,
                                          "frac": fused_b / (ms_step * 1e-3) / 1e9 / HBM_PEAK_GBS}
    ctx.close()
    # ---- the plain FP64 form on slot 0's input (untimed reference) -----------
    if a.mode == "mixed" and not a.no_ref:
        pr, _ = workload(a.n, "fp64", 1)
        with ep.Context(pr) as cr:
            cr.autotune(reps=5)
            tr0 = time.perf_counter()
            (r64,), _ = solve_batch(cr, pr, R, zs[:1], opts, seed=a.seed)
            tr = time.perf_counter() - tr0
            rho64 = np.abs(cr.psi(0)) ** 2
        line["fp64_reference"] = {"solves_per_s": 1.0 / tr, "E_imp_K": r64.e.eimpu, "steps": r64.steps,
                                  "residual": r64.e.resid,
                                  "dE_imp_K": abs(results[0].e.eimpu - r64.e.eimpu),
                                  # ||rho_mixed - rho_fp64||_2 / ||rho_fp64||_2 of slot 0
                                  "L2_density": float(np.linalg.norm(rho0 - rho64) / np.linalg.norm(rho64)),
                                  "max_dE_imp_K_all_slots": max(abs(r.e.eimpu - r64.e.eimpu) for r in results),
                                  "speedup": line["value"] * tr}
    # ---- CPU baseline: the FP64 oracle on the same workload, bounded sample ---
    if not a.no_cpu:
        from oracle.oracle import Oracle
        import eff_pot_amd as ep2
        pc = ep2.make_params(a.n, xmax=p.xmax, n4=1000.0, core=ep2.CORE_OTC, h_core=2.377)
        o = Oracle(pc)
        o.init_fermi(0.0218, R, 0.6, 1e-6, 20261203)
        o.set_impurity(ep2.make_morse(rimp=(0.0, 0.0, zs[0])))
        o.poten()
        o.step_rk(dt)
        o.renormalize()
        o.step_rk(dt)
        o.renormalize()
        o.step_rk(dt)
        o.renormalize()
        tc0 = time.perf_counter()
        ncpu = 0
        while ncpu < 2 or time.perf_counter() - tc0 < a.cpu_seconds:
            o.step_pc(dt)
            o.renormalize()
            ncpu += 1
        tc = (time.perf_counter() - tc0) / ncpu
        o.close()
        cores = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
        line["cpu_baseline"] = {"value": 1.0 / (tc * steps_per_solve), "unit": "solves/s", "cores": cores,
                                "kind": "port",
                                "sample": f"{ncpu} FP64 PC steps (+ renormalisation) of the same {a.n}^3 workload "
                                          f"after 3 RK steps, {tc:.3f} s/step, scaled by the GPU's "
                                          f"{steps_per_solve:.0f} steps per converged solve",
                                "host": host_info()}
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
