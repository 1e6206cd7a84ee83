"""The workloads beyond the headline at their own shapes, against the FP64
oracle over bounded step counts, and each batched result against the same
solve alone:

* the 256^3 stress shape (h = 0.4 A, N = 5000, R = 37.96 A): FP64 mean field,
  energies and 3 RK + 3 PC steps at the FP64 bars (fields 1e-12, energies
  1e-11, psi 1e-11); a batch of 8 impurity positions in the mixed form,
  slot 0 against the oracle and every slot against the FP64 batch;
* the 96^3 scan shape (h = 0.6 A, N = 1000): four of the 48 positions
  z0 = k 1.6 R / 47 (k = 0, 16, 32, 44; the last, 33.2 A, lies beyond the box
  edge at 28.8 A) as one mixed batch, each against the oracle and against its
  own one-slot run;
* the serving mix: 64^3, 96^3 and 128^3 requests admitted into one server,
  each equal to the same request served alone, one against the oracle.

Mixed against FP64: an energy check recomputes pot4 from the corrected
density, so it changes the trajectory; the runs compared here check at the
same steps (the mixed cap falls on a regular check, 3 + k pener, and the FP64
run ends at the mixed run's last step, one FP64 step after its FP32 phase).
On the same schedule only the FP32 rounding of the mean field (defect-
corrected at every check) separates them: measured up to 1e-9 relative in
E_tot and 8e-9 in E_imp at 96^3, 128^3 and 256^3; bars 1e-8 and 1e-6 (plus
1e-6 K for an impurity outside the droplet).  Alone against in the mix:
1e-12 relative."""
import numpy as np
import pytest

import eff_pot_amd as ep
from eff_pot_amd import serve
from eff_pot_amd.abi import make_params, make_morse, CORE_OTC, PREC_MIXED, SolveOpts
from oracle.oracle import Oracle

pytestmark = pytest.mark.gpu

S256 = dict(n=256, xmax=51.2, n4=5000.0)
R256 = 2.22 * 5000.0 ** (1 / 3)


def rel(a, b):
    return np.linalg.norm((a - b).ravel()) / np.linalg.norm(np.asarray(b).ravel())


def close_mixed(got, want):
    assert abs(got.etot - want.etot) <= 1e-8 * abs(want.etot), (got.etot, want.etot)
    assert abs(got.eimpu - want.eimpu) <= 1e-6 * abs(want.eimpu) + 1e-6, (got.eimpu, want.eimpu)


def test_stress256_fp64_matches_oracle(gpu_available):
    print("[cfg 256] FP64 single vs oracle", flush=True)
    p = make_params(S256["n"], xmax=S256["xmax"], n4=S256["n4"], core=CORE_OTC, h_core=2.377, psi_complex=0)
    imp = make_morse()
    dt = ep.default_dt(p)
    o = Oracle(make_params(S256["n"], xmax=S256["xmax"], n4=S256["n4"], core=CORE_OTC, h_core=2.377))
    o.init_fermi(0.0218, R256, 0.6, 1e-6, 20261203)
    o.set_impurity(imp)
    o.poten()
    eo = o.energy()
    with ep.Context(p) as c:
        c.init_fermi(0.0218, R256, 0.6, 1e-6, 20261203)
        assert np.array_equal(c.psi(), o.psi())
        c.set_impurity(imp)
        c.poten()
        for nm in ("pot4", "dencg", "delj4", "sto1", "denalf", "ualphas", "uimp"):
            r = rel(c.field(nm), o.field(nm))
            assert r < 1e-12, (nm, r)
        eg = c.energy()[0]
        for k in ("etot", "ekin4", "elj4", "ecor4", "ealphas", "eHeX", "natoms", "mu"):
            assert abs(eg[k] - eo[k]) <= 1e-11 * max(1.0, abs(eo[k])), (k, eg[k], eo[k])
        print("[cfg 256] 3 RK + 3 PC", flush=True)
        for _ in range(3):
            o.step_rk(dt); o.renormalize()
            c.step_rk(dt); c.renormalize()
        for _ in range(3):
            e1 = o.step_pc(dt); o.renormalize()
            e2 = c.step_pc(dt)[0]; c.renormalize()
            assert abs(e2 - e1) <= 1e-9 * abs(e1)
        assert rel(c.psi(), o.psi()) < 1e-11
    o.close()


def test_stress256_mixed_batch8_tracks_fp64(gpu_available):
    print("[cfg 256] mixed batch of 8 vs FP64 batch and oracle", flush=True)
    zs = [0.5 * R256 * b / 7 for b in range(8)]
    # the FP32 phase ends at a regular check (step 8 = 3 + 5 pener), one FP64
    # step follows; the FP64 batch runs to that same step
    opts = SolveOpts(dt=0.0, tol=0.0, n_rk=3, pener=5, max_steps=8, polish_steps=0, tol_fp32=0.0, check_finite=1)
    res = {}
    for prec in (PREC_MIXED, 0):
        p = make_params(S256["n"], xmax=S256["xmax"], n4=S256["n4"], core=CORE_OTC, h_core=2.377, batch=8,
                        precision=prec, psi_complex=0)
        opts.dt = ep.default_dt(p)
        if prec == 0:
            opts.max_steps = res[PREC_MIXED][0].steps
        with ep.Context(p) as c:
            for b, z in enumerate(zs):
                c.init_fermi(0.0218, R256, 0.6, 1e-6, 11 + b, slot=b)
                c.set_impurity(make_morse(rimp=(0.0, 0.0, z)), slot=b)
            res[prec] = c.solve(opts)
    mixed, fp64 = res[PREC_MIXED], res[0]
    steps = mixed[0].steps
    assert all(r.steps == steps for r in mixed)
    o = Oracle(make_params(S256["n"], xmax=S256["xmax"], n4=S256["n4"], core=CORE_OTC, h_core=2.377))
    o.init_fermi(0.0218, R256, 0.6, 1e-6, 11)
    o.set_impurity(make_morse(rimp=(0.0, 0.0, zs[0])))
    opts.max_steps = steps
    ro = o.solve(opts)
    o.close()
    print("steps", steps, ro.steps, "etot", mixed[0].e.etot, ro.e.etot, "eimpu", mixed[0].e.eimpu, ro.e.eimpu,
          flush=True)
    assert ro.steps == steps
    close_mixed(mixed[0].e, ro.e)
    for b in range(8):
        print(b, mixed[b].e.eimpu, fp64[b].e.eimpu, flush=True)
        assert fp64[b].steps == steps
        close_mixed(mixed[b].e, fp64[b].e)
        assert abs(mixed[b].e.natoms - 5000.0) <= 1e-9 * 5000.0


SCAN_K = [0, 16, 32, 44]


def test_scan96_positions_match_oracle_and_alone(gpu_available):
    R = 2.22 * 1000.0 ** (1 / 3)
    zs = [k * 1.6 * R / 47 for k in SCAN_K]
    assert zs[-1] > 28.8  # beyond the box edge
    print("[cfg 96] 4 scan positions, mixed batch", flush=True)
    # cap at a regular check (3 + 2 pener): the oracle's checks fall on the same steps
    opts = SolveOpts(dt=0.0, tol=0.0, n_rk=3, pener=5, max_steps=13, polish_steps=0, tol_fp32=0.0, check_finite=1)

    def run(slots):
        p = make_params(96, xmax=28.8, n4=1000.0, core=CORE_OTC, h_core=2.377, batch=len(slots),
                        precision=PREC_MIXED, psi_complex=0)
        opts.dt = ep.default_dt(p)
        with ep.Context(p) as c:
            for b, k in enumerate(slots):
                c.init_fermi(0.0218, R, 0.6, 1e-6, 700 + k, slot=b)
                c.set_impurity(make_morse(rimp=(0.0, 0.0, k * 1.6 * R / 47)), slot=b)
            return c.solve(opts)

    mix = run(SCAN_K)
    steps = mix[0].steps
    for b, k in enumerate(SCAN_K):
        alone = run([k])[0]
        assert alone.steps == mix[b].steps
        for f in ("etot", "eimpu", "natoms"):
            x, y = getattr(alone.e, f), getattr(mix[b].e, f)
            assert abs(x - y) <= 1e-12 * abs(y) + 1e-300, (k, f, x, y)
        o = Oracle(make_params(96, xmax=28.8, n4=1000.0, core=CORE_OTC, h_core=2.377))
        o.init_fermi(0.0218, R, 0.6, 1e-6, 700 + k)
        o.set_impurity(make_morse(rimp=(0.0, 0.0, k * 1.6 * R / 47)))
        o_opts = SolveOpts(dt=opts.dt, tol=0.0, n_rk=3, pener=5, max_steps=steps, polish_steps=0, tol_fp32=0.0,
                           check_finite=1)
        ro = o.solve(o_opts)
        o.close()
        print(k, "eimpu", mix[b].e.eimpu, ro.e.eimpu, "etot", mix[b].e.etot, ro.e.etot, flush=True)
        assert ro.steps == steps
        close_mixed(mix[b].e, ro.e)


def test_served_mix_of_grid_classes(gpu_available):
    """A 64^3, a 96^3 and a 128^3 request (N = 150, 600, 2500) arriving
    together: each result equals the same request served alone, and the
    64^3 one equals the FP64 oracle run over the same steps to the mixed bars."""
    print("[cfg serve] 64/96/128 mix", flush=True)
    spec = [(150.0, 30.0, (0.0, 0.0, 2.0)), (600.0, 45.0, (1.0, 0.0, -4.0)), (2500.0, 20.0, (0.0, 3.0, 10.0))]

    def reqs():
        return [serve.Request(rid=i, natoms=n, D=d, re=4.0, a=1.0, pos=pos, arrival=0.01 * i, seed=90 + i)
                for i, (n, d, pos) in enumerate(spec)]

    kw = dict(tol=1e-7, tol_fp32=1e-6, max_steps=600, precision=PREC_MIXED)
    mix = reqs()
    srv = serve.Server(budget_bytes=24 << 30, **kw)
    try:
        srv.run(mix)
    finally:
        srv.close()
    assert [r.grid for r in mix] == [64, 96, 128]
    for r in mix:
        alone = reqs()[r.rid]
        s1 = serve.Server(budget_bytes=24 << 30, **kw)
        try:
            s1.run([alone])
        finally:
            s1.close()
        print(r.rid, r.grid, r.steps, r.eimp, alone.steps, alone.eimp, flush=True)
        assert not r.failed and r.steps == alone.steps and r.steps > 0
        assert abs(r.eimp - alone.eimp) <= 1e-12 * abs(alone.eimp)
        assert abs(r.etot - alone.etot) <= 1e-12 * abs(alone.etot)
    r = mix[0]
    o_p = serve.class_params(64, 1, 0)
    o_p.n4 = r.natoms
    o = Oracle(o_p)
    o.init_fermi(0.0218, serve.radius(r.natoms), 0.6, 1e-6, r.seed)
    o.set_impurity(make_morse(rimp=r.pos, D=30.0, re=4.0, a=1.0))
    ro = o.solve(SolveOpts(dt=r.dt, tol=0.0, n_rk=3, pener=50, max_steps=r.steps, polish_steps=0, tol_fp32=0.0,
                           check_finite=1))
    o.close()
    print("served vs oracle", r.steps, ro.steps, r.eimp, ro.e.eimpu, r.etot, ro.e.etot, flush=True)
    assert ro.steps == r.steps
    close_mixed(type("E", (), {"etot": r.etot, "eimpu": r.eimp})(), ro.e)


def test_admission_by_footprint_waits_for_budget(gpu_available):
    """Two 96^3 requests arriving together with a budget that holds one block
    of one slot: the second is admitted only when the first has finished, no
    second block is created, the reserved bytes never exceed the budget, and
    both results equal the same requests served side by side with room."""
    print("[cfg serve] admission by footprint", flush=True)
    fp = serve.footprint_bytes(serve.class_params(96, 1, PREC_MIXED))
    kw = dict(tol=1e-7, tol_fp32=1e-6, max_steps=200, precision=PREC_MIXED)

    def reqs():
        return [serve.Request(rid=i, natoms=600.0, D=30.0, re=4.0, a=1.0, pos=(0.0, 0.0, z), arrival=0.0,
                              seed=60 + i) for i, z in enumerate((1.0, -2.0))]

    tight = reqs()
    srv = serve.Server(budget_bytes=int(1.5 * fp), block_slots={96: 1}, **kw)
    try:
        srv.run(tight)
        st = srv.stats()
    finally:
        srv.close()
    assert st["peak_bytes"] == fp and st["blocks_created"][96] == 1
    first, second = sorted(tight, key=lambda r: r.start)
    assert second.start >= first.finish
    roomy = reqs()
    srv = serve.Server(budget_bytes=8 * fp, block_slots={96: 2}, **kw)
    try:
        srv.run(roomy)
    finally:
        srv.close()
    for a, b in zip(tight, roomy):
        assert not a.failed and a.steps == b.steps
        assert abs(a.eimp - b.eimp) <= 1e-12 * abs(b.eimp) and abs(a.etot - b.etot) <= 1e-12 * abs(b.etot)
