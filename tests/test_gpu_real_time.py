"""Real-time propagation and a moving impurity on the GPU against the oracle.

Two slots relax in imaginary time (3 RK-Gill + 200 PC steps), then switch to
real time; slot 0's Morse impurity moves (40 u, v = 2 A per K^-1 along z),
slot 1's stays.  After 3 RK-Gill + 100 PC real-time steps the impurity's r, v,
a, the energies (with ekinx) and psi agree with the oracle's to 1e-9."""
import numpy as np
import pytest

import eff_pot_amd as ep
from eff_pot_amd.abi import make_params, make_morse, CORE_OTC
from oracle import oracle as O

pytestmark = pytest.mark.gpu

SEEDS = (3, 4)
RIMP = ((0.0, 0.0, 1.0), (0.5, 0.0, -1.0))


def test_real_time_matches_oracle(gpu_available):
    print("[rt] 32^3 x 2 slots", flush=True)
    p = make_params(32, xmax=12.0, n4=20.0, core=CORE_OTC, h_core=2.377, batch=2, psi_complex=1)
    dt_i, dt_r = 1e-4, 5e-5
    with ep.Context(p) as ctx:
        for s in range(2):
            ctx.init_fermi(0.0218, 2.22 * 20 ** (1 / 3), 0.6, 1e-6, SEEDS[s], slot=s)
            ctx.set_impurity(make_morse(rimp=RIMP[s]), slot=s)
        ctx.poten()
        for _ in range(3):
            ctx.step_rk(dt_i)
            ctx.renormalize()
        for _ in range(200):
            ctx.step_pc(dt_i)
            ctx.renormalize()
        ctx.set_real_time(True)
        ctx.set_impurity_motion([0.0, 0.0, 2.0], 40.0, slot=0)
        for _ in range(3):
            ctx.step_rk(dt_r)
        for _ in range(100):
            ctx.step_pc(dt_r)
        got = [ctx.impurity(s) for s in range(2)]
        psis = [ctx.psi(s) for s in range(2)]
        ctx.poten()
        es = ctx.energy()
    for s in range(2):
        o = O.Oracle(make_params(32, xmax=12.0, n4=20.0, core=CORE_OTC, h_core=2.377))
        o.init_fermi(0.0218, 2.22 * 20 ** (1 / 3), 0.6, 1e-6, SEEDS[s])
        o.set_impurity(make_morse(rimp=RIMP[s]))
        o.poten()
        for _ in range(3):
            o.step_rk(dt_i)
            o.renormalize()
        for _ in range(200):
            o.step_pc(dt_i)
            o.renormalize()
        o.set_real_time(True)
        if s == 0:
            o.set_impurity_motion([0.0, 0.0, 2.0], 40.0)
        for _ in range(3):
            o.step_rk(dt_r)
        for _ in range(100):
            o.step_pc(dt_r)
        r, v, a = o.impurity()
        psi = o.psi()
        o.poten()
        e = o.energy()
        o.close()
        print(s, got[s], (r, v, a))
        for x, y in zip(got[s], (r, v, a)):
            assert np.allclose(x, y, rtol=1e-9, atol=1e-9 * max(1.0, np.abs(y).max()))
        assert np.linalg.norm(psis[s] - psi) <= 1e-9 * np.linalg.norm(psi)
        for k in ("etot", "eimpu", "ekinx", "eHeX", "natoms"):
            assert abs(es[s][k] - e[k]) <= 1e-9 * max(1.0, abs(e[k])), (s, k, es[s][k], e[k])
    assert es[0]["ekinx"] > 0.0 and es[1]["ekinx"] == 0.0
    assert got[0][0][2] != 1.0 and tuple(got[1][0]) == RIMP[1]


def test_force_filter_matches_oracle(gpu_available):
    """Lfilter_exciplex_force with the exclusion 2h: two moving impurities,
    one below and one above the centre plane, 3 RK + 30 PC real-time steps;
    r, v and a against the oracle to 1e-9."""
    print("[rt] force filter 32^3 x 2 slots", flush=True)
    p = make_params(32, xmax=12.0, n4=20.0, core=CORE_OTC, h_core=2.377, batch=2, psi_complex=1)
    rz = (-1.5, 1.5)
    dt_i, dt_r = 1e-4, 5e-5
    with ep.Context(p) as ctx:
        for s in range(2):
            ctx.init_fermi(0.0218, 2.22 * 20 ** (1 / 3), 0.6, 1e-6, SEEDS[s], slot=s)
            ctx.set_impurity(make_morse(rimp=(0.0, 0.0, rz[s])), slot=s)
            ctx.set_exclusion(1.5, slot=s)
            ctx.set_force_filter(True, slot=s)
        ctx.poten()
        for _ in range(3):
            ctx.step_rk(dt_i)
            ctx.renormalize()
        for _ in range(20):
            ctx.step_pc(dt_i)
            ctx.renormalize()
        ctx.set_real_time(True)
        for s in range(2):
            ctx.set_impurity_motion([0.0, 0.0, -rz[s]], 40.0, slot=s)
        for _ in range(3):
            ctx.step_rk(dt_r)
        for _ in range(30):
            ctx.step_pc(dt_r)
        got = [ctx.impurity(s) for s in range(2)]
    for s in range(2):
        o = O.Oracle(make_params(32, xmax=12.0, n4=20.0, core=CORE_OTC, h_core=2.377))
        o.init_fermi(0.0218, 2.22 * 20 ** (1 / 3), 0.6, 1e-6, SEEDS[s])
        o.set_impurity(make_morse(rimp=(0.0, 0.0, rz[s])))
        o.set_exclusion(1.5)
        o.set_force_filter(True)
        o.poten()
        for _ in range(3):
            o.step_rk(dt_i)
            o.renormalize()
        for _ in range(20):
            o.step_pc(dt_i)
            o.renormalize()
        o.set_real_time(True)
        o.set_impurity_motion([0.0, 0.0, -rz[s]], 40.0)
        for _ in range(3):
            o.step_rk(dt_r)
        for _ in range(30):
            o.step_pc(dt_r)
        want = o.impurity()
        o.close()
        for g, w, nm in zip(got[s], want, ("r", "v", "a")):
            assert np.max(np.abs(np.asarray(g) - w)) <= 1e-9 * max(1.0, np.max(np.abs(w))), (s, nm, g, w)
        print(s, got[s][2], want[2], flush=True)
