"""uplus (X+-He interaction, energy.f90:146-163) on the GPU against the oracle.

Four slots of one context with different X+ potentials (the reference's
default Ba_plus_2D2, Ba_plus_gs with its own cutoff, none, and the
Rb_plus_Koutselos of the reference's snapshot workflow) and impurities on
grid nodes, so that points sit exactly on the r = 2 A boundary; uplus within
1e-12 relative of the oracle (double-double series on the device, REAL*16-style
sum in the oracle), the other energies unchanged by the X+ term."""
import numpy as np
import pytest

import eff_pot_amd as ep
from eff_pot_amd.abi import make_params, make_morse, CORE_OTC, PREC_MIXED
from oracle.oracle import Oracle

pytestmark = pytest.mark.gpu

# the fourth is the X+ potential of the reference's snapshot workflow (prep.sh:5-7)
PLUS = [("Ba_plus_2D2", 1.05, 6.260015e4), ("Ba_plus_gs", 1.05, 6.836177e3), None,
        ("Rb_plus_Koutselos", 2.0, 6.1105922888e3)]
RIMP = [(0.0, 0.0, 0.5), (1.0, -0.5, 2.0), (-1.5, 0.0, 0.0), (0.5, 0.5, -1.0)]


@pytest.mark.parametrize("precision", [0, PREC_MIXED])
def test_uplus_matches_oracle(gpu_available, precision):
    print(f"[plus] 32^3 x 4 slots precision={precision}", flush=True)
    p = make_params(32, xmax=8.0, n4=100.0, core=CORE_OTC, h_core=2.377, batch=4, precision=precision)
    with ep.Context(p) as ctx:
        psis = []
        for s in range(4):
            ctx.init_fermi(0.0218, 2.22 * 100 ** (1 / 3), 0.6, 1e-6, 70 + s, slot=s)
            ctx.set_impurity(make_morse(rimp=RIMP[s]), slot=s)
            ctx.set_plus(ep.make_plus(*PLUS[s]) if PLUS[s] else None, slot=s)
            psis.append(ctx.psi(s))
        ctx.poten()
        got = ctx.energy()
        ctx.set_plus(None, slot=0)
        ctx.poten()
        off = ctx.energy()
    assert off[0]["uplus"] == 0.0
    for k in ("etot", "eHeX", "ekin4", "natoms"):
        assert off[0][k] == got[0][k]
    for s in range(4):
        o = Oracle(make_params(32, xmax=8.0, n4=100.0, core=CORE_OTC, h_core=2.377))
        o.set_psi(psis[s])
        o.set_impurity(make_morse(rimp=RIMP[s]))
        o.set_plus(ep.make_plus(*PLUS[s]) if PLUS[s] else None)
        o.poten()
        e = o.energy()
        o.close()
        print(s, got[s]["uplus"], e["uplus"])
        if PLUS[s] is None:
            assert got[s]["uplus"] == 0.0 and e["uplus"] == 0.0
        else:
            assert e["uplus"] != 0.0
            assert abs(got[s]["uplus"] - e["uplus"]) <= 1e-12 * abs(e["uplus"])
        assert abs(got[s]["etot"] - e["etot"]) <= 1e-10 * abs(e["etot"])


EXCL = [2.0, 0.0, 3.5]  # radii hitting grid distances exactly (h = 0.5 A, impurities on nodes)


def test_exclusion_matches_oracle(gpu_available):
    """z_exciplex_exclusion per slot: eHeX, uplus and E_imp against the
    oracle with the same sphere (points at r = radius stay in), and a
    negative radius refused."""
    print("[plus] exclusion sphere 32^3 x 3 slots", flush=True)
    p = make_params(32, xmax=8.0, n4=100.0, core=CORE_OTC, h_core=2.377, batch=3)
    with ep.Context(p) as ctx:
        psis = []
        for s in range(3):
            ctx.init_fermi(0.0218, 2.22 * 100 ** (1 / 3), 0.6, 1e-6, 80 + s, slot=s)
            ctx.set_impurity(make_morse(rimp=RIMP[s]), slot=s)
            ctx.set_plus(ep.make_plus(*PLUS[0]), slot=s)
            ctx.set_exclusion(EXCL[s], slot=s)
            psis.append(ctx.psi(s))
        with pytest.raises(ep.EffpotError) as err:
            ctx.set_exclusion(-1.0, slot=1)
        assert err.value.code == 1
        ctx.poten()
        got = ctx.energy()
    for s in range(3):
        o = Oracle(make_params(32, xmax=8.0, n4=100.0, core=CORE_OTC, h_core=2.377))
        o.set_psi(psis[s])
        o.set_impurity(make_morse(rimp=RIMP[s]))
        o.set_plus(ep.make_plus(*PLUS[0]))
        o.poten()
        full = o.energy()
        o.set_exclusion(EXCL[s])
        e = o.energy()
        o.close()
        if EXCL[s] > 0:
            assert e["eHeX"] != full["eHeX"] and e["uplus"] != full["uplus"]
        for k in ("eHeX", "uplus", "eimpu", "etot"):
            assert abs(got[s][k] - e[k]) <= 1e-11 * max(1.0, abs(e[k])), (s, k, got[s][k], e[k])
