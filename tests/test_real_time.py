"""Real-time propagation with a moving impurity in the oracle (host only).

Pins: a free impurity ('null' potential, no force) moves on a straight line
at constant velocity through the RK-Gill starter and the Hamming
predictor-modifier-corrector exactly, so every history index of the r/v
integrators is right; in real time the norm stays put (imaginary time would
drain it); with a smooth uncapped impurity potential the total momentum,
helium sum Im(psi* D psi) dV (D the 13-point first derivative) plus m v of the
impurity, is conserved to 2e-3 of the impulse exchanged, which fixes the sign
of the force, the mass units (mp_u) and the direction of time (i H)."""
import numpy as np
import pytest

from eff_pot_amd.abi import make_params, make_morse, CORE_OT, POT_NULL
from oracle import oracle as O

MP_U = 0.0207569277


def _droplet(n=32, xmax=12.0, D=5.0, a=0.5, seed=3, relax=200):
    p = make_params(n, xmax=xmax, n4=20.0, core=CORE_OT, h_core=2.377)
    o = O.Oracle(p)
    o.init_fermi(0.0218, 2.22 * 20 ** (1 / 3), 0.6, 1e-6, seed)
    imp = make_morse(rimp=(0.0, 0.0, 1.0), D=D, a=a)
    if D == 0.0:
        imp.kind = POT_NULL
    o.set_impurity(imp)
    o.poten()
    for _ in range(3):
        o.step_rk(1e-4)
        o.renormalize()
    for _ in range(relax):
        o.step_pc(1e-4)
        o.renormalize()
    return o, p


def _p_fd(o, psi, h):
    out = []
    for iv in (1, 2, 3):
        dr = o.derivnd(1, iv, np.ascontiguousarray(psi.real))
        di = o.derivnd(1, iv, np.ascontiguousarray(psi.imag))
        out.append(np.sum(psi.real * di - psi.imag * dr) * h ** 3)
    return np.array(out)


def test_free_impurity_moves_uniformly():
    o, p = _droplet(n=16, xmax=6.0, D=0.0, relax=5)
    o.set_real_time(True)
    v0 = np.array([0.7, -1.3, 2.1])
    o.set_impurity_motion(v0, 40.0)
    dt = 1e-4
    for _ in range(3):
        o.step_rk(dt)
    for _ in range(40):
        o.step_pc(dt)
    r, v, a = o.impurity()
    t = 43 * dt
    assert np.array_equal(a, np.zeros(3))
    assert np.allclose(v, v0, rtol=0, atol=1e-14)
    assert np.allclose(r, np.array([0.0, 0.0, 1.0]) + v0 * t, rtol=0, atol=1e-13)
    e = o.energy()
    assert abs(e["ekinx"] - 0.5 * 40.0 * MP_U * np.dot(v0, v0)) < 1e-13
    o.close()


def test_real_time_conserves_norm_and_momentum():
    o, p = _droplet()
    h = 2 * 12.0 / 32
    o.set_real_time(True)
    m = 40.0 * MP_U
    o.set_impurity_motion([0.0, 0.0, 2.0], 40.0)
    _, v0, a0 = o.impurity()
    assert a0[2] < 0.0  # pulled back towards the droplet centre
    P0 = _p_fd(o, o.psi(), h) + m * v0
    n0 = o.energy()["natoms"]
    dt = 5e-5
    for _ in range(3):
        o.step_rk(dt)
    for _ in range(300):
        o.step_pc(dt)
    _, v, _ = o.impurity()
    impulse = m * (v - v0)
    assert abs(impulse[2]) > 0.02
    dP = _p_fd(o, o.psi(), h) + m * v - P0
    assert np.max(np.abs(dP)) <= 2e-3 * abs(impulse[2]), (dP, impulse)
    assert abs(o.energy()["natoms"] - n0) <= 1e-7 * n0
    o.close()


@pytest.mark.parametrize("rz", [-1.5, 1.5])
def test_exciplex_force_filter(rz):
    """Lfilter_exciplex_force (potenimp.f90:216-245): with the impurity on a
    grid plane and the exclusion 2h, F_z sums exactly the planes z >= rz +
    2h (rz < 0) or z <= rz - 2h (rz > 0); F_x and F_y keep every plane."""
    o, p = _droplet(relax=10)
    h = 0.75
    o.set_impurity(make_morse(rimp=(0.0, 0.0, rz), D=5.0, a=0.5))
    o.poten()
    o.set_impurity_motion([0.0, 0.0, 0.0], 40.0)
    _, _, a_all = o.impurity()
    o.set_exclusion(2 * h)
    o.set_force_filter(True)
    o.set_impurity_motion([0.0, 0.0, 0.0], 40.0)
    _, _, a_f = o.impurity()
    z = -12.0 + h * np.arange(32)
    keep = z >= rz + 2 * h if rz < 0 else z <= rz - 2 * h
    assert 0 < keep.sum() < 32
    f = o.field("dzden") * o.field("uimp")
    want = -np.sum(f[keep]) * h ** 3 / (40.0 * MP_U)
    assert abs(a_f[2] - want) <= 1e-12 * max(abs(want), abs(a_all[2]))
    assert a_f[0] == a_all[0] and a_f[1] == a_all[1]
    o.set_force_filter(False)
    o.set_impurity_motion([0.0, 0.0, 0.0], 40.0)
    assert o.impurity()[2][2] == a_all[2]
    o.close()
