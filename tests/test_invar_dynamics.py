"""Internal-state dynamics (potinvar + the invar integrators) in the oracle.

Pin: with a negligible helium density (psi = 1e-6 everywhere, real time) H(rho) is the
spin-orbit matrix alone, so invar must follow exp(-i (H - Ev0) t) invar0
through the RK-Gill starter and the Hamming PMC; the closed form comes from
LAPACK (numpy.linalg.eigh).  A wrong sign of i, a missing Ev0, a wrong
spin-orbit entry or a history index out of place fails it.  The norm of invar
stays 1 and eso follows invar."""
import numpy as np
import pytest

import eff_pot_amd as ep
from eff_pot_amd.abi import make_params, CORE_OT
from oracle import oracle as O
from test_state_impurity import POTS


@pytest.mark.parametrize("state", [ep.STATE_P, ep.STATE_D])
def test_spin_orbit_precession(state):
    n = 6 if state == ep.STATE_P else 10
    o = O.Oracle(make_params(16, xmax=6.0, n4=1e-9, core=CORE_OT, h_core=2.377))
    o.set_psi(np.full((16, 16, 16), 1e-6 + 0j))
    rng = np.random.default_rng(7 + state)
    iv0 = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    iv0 /= np.linalg.norm(iv0)
    als, ev0 = 30.0, 4.0
    imp = ep.make_state_imp(state, rimp=(0.0, 0.0, 0.0), instate=-1, als=als, invar=iv0, **POTS)
    o.set_impurity_state(imp)
    o.poten()
    o.set_real_time(True)  # the helium part stays put (unitary); invar's equation is the same either way
    o.set_state_dynamics(imp, ev0=ev0)
    dt, nrk, npc = 5e-5, 3, 200
    for _ in range(nrk):
        o.step_rk(dt)
    for _ in range(npc):
        o.step_pc(dt)
    t = (nrk + npc) * dt
    H = O.state_matrix(state, np.zeros(31), als) - ev0 * np.eye(n)
    w, V = np.linalg.eigh(H)
    want = V @ (np.exp(-1j * w * t) * (V.conj().T @ iv0))
    got = o.invar()[:n]
    assert np.max(np.abs(got - want)) < 1e-9, np.max(np.abs(got - want))
    assert abs(np.linalg.norm(got) - 1.0) < 1e-10
    if n < 10:
        assert np.max(np.abs(o.invar()[n:])) == 0.0
    o.poten()
    e = o.energy()
    assert abs(e["eso"] - O.state_eso(state, got, als)) < 1e-13
    o.close()


@pytest.mark.parametrize("state", [ep.STATE_P, ep.STATE_D])
def test_moving_state_conserves_momentum(state):
    """A P/D impurity moving in real time while its invar evolves: with soft
    potentials (-c/r^4 beyond 2 A) the total momentum, helium plus m v, is
    conserved to 1e-4 of the impulse exchanged, and |invar| stays 1.  (Steep
    fields near the impurity break it on a 0.75 A grid, as for a steep
    isotropic potential.)"""
    from test_real_time import _droplet, _p_fd
    from test_state_impurity import _series
    soft = dict(v_pi=_series(-300.0, 3, 2.0), v_sigma=_series(-600.0, 3, 2.0), v_delta=_series(-150.0, 3, 2.0))
    o, p = _droplet()
    imp = ep.make_state_imp(state, rimp=(0.0, 0.0, 7.5), instate=0, als=3.0, **soft)
    w, V = np.linalg.eigh(O.state_matrix(state, o.state_moments(imp), 3.0))
    inv = (V[:, 0] + V[:, 1]) / np.sqrt(2.0)
    imp = ep.make_state_imp(state, rimp=(0.0, 0.0, 7.5), instate=-1, als=3.0, invar=inv, **soft)
    o.set_impurity_state(imp)
    o.poten()
    o.set_state_dynamics(imp, ev0=0.0)
    o.set_real_time(True)
    o.set_impurity_motion([0.0, 0.0, -2.0], 40.0)
    m = 40.0 * 0.0207569277
    _, v0, _ = o.impurity()
    P0 = _p_fd(o, o.psi(), 0.75) + m * v0
    for _ in range(3):
        o.step_rk(5e-5)
    for _ in range(300):
        o.step_pc(5e-5)
    _, v, _ = o.impurity()
    impulse = m * (v - v0)
    dP = _p_fd(o, o.psi(), 0.75) + m * v - P0
    assert abs(impulse[2]) > 0.05
    # along the impulse to 1e-4 of it; across it the helium's own drift of
    # ~1e-5 (its seeded noise, the same with no impurity coupling) remains
    assert abs(dP[2]) <= 1e-4 * abs(impulse[2]), (dP, impulse)
    assert np.max(np.abs(dP[:2])) <= 5e-5, dP
    assert abs(np.linalg.norm(o.invar()) - 1.0) < 1e-8
    o.close()


@pytest.mark.parametrize("state,als,dt", [(ep.STATE_P, 1000.0, 5e-5), (ep.STATE_D, 600.0, 5e-5)])
def test_erriv_is_the_corrector_truncation_error(state, als, dt):
    """steppc's erriv (steppc.f90:213-218) is Milne's device of the Hamming
    method: 9/121 (p - c) equals the corrector's local truncation error
    h^5 y^(5) / 40 to leading order in h (the predictor's is 14/45 h^5 y^(5)).
    For the spin-orbit precession y' = -i (H - Ev0) y that is
    sum_k |h^5 [(-i (H - Ev0))^5 y]_k| / 40 times the reference's weight, 1/6
    (P) or 1/10 (D): with |lambda| dt ~ 0.05 the two agree to O(lambda dt) ~ 1e-3
    (measured 2e-4), while a lost 9/121, a wrong weight or the predictor's
    error in place of the corrector's are off by factors of 13, 1.7 and 12."""
    n = 6 if state == ep.STATE_P else 10
    o = O.Oracle(make_params(16, xmax=6.0, n4=1e-9, core=CORE_OT, h_core=2.377))
    o.set_psi(np.full((16, 16, 16), 1e-6 + 0j))
    rng = np.random.default_rng(7 + state)
    iv0 = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    iv0 /= np.linalg.norm(iv0)
    ev0 = 4.0
    imp = ep.make_state_imp(state, rimp=(0.0, 0.0, 0.0), instate=-1, als=als, invar=iv0, **POTS)
    o.set_impurity_state(imp)
    o.poten()
    o.set_real_time(True)
    o.set_state_dynamics(imp, ev0=ev0)
    for _ in range(3):
        o.step_rk(dt)
    for _ in range(50):
        o.step_pc(dt)
    A = -1j * (O.state_matrix(state, np.zeros(31), als) - ev0 * np.eye(n))
    assert 0.03 < np.abs(np.linalg.eigvals(A)).max() * dt < 0.06
    y5 = np.linalg.matrix_power(A, 5) @ o.invar()[:n]
    want = (1.0 / 6.0 if state == ep.STATE_P else 0.1) * np.sum(np.abs(dt ** 5 * y5)) / 40.0
    errimp, errvimp, erriv = o.step_errors()
    assert abs(erriv / want - 1.0) < 2e-3, (erriv, want)
    assert errimp == 0.0 and errvimp == 0.0  # the impurity does not move
    o.close()
