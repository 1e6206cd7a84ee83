"""P/D internal-state impurity (Lstate 'P'/'D'): instates + potenimp, host only.

Pins: the library's host instates (matrix + Jacobi + Kramers-pair selection)
against LAPACK (numpy.linalg.eigh) on the oracle's matrix; the spin-orbit
splittings of a spherical field (P: V0 - Als twice, V0 + Als/2 four times;
D: V0 - 3Als/2 four times, V0 + Als six times) and their eso; the reference's
own closed form for the P levels (the cubic of instates.f90:366-377); for a
density on one grid point, the D matrix on the z axis is diagonal
(V_delta, V_sigma, V_pi, V_sigma, V_delta) -- the reference's naming of the
d orbitals -- and its spectrum is the same in every direction; and the energy
identity eHeX + eso = the selected eigenvalue of V + SO, which ties the
oracle's potenimp field to its instates sums."""
import numpy as np
import pytest

import eff_pot_amd as ep
from eff_pot_amd.abi import make_params, CORE_OT, Plus, POT_SERIES
from oracle import oracle as O

NM = {ep.STATE_P: 7, ep.STATE_D: 31}


def _lapack_select(H, instate):
    w, V = np.linalg.eigh(H)
    return w, (V[:, 2 * instate] + V[:, 2 * instate + 1]) / np.sqrt(2.0)


def _rho(inv):
    """Spin-summed orbital density matrix (real part), the part uimp reads."""
    c = np.asarray(inv).reshape(-1, 2)
    return (np.conj(c) @ c.T).real


@pytest.mark.parametrize("state", [ep.STATE_P, ep.STATE_D])
def test_host_instates_matches_lapack(state):
    rng = np.random.default_rng(11 + state)
    n = 6 if state == ep.STATE_P else 10
    for trial in range(20):
        m = rng.standard_normal(NM[state]) * 10.0 ** rng.uniform(-1, 3)
        als = rng.uniform(-5.0, 20.0)
        H = O.state_matrix(state, m, als)
        assert np.allclose(H, H.conj().T, atol=0)
        wl = np.linalg.eigvalsh(H)
        scale = np.abs(wl).max()
        for instate in range(n // 2):
            w, inv, eso = ep.instates(state, m, als, instate)
            assert np.max(np.abs(w - wl)) <= 1e-13 * scale
            _, inv_l = _lapack_select(H, instate)
            assert abs(np.vdot(inv, inv) - 1.0) < 1e-13
            # quantities of the state that do not depend on the basis inside the Kramers pair
            assert np.max(np.abs(_rho(inv) - _rho(inv_l))) < 1e-10
            e_l = O.state_eso(state, inv_l, als)
            assert abs(eso - e_l) <= 1e-11 * max(1.0, abs(als))
            assert abs(np.vdot(inv, H @ inv).real - w[2 * instate]) <= 1e-12 * scale


def test_spin_orbit_splittings():
    for als in (1.0, 7.5, -2.0):
        w, _, e0 = ep.instates(ep.STATE_P, [1, 5, 5, 5, 0, 0, 0], als, 0)
        lo, hi = sorted([5 - als, 5 + 0.5 * als])
        assert np.allclose(w, [lo] * 2 + [hi] * 4 if als > 0 else [lo] * 4 + [hi] * 2, atol=1e-13)
        if als > 0:
            assert abs(e0 - (-als)) < 1e-13
            assert abs(ep.instates(ep.STATE_P, [1, 5, 5, 5, 0, 0, 0], als, 1)[2] - 0.5 * als) < 1e-13
        m = np.zeros(31)
        m[30] = 7.0
        w, _, _ = ep.instates(ep.STATE_D, m, als, 0)
        lo, hi = sorted([7 - 1.5 * als, 7 + als])
        assert np.allclose(w, [lo] * 4 + [hi] * 6 if als > 0 else [lo] * 6 + [hi] * 4, atol=1e-13)
        if als > 0:
            esos = [ep.instates(ep.STATE_D, m, als, k)[2] for k in range(5)]
            assert np.allclose(esos, [-1.5 * als] * 2 + [als] * 3, atol=1e-13)


def test_p_levels_solve_the_reference_cubic():
    """instates.f90:366-377: t^3 + C t^2 + B t + A with the spin-orbit terms."""
    rng = np.random.default_rng(3)
    for _ in range(10):
        Vx, Vy, Vz, Vxy, Vxz, Vyz = rng.standard_normal(6) * 20
        als = rng.uniform(0.5, 10)
        A = Vxz ** 2 * Vy - 2 * Vxy * Vxz * Vyz + Vx * Vyz ** 2 + Vxy ** 2 * Vz - Vx * Vy * Vz
        B = -Vxy ** 2 - Vxz ** 2 - Vyz ** 2 + Vx * Vy + Vx * Vz + Vy * Vz
        C = -(Vx + Vy + Vz)
        A = A + 0.25 * (als ** 3 - C * als ** 2)
        B = B - 0.75 * als ** 2
        w, _, _ = ep.instates(ep.STATE_P, [1, Vx, Vy, Vz, Vxy, Vxz, Vyz], als, 0)
        roots = np.sort(np.roots([1.0, C, B, A]).real)
        assert np.allclose(w[0::2], roots, atol=1e-9 * np.abs(roots).max())
        assert np.allclose(w[0::2], w[1::2], atol=1e-10)  # Kramers pairs


def _series(coef, k0, r0):
    """A smooth attractive inverse-power potential with a plateau inside r0."""
    p = Plus()
    p.kind, p.nterms, p.k0 = POT_SERIES, 1, k0
    p.pg[0] = coef
    p.rcutoff = p.r0 = r0
    p.fcutoff = coef / r0 ** (k0 + 1)
    return p


POTS = dict(v_pi=_series(-4000.0, 5, 1.5), v_sigma=_series(-9000.0, 5, 1.5), v_delta=_series(-1500.0, 5, 1.5))


def _oracle(n=16, xmax=4.0):
    return O.Oracle(make_params(n, xmax=xmax, n4=20.0, core=CORE_OT, h_core=2.377))


def test_point_density_d_matrix():
    o = _oracle()
    imp = ep.make_state_imp(ep.STATE_D, rimp=(0.0, 0.0, 0.0), instate=0, **POTS)
    spectra = []
    for off in ((0, 0, 5), (3, 4, 0), (0, 3, 4), (5, 0, 0)):  # r = 2.5 A in four directions
        psi = np.zeros((16, 16, 16), dtype=complex)
        psi[8 + off[2], 8 + off[1], 8 + off[0]] = 1.0
        o.set_psi(psi)
        m = o.state_moments(imp)
        H = O.state_matrix(ep.STATE_D, m, 0.0)
        M = H[0::2, 0::2].real
        assert np.allclose(H[1::2, 1::2], M) and np.allclose(H[0::2, 1::2], 0)
        if off == (0, 0, 5):
            d = np.diag(M)
            assert np.allclose(M - np.diag(d), 0, atol=1e-12 * np.abs(d).max())
            # (d1, d2, d3, d4, d5) = (xy, yz, z^2, xz, x^2-y^2): V_delta, V_sigma, V_pi, V_sigma, V_delta
            assert d[0] == d[4] and d[1] == d[3]
            vd, vs, vp = d[0], d[1], d[2]
            ref = 0.125 * np.array([ep.plus_eval(POTS[k], 2.5) for k in ("v_delta", "v_sigma", "v_pi")])
            assert np.allclose([vd, vs, vp], ref, rtol=1e-5)  # table interpolation
        spectra.append(np.linalg.eigvalsh(M))
    for s in spectra[1:]:
        assert np.allclose(s, spectra[0], rtol=1e-12, atol=1e-12 * np.abs(spectra[0]).max())
    o.close()


def test_point_density_p_matrix():
    o = _oracle()
    imp = ep.make_state_imp(ep.STATE_P, rimp=(0.0, 0.0, 0.0), instate=0, **POTS)
    for off in ((0, 0, 5), (3, 4, 0), (4, 0, 3)):
        psi = np.zeros((16, 16, 16), dtype=complex)
        psi[8 + off[2], 8 + off[1], 8 + off[0]] = 1.0
        o.set_psi(psi)
        m = o.state_moments(imp)
        M = O.state_matrix(ep.STATE_P, m, 0.0)[0::2, 0::2].real
        w = np.linalg.eigvalsh(M)
        vp = 0.125 * ep.plus_eval(POTS["v_pi"], 2.5)
        vs = 0.125 * ep.plus_eval(POTS["v_sigma"], 2.5)
        assert np.allclose(np.sort(w), np.sort([vp, vp, vs]), rtol=1e-5)
        u = np.array(off, float) / 5.0  # the sigma direction is the bond
        assert np.allclose(M @ u, w.min() * u, rtol=1e-9, atol=1e-9 * abs(w.min()))
    o.close()


@pytest.mark.parametrize("state", [ep.STATE_P, ep.STATE_D])
def test_energy_is_the_selected_level(state):
    """eHeX + eso = <invar|V + SO|invar> = the pair's eigenvalue when the
    1000 K cap never acts (attractive potentials here)."""
    o = _oracle(16, 6.0)
    o.init_fermi(0.0218, 4.0, 0.6, 1e-6, 9)
    rimp = (0.75, -0.375, 1.125)
    for instate in range(3):
        imp = ep.make_state_imp(state, rimp=rimp, instate=instate, als=3.0, **POTS)
        m = o.state_moments(imp)
        H = O.state_matrix(state, m, imp.als)
        w, inv = _lapack_select(H, instate)
        for i, c in enumerate(inv):
            imp.invar[2 * i], imp.invar[2 * i + 1] = c.real, c.imag
        o.set_impurity_state(imp)
        o.poten()
        e = o.energy()
        assert abs(e["eso"] - O.state_eso(state, inv, imp.als)) < 1e-14
        assert abs((e["eHeX"] + e["eso"]) - w[2 * instate]) <= 1e-11 * abs(w[2 * instate])
        assert e["eimpu"] == e["eHeX"] + e["eso"]
    o.close()
