"""instates' other choices of invar (host only): Ldiag_jz / Ljz,
Lexcite_state_fix, Laverage_P_value, Lmixture and Lexciplex_state_fix
(instates.f90:54-113, 228-364 for P; 613-726 for D).

Pins:
* Jz = Lz + Sz: its orbital part read off the library's spin-orbit matrix
  (the up-up block of L.S is Lz/2) equals the textbook Lz = -i (x d/dy -
  y d/dx) on the reference's orbitals (P: x, y, z; D: xy, yz, z^2, xz,
  (x^2-y^2)/2), so both conventions are fixed independently of the code;
* Ldiag_jz against LAPACK (numpy.linalg.eigh) on the oracle's matrix, for
  every instate and every Ljz: the same state up to a global phase, the same
  eso; for axially symmetric moments the selected state has <Jz> = Ljz exactly
  and is an eigenvector of V + SO;
* Lexcite_state_fix: the closed-form orbitals, eso = 0 (a real orbital with
  spin up has <L.S> = 0);
* Laverage_P_value and Lmixture: unit norm and <V + SO> = the weighted mean of
  the levels (cross terms between eigenvectors vanish); Lmixture with weights
  (0, 0), (1, 0), (0, 1) is the pair 0, 1, 2 selection exactly.  Their field
  beyond these invariants depends on the eigensolver's phases (as the
  reference's does on EISPACK's): parity unpinned there;
* exciplex moments: closed forms for a 'Linear' and a 'Ring' exciplex on each
  axis.
"""
import numpy as np
import pytest

import eff_pot_amd as ep
from oracle import oracle as O

N = {ep.STATE_P: 6, ep.STATE_D: 10}
NM = {ep.STATE_P: 7, ep.STATE_D: 31}
MANIFOLD = {  # (state, instate) -> (first eigenvector, count, default 2 jz): modules.f90:310-318
    (ep.STATE_P, 0): (0, 2, -1), (ep.STATE_P, 1): (2, 4, -3), (ep.STATE_P, 2): (2, 4, -1),
    (ep.STATE_D, 0): (0, 4, -3), (ep.STATE_D, 1): (0, 4, -1), (ep.STATE_D, 2): (4, 6, -5),
    (ep.STATE_D, 3): (4, 6, -3), (ep.STATE_D, 4): (4, 6, -1)}


def _textbook_lz(state):
    """<a|Lz|b> on real orbitals written as symmetric quadratic forms
    (D) or vectors (P); Lz = -i d/dphi of a rotation about z."""
    G = np.array([[0.0, -1.0, 0.0], [1.0, 0.0, 0.0], [0.0, 0.0, 0.0]])  # d/dphi R_phi p = G p
    if state == ep.STATE_P:
        E = np.eye(3)  # x, y, z: f = e.p, d/dphi f = e.G p
        return -1j * np.array([[E[a] @ (E[b] @ G) for b in range(3)] for a in range(3)])
    def q(*pairs):
        Q = np.zeros((3, 3))
        for (i, j, c) in pairs:
            Q[i, j] += 0.5 * c
            Q[j, i] += 0.5 * c
        return Q
    forms = [q((0, 1, 1.0)), q((1, 2, 1.0)), q((2, 2, 2.0), (0, 0, -1.0), (1, 1, -1.0)), q((0, 2, 1.0)),
             q((0, 0, 0.5), (1, 1, -0.5))]
    forms = [Q / np.sqrt(np.trace(Q @ Q)) for Q in forms]  # sphere inner product ~ tr(Q1 Q2) for l = 2
    return -1j * np.array([[np.trace(A @ (B @ G + G.T @ B)) for B in forms] for A in forms])


def _jz(state, als=2.0):
    n = N[state]
    LS = O.state_matrix(state, np.zeros(NM[state]), als) / als
    lz = 2.0 * LS[0::2, 0::2]
    return np.kron(lz, np.eye(2)) + np.kron(np.eye(n // 2), np.diag([0.5, -0.5])), lz


@pytest.mark.parametrize("state", [ep.STATE_P, ep.STATE_D])
def test_jz_is_the_textbook_one(state):
    _, lz = _jz(state)
    assert np.max(np.abs(lz - _textbook_lz(state))) < 1e-14


def _lapack_jz(state, H, instate, jz2):
    k0, nj, jdef = MANIFOLD[(state, instate)]
    w, V = np.linalg.eigh(H)
    Vm = V[:, k0:k0 + nj]
    J, _ = _jz(state)
    wj, W = np.linalg.eigh(Vm.conj().T @ J @ Vm)
    il = ((jz2 or jdef) + nj - 1) // 2
    return Vm @ W[:, il], wj, w


@pytest.mark.parametrize("state", [ep.STATE_P, ep.STATE_D])
def test_jz_selection_matches_lapack(state):
    rng = np.random.default_rng(70 + state)
    n = N[state]
    done = 0
    for trial in range(40):
        als = rng.uniform(3.0, 15.0)
        m = rng.standard_normal(NM[state]) * als * rng.uniform(0.05, 1.0)
        H = O.state_matrix(state, m, als)
        for instate in range(n // 2):
            k0, nj, _ = MANIFOLD[(state, instate)]
            for jz2 in [0] + list(range(-(nj - 1), nj, 2)):
                v_l, wj, w = _lapack_jz(state, H, instate, jz2)
                gaps = np.diff(wj).min(), min(w[k0] - w[k0 - 1] if k0 else np.inf,
                                              w[k0 + nj] - w[k0 + nj - 1] if k0 + nj < n else np.inf)
                if gaps[0] < 1e-3 or gaps[1] < 1e-3 * als:
                    continue
                sel = ep.make_select(ep.SELECT_JZ, "" if jz2 == 0 else f"{jz2:+d}/2")
                wg, inv, eso = ep.instates(state, m, als, instate, sel)
                assert abs(np.vdot(inv, inv).real - 1.0) < 1e-13
                assert abs(abs(np.vdot(v_l, inv)) - 1.0) < 1e-10, (trial, instate, jz2)
                assert abs(eso - O.state_eso(state, v_l, als)) <= 1e-10 * als
                done += 1
    assert done > 100


@pytest.mark.parametrize("state", [ep.STATE_P, ep.STATE_D])
def test_jz_axial_states_have_the_asked_jz(state):
    """Moments symmetric about z: Jz commutes with V + SO, so the chosen
    state is a Jz eigenvector with eigenvalue Ljz and an eigenvector of H."""
    if state == ep.STATE_P:
        m = np.array([1.0, -40.0, -40.0, -95.0, 0.0, 0.0, 0.0])
    else:  # helium on the z axis only (two points), through the oracle's moment sums
        from test_state_impurity import POTS, _oracle
        o = _oracle()
        psi = np.zeros((16, 16, 16), dtype=complex)
        psi[8 + 5, 8, 8] = 1.0
        psi[8 - 3, 8, 8] = 0.7
        o.set_psi(psi)
        m = o.state_moments(ep.make_state_imp(ep.STATE_D, instate=0, **POTS))
        o.close()
    J, _ = _jz(state)
    for als in (4.0, 25.0):
        H = O.state_matrix(state, m, als)
        assert np.max(np.abs(H @ J - J @ H)) < 1e-10 * np.abs(H).max()
        for instate in range(N[state] // 2):
            k0, nj, jdef = MANIFOLD[(state, instate)]
            for jz2 in [0] + list(range(-(nj - 1), nj, 2)):
                sel = ep.make_select(ep.SELECT_JZ, "" if jz2 == 0 else f"{jz2:+d}/2")
                _, inv, _ = ep.instates(state, m, als, instate, sel)
                want = 0.5 * (jz2 or jdef)
                assert np.max(np.abs(J @ inv - want * inv)) < 1e-10, (als, instate, jz2)
                e = np.vdot(inv, H @ inv).real
                assert np.max(np.abs(H @ inv - e * inv)) < 1e-9 * np.abs(H).max()


def test_jz_rejects_ljz_outside_the_manifold():
    m = [1.0, 5.0, 5.0, 6.0, 0.0, 0.0, 0.0]
    for instate, ljz in ((0, "-3/2"), (1, "+5/2"), (2, "+7/2")):
        with pytest.raises(ep.EffpotError):
            ep.instates(ep.STATE_P, m, 3.0, instate, ep.make_select(ep.SELECT_JZ, ljz))
    with pytest.raises(ep.EffpotError):
        ep.instates(ep.STATE_D, np.ones(31), 3.0, 0, ep.make_select(ep.SELECT_JZ, "-5/2"))
    with pytest.raises(ep.EffpotError):
        ep.make_select(ep.SELECT_JZ, "-3/4")
    for mode in (ep.SELECT_FIX, ep.SELECT_AVERAGE, ep.SELECT_MIXTURE):  # P only
        with pytest.raises(ep.EffpotError):
            ep.instates(ep.STATE_D, np.ones(31), 3.0, 0, ep.make_select(mode))
    with pytest.raises(ep.EffpotError):
        ep.instates(ep.STATE_P, m, 3.0, 0, ep.make_select(ep.SELECT_MIXTURE, a_pi_3o2=-0.1))


def test_fixed_states():
    m = [1.0, -30.0, -20.0, -50.0, 3.0, -2.0, 1.0]
    want = {0: [1, 0, 0, 0, 0, 0], 1: [1, 0, 1, 0, 0, 0], 2: [1, 0, 1, 0, 1, 0]}
    for instate, v in want.items():
        w, inv, eso = ep.instates(ep.STATE_P, m, 7.0, instate, ep.make_select(ep.SELECT_FIX))
        v = np.array(v, float) / np.sqrt(sum(v))
        assert np.array_equal(inv, v.astype(complex))
        assert eso == 0.0
        assert np.allclose(w, np.linalg.eigvalsh(O.state_matrix(ep.STATE_P, m, 7.0)), rtol=0, atol=1e-12)


def test_average_and_mixture_invariants():
    rng = np.random.default_rng(5)
    for trial in range(10):
        als = rng.uniform(2.0, 12.0)
        m = np.concatenate([[1.0], rng.standard_normal(6) * 30.0])
        H = O.state_matrix(ep.STATE_P, m, als)
        w = np.linalg.eigvalsh(H)
        _, inv, _ = ep.instates(ep.STATE_P, m, als, 0, ep.make_select(ep.SELECT_AVERAGE))
        assert abs(np.vdot(inv, inv).real - 1.0) < 1e-13
        assert abs(np.vdot(inv, H @ inv).real - w.mean()) < 1e-11 * np.abs(w).max()
        for a, b in ((0.0, 0.0), (1.0, 0.0), (0.0, 1.0), (0.94, 0.34117444), (0.2, 0.3), (0.5, 0.5)):
            _, inv, eso = ep.instates(ep.STATE_P, m, als, 2, ep.make_select(ep.SELECT_MIXTURE, a_pi_3o2=a,
                                                                             a_sig_1o2=b))
            amp2 = np.array([max(1.0 - a - b, 0.0), a, b])
            assert abs(np.vdot(inv, inv).real - amp2.sum()) < 1e-13
            assert abs(np.vdot(inv, H @ inv).real - amp2 @ w[0::2]) < 1e-11 * np.abs(w).max() * amp2.sum()
            hot = np.flatnonzero(amp2 == 1.0)
            if len(hot):  # one pair: exactly the default selection of that pair
                _, inv_p, eso_p = ep.instates(ep.STATE_P, m, als, int(hot[0]))
                assert np.array_equal(inv, inv_p) and eso == eso_p


@pytest.mark.parametrize("axis", [0, 1, 2])
def test_exciplex_moments(axis):
    vp, vs = ep.make_plus("Ba_plus_2D1", 1.05, 5.216077e4), ep.make_plus("Ba_plus_2D0", 1.05, 9.674836e4)
    r = 3.9
    Vp, Vs = ep.plus_eval(vp, r), ep.plus_eval(vs, r)
    ax = "XYZ"[axis]
    lin = ep.exciplex_moments(vp, vs, "Linear", ax, r)
    want = np.zeros(7)
    want[1:4] = 2 * Vp
    want[1 + axis] = 2 * Vs
    assert np.allclose(lin, want, rtol=1e-13, atol=1e-12 * abs(Vs))
    ring = ep.exciplex_moments(vp, vs, "Ring", ax, r)
    want = np.zeros(7)
    want[1:4] = 6 * Vp + 3 * (Vs - Vp)
    want[1 + axis] = 6 * Vp
    assert np.allclose(ring, want, rtol=1e-13, atol=1e-12 * abs(Vs))
    # for the ring, the bond directions lie in the plane: the sigma level is
    # the in-plane one, twice degenerate
    w = np.linalg.eigvalsh(O.state_matrix(ep.STATE_P, ring, 0.0)[0::2, 0::2].real)
    assert np.allclose(np.sort(w), np.sort([6 * Vp, 6 * Vp + 3 * (Vs - Vp), 6 * Vp + 3 * (Vs - Vp)]), rtol=1e-12)
    with pytest.raises(ep.EffpotError):
        ep.exciplex_moments(vp, vs, "Square", ax, r)
