"""Pins of the oracle's k-space kernels, impurity field, energies and
integrators (CPU only).

- Si(x) against scipy.special.sici; the OT-screened LJ transform fvlj4
  (fforma.f90:40-60) against scipy quadrature of (2/p) int_h^inf r V_LJ(r)
  sin(2 pi p r) dr, and its p=0 limit against b4=-718.99 K A^3 (modules.f90:195);
  wcgk (initcg.f90) and kalfs closed forms.
- Uniform bulk at rho0=0.021836 with h=2.190323: energy per atom
  1/2 b rho + 1/2 c2 rho^2 + c3 rho^3/3 = -7.150 K and pot4 = mu.
- Gaussian psi: kinetic energy 3 h2o2m4/(2 s^2) per atom.
- Impurity field: parabolic table interpolation (potenimp.f90:588-632) against
  the direct Morse potential, the 1000 K cap; E_imp of a Gaussian density
  against a radial quadrature.
- RK-Gill and Hamming PC on a linear plane-wave eigenmode: 4th- / 5th-order
  convergence to exp(lambda t) (a wrong coefficient breaks the order).
"""
import numpy as np
import pytest
import scipy.integrate as si
import scipy.special as sp

from eff_pot_amd.abi import (make_params, make_morse, Impurity, CORE_OT, CORE_OTC, ICON_PERIODIC,
                             H4OT, POT_MORSE)
from oracle.oracle import Oracle, lib, fd_coeffs


@pytest.mark.parametrize("x", [1e-3, 0.3, 1.0, 2.5, 3.99, 4.01, 7.3, 15.0, 25.9, 60.0, 300.0])
def test_si(x):
    assert abs(lib().oracle_si(x) - sp.sici(x)[0]) < 3e-16 * max(1.0, abs(sp.sici(x)[0]))
    assert lib().oracle_si(-x) == -lib().oracle_si(x)


def _vlj(r, eps, sig):
    return 4 * eps * ((sig / r) ** 12 - (sig / r) ** 6)


@pytest.mark.parametrize("p", [0.02, 0.1, 0.25, 0.5, 0.9])
def test_fvlj_quadrature(p):
    eps, sig, hc = 10.22, 2.556, 2.190323
    got = lib().oracle_fvlj(p, eps, sig, hc)
    k = 2 * np.pi * p
    val, err = si.quad(lambda r: r * _vlj(r, eps, sig), hc, np.inf, weight="sin", wvar=k, limlst=200)
    ref = 2.0 / p * val
    assert abs(got - ref) < 1e-7 * abs(lib().oracle_bforce(eps, sig, hc)), (got, ref)


def test_fvlj_exact_mpmath():
    mp = pytest.importorskip("mpmath")
    mp.mp.dps = 40
    eps, sig, hc = 10.22, 2.556, 2.377
    for p, tol in ((0.3, 1e-10), (1.0, 5e-8), (1.7320508075688772, 5e-6)):
        a = 2 * mp.pi * mp.mpf(p)

        def I(n):
            return a ** (n - 1) * mp.quadosc(lambda u: mp.sin(u) / u ** n, [a * hc, mp.inf], omega=1)
        s6 = mp.mpf(sig) ** 6
        exact = float(8 * eps * s6 * (s6 * I(11) - I(5)) / p)
        assert abs(lib().oracle_fvlj(p, eps, sig, hc) - exact) < tol, p


def test_bforce_limit():
    b = lib().oracle_bforce(10.22, 2.556, H4OT)
    assert abs(b - (-718.99)) < 5e-3
    # small-k expansion (k = 2 pi p): b - k^2/6 int_h^inf 4 pi r^4 V dr plus the
    # non-analytic |k|^3 term of the -4 eps sigma^6 / r^6 tail, -(pi^2/12) 4 eps sigma^6 k^3
    eps, sig, h = 10.22, 2.556, H4OT
    m4 = 4 * np.pi * 4 * eps * (sig ** 12 / (7 * h ** 7) - sig ** 6 / h)
    for pp in (1e-4, 3e-4, 1e-3):
        k = 2 * np.pi * pp
        approx = b - k ** 2 / 6 * m4 - np.pi ** 2 / 12 * 4 * eps * sig ** 6 * k ** 3
        assert abs(lib().oracle_fvlj(pp, eps, sig, h) - approx) < 2e-5


def test_wcg_kalfs_closed_forms():
    o = Oracle(make_params(32, xmax=8.0, h_core=2.377))
    w, v, k = o.kernels()
    assert w[0, 0, 0] == 1.0 and k[0, 0, 0] == 1.0
    h = 16.0 / 32
    px = np.arange(17) / (32 * h)
    a = 2 * np.pi * px[1:] * 2.377
    assert np.allclose(w[0, 0, 1:], 3 * (np.sin(a) - a * np.cos(a)) / a ** 3, rtol=1e-13, atol=1e-15)
    assert np.allclose(k[0, 0, :], np.exp(-(np.pi * 1.0 * px) ** 2), rtol=1e-14)


def test_bulk_known_answer():
    rho = 0.021836
    p = make_params(16, xmax=6.0, core=CORE_OT, h_core=H4OT, icon=ICON_PERIODIC, n4=1.0)
    o = Oracle(p)
    o.set_psi(np.full(o.shape, np.sqrt(rho), dtype=complex))
    o.set_impurity(make_morse(D=0.0))
    o.poten()
    e = o.energy()
    b = lib().oracle_bforce(10.22, 2.556, H4OT)
    c2, c3 = -2.41186e4, 1.85850e6
    epn = 0.5 * b * rho + 0.5 * c2 * rho ** 2 + c3 * rho ** 3 / 3
    assert abs(e["etot4"] / e["natoms"] - epn) < 1e-10
    assert abs(epn - (-7.150)) < 1e-3
    mu = b * rho + 1.5 * c2 * rho ** 2 + 4.0 / 3.0 * c3 * rho ** 3
    assert np.allclose(o.field("pot4"), mu, atol=1e-6)
    assert abs(e["mu"] - mu) < 1e-6
    assert abs(e["ekin4"]) < 1e-20 and abs(e["ealphas"]) == 0.0


def test_gaussian_kinetic_energy():
    n, L, s = 48, 10.0, 1.6
    p = make_params(n, xmax=L, core=CORE_OT, n4=1.0, icon=ICON_PERIODIC)
    o = Oracle(p)
    h = 2 * L / n
    x = -L + h * np.arange(n)
    r2 = x[None, None, :] ** 2 + x[None, :, None] ** 2 + x[:, None, None] ** 2
    psi = np.exp(-r2 / (2 * s * s)).astype(complex)
    o.set_psi(psi)
    o.set_impurity(make_morse(D=0.0))
    o.poten()
    e = o.energy()
    assert abs(e["ekin4"] / e["natoms"] - 3 * 6.05969638298 / (2 * s * s)) < 1e-6


def test_impurity_field_interpolation_and_cap():
    p = make_params(32, xmax=8.0)
    o = Oracle(p)
    imp = make_morse(rimp=(0.3, -0.2, 1.1))
    o.set_impurity(imp)
    u = o.field("uimp")
    h = 0.5
    x = -8.0 + h * np.arange(32)
    r = np.sqrt((x[None, None, :] - 0.3) ** 2 + (x[None, :, None] + 0.2) ** 2 + (x[:, None, None] - 1.1) ** 2)
    e = np.exp(-1.0 * (r - 4.0))
    v = 30.0 * (e * e - 2 * e)
    direct = np.minimum(v, 1000.0)
    assert u.max() <= 1000.0
    assert np.all(u[v > 1001.0] == 1000.0)
    mask = v < 900.0
    # parabolic interpolation on the table spacing D: |err| <= D^3/(9 sqrt 3) max|V'''|
    D = 4 * 8.0 / (100 * 32)
    v3 = np.abs(30.0 * (-8 * e * e + 2 * e))   # |V'''| for a = 1
    bound = D ** 3 / (9 * np.sqrt(3)) * v3 * np.exp(2 * D) * 1.05 + 5e-8
    assert np.all(np.abs(u[mask] - direct[mask]) <= bound[mask])
    assert abs(u.min() - (-30.0)) < 0.05


def test_eimp_gaussian_density_quadrature():
    n, L, w = 64, 9.6, 2.5
    p = make_params(n, xmax=L, core=CORE_OT, n4=1.0)
    o = Oracle(p)
    h = 2 * L / n
    x = -L + h * np.arange(n)
    r2 = x[None, None, :] ** 2 + x[None, :, None] ** 2 + x[:, None, None] ** 2
    A = 0.03
    o.set_psi(np.sqrt(A * np.exp(-r2 / w ** 2)).astype(complex))
    o.set_impurity(make_morse())
    o.poten()
    e = o.energy()

    def u(r):
        ee = np.exp(-(r - 4.0))
        return min(30.0 * (ee * ee - 2 * ee), 1000.0)
    ref = si.quad(lambda r: 4 * np.pi * r * r * A * np.exp(-r * r / w ** 2) * u(r), 0, 12, points=[1.3, 4.0],
                  limit=200)[0]
    assert abs(e["eHeX"] - ref) < 2e-3 * abs(ref)


def _linear_setup(n=16, V=3.0):
    """pot4 = 0 (eps=c2=c3=alphas=C=0) and uimp = V everywhere (r < r_cutoff -> umax):
    H psi = h2o2m4 lap psi - V psi is linear."""
    p = make_params(n, xmax=4.0, core=CORE_OT, icon=ICON_PERIODIC, eps4=0.0, cp4=0.0, cpp4=0.0,
                    alphas=0.0, C=0.0, n4=1.0)
    o = Oracle(p)
    imp = Impurity(kind=POT_MORSE, D=0.0, re=1.0, a=1.0, r_cutoff=1e9, umax=V, cap=1e9)
    o.set_impurity(imp)
    h = 8.0 / n
    x = -4.0 + h * np.arange(n)
    kx = 2 * np.pi * 2 / (n * h)
    psi0 = np.cos(kx * x)[None, None, :] * np.ones((n, n, 1))
    c2 = fd_coeffs(2)
    sym = (c2 * np.cos(kx * h * np.arange(-6, 7))).sum() / h ** 2
    lam = 6.05969638298 * sym - V
    return o, psi0.astype(complex), lam


@pytest.mark.parametrize("kind", ["rk", "pc"])
def test_integrator_order(kind):
    errs = []
    for dt in (4e-3, 2e-3):
        o, psi0, lam = _linear_setup()
        o.set_psi(psi0)
        o.poten()
        T = 0.12
        nst = int(round(T / dt))
        if kind == "rk":
            for _ in range(nst):
                o.step_rk(dt)
        else:
            for _ in range(3):
                o.step_rk(dt)
            for _ in range(nst - 3):
                o.step_pc(dt)
        got = o.psi()
        errs.append(np.abs(got - np.exp(lam * T) * psi0).max())
    ratio = errs[0] / errs[1]
    # RK-Gill is 4th order (ratio 16); Hamming's predictor-modifier-corrector
    # with the 112/121 modifier and 9/121 final correction extrapolates away
    # the leading local error: 5th order (ratio 32) on this linear problem.
    lo, hi = (14.0, 19.0) if kind == "rk" else (28.0, 38.0)
    assert lo < ratio < hi, (errs, ratio)


def test_rk_gill_single_step_amplification():
    o, psi0, lam = _linear_setup()
    o.set_psi(psi0)
    o.poten()
    dt = 0.01
    o.step_rk(dt)
    z = lam * dt
    g = 1 + z + z * z / 2 + z ** 3 / 6 + z ** 4 / 24
    assert np.allclose(o.psi(), g * psi0, rtol=1e-12, atol=1e-14)


def test_residual_vanishes_for_an_eigenstate():
    """The residual |H psi - mu psi| / (|mu| |psi|) of energy() is 0 for an
    exact eigenstate: a uniform bulk density on a periodic grid (OT, so every
    gradient term vanishes) with a constant impurity field (a Morse well
    capped below its own minimum: uimp = cap everywhere), mu = pot4 + cap."""
    rho = 0.021836
    p = make_params(16, xmax=6.0, core=CORE_OT, h_core=H4OT, icon=ICON_PERIODIC, n4=1.0)
    o = Oracle(p)
    o.set_psi(np.full(o.shape, np.sqrt(rho), dtype=complex))
    o.set_impurity(make_morse(D=30.0, cap=-40.0))  # V >= -30 K everywhere: uimp = -40 K
    assert np.all(o.field("uimp") == -40.0)
    o.poten()
    e = o.energy()
    assert abs(e["mu"] - (o.field("pot4")[0, 0, 0] - 40.0)) < 1e-9
    assert e["resid"] < 1e-12
    o.close()


def test_residual_of_two_plane_waves():
    """In the linear setting (pot4 = 0, uimp = V) plane waves are exact
    eigenmodes of the 13-point Laplacian with E_j = -h2o2m4 sigma(k_j) + V,
    sigma the stencil's symbol; for psi = a cos(k1 x) + b cos(k2 x):
    mu = (a^2 E1 + b^2 E2) / (a^2 + b^2) and
    resid = sqrt((a^2 (E1 - mu)^2 + b^2 (E2 - mu)^2) / (a^2 + b^2)) / |mu|.
    One mode alone gives 0.  A kinetic sign slip, a lost V or a missing dV
    weight moves it."""
    o, psi1, lam1 = _linear_setup(V=3.0)
    n, h = 16, 0.5
    x = -4.0 + h * np.arange(n)
    c2 = fd_coeffs(2)
    k2 = 2 * np.pi * 5 / (n * h)
    sym2 = (c2 * np.cos(k2 * h * np.arange(-6, 7))).sum() / h ** 2
    e1, e2 = -lam1, -(6.05969638298 * sym2 - 3.0)  # lam = h2o2m4 sigma - V; E = -lam
    o.set_psi(psi1)
    o.poten()
    assert o.energy()["resid"] < 1e-13
    a, b = 0.8, 0.35
    psi = (a * psi1.real + b * np.cos(k2 * x)[None, None, :] * np.ones((n, n, 1))).astype(complex)
    o.set_psi(psi)
    o.poten()
    e = o.energy()
    mu = (a * a * e1 + b * b * e2) / (a * a + b * b)
    want = np.sqrt((a * a * (e1 - mu) ** 2 + b * b * (e2 - mu) ** 2) / (a * a + b * b)) / abs(mu)
    assert abs(e["mu"] - mu) < 1e-11 * abs(mu)
    assert abs(e["resid"] - want) < 1e-10 * want, (e["resid"], want)
    o.close()
