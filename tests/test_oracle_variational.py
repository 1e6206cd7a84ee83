"""The mean field of the oracle is the functional derivative of its energy.

poten (poten.f90:27-126, term_alfa.f90:26-105) builds pot4 as dE/drho of the
potential part of the energy that energy (energy.f90:47-109) sums:

  E_pot = elj4 + ecor4 + ealphas + esolid

and each term separately has its own field:

  elj4    = 1/2 sum rho (V_LJ * rho) dV              ->  delj4
  ecor4   = sum rho rhob^2 (c2/2 + c3/3 rhob) dV     ->  rhob^2 (c2/2 + c3/3 rhob) + sto1
  esolid  = C sum rho (1 + tanh(beta (rho - rho_m))) ->  the solid penalty of pot4
  ealphas = -(h2o2m4 alpha_s / 2) sum G_i (F * G_i)  ->  ualphas
            G_i = (1 - rhot/rho0s) d_i rho, rhot = F * rho

On the grid these are exact identities: the 13-point centred first derivative
is an antisymmetric matrix (zero extension, icon 13, or periodic wrap, icon 8),
and every k-space kernel (V_LJ, the coarse-graining sphere, the alpha_s
Gaussian F) is real and even in p, so each convolution is a symmetric matrix.
So sum field * drho dV equals the directional derivative d/de E(rho + e drho)
to rounding, for a non-uniform droplet and any smooth drho.  The derivative is
taken by a Richardson-extrapolated central difference.  This pins the alpha_s
field (a lost 1/rho0s, a sign or a wrong kernel in term_alfa breaks it) and the
correlation convolution sto1 at non-uniform density, which the uniform-bulk
known answer cannot see (there every gradient and the alpha_s term vanish).
"""
import numpy as np
import pytest

import eff_pot_amd as ep
from oracle.oracle import Oracle

A0 = 0.5  # c2/2 and c3/3 are applied with the parameters below


def droplet(icon, core, n=32, natoms=100.0, spacing=0.75):
    p = ep.make_params(n, xmax=n * spacing / 2, n4=natoms, core=core, h_core=2.377, icon=icon)
    o = Oracle(p)
    R = 2.22 * natoms ** (1.0 / 3.0)
    o.init_fermi(0.0218, R, 0.6, 1e-6, 4242, centre=(0.7, -0.4, 1.1))  # off-grid, no mirror symmetry
    return p, o


def smooth_perturbation(rho, seed):
    """A smooth random drho: low-pass filtered noise times the density (so
    rho + e drho stays positive) plus a smooth dipole that moves the droplet."""
    rng = np.random.default_rng(seed)
    nz, ny, nx = rho.shape
    k = [np.fft.fftfreq(m) for m in (nz, ny, nx)]
    kk = k[0][:, None, None] ** 2 + k[1][None, :, None] ** 2 + k[2][None, None, :] ** 2
    noise = np.real(np.fft.ifftn(np.fft.fftn(rng.standard_normal(rho.shape)) * np.exp(-kk / (2 * 0.06 ** 2))))
    noise /= np.abs(noise).max()
    z = np.arange(nz)[:, None, None] - nz / 2
    x = np.arange(nx)[None, None, :] - nx / 2
    return rho * (0.5 * noise + 0.02 * (z + 0.5 * x))


def energy_terms(o, rho):
    o.set_den(np.ascontiguousarray(rho))
    o.poten()
    e = o.energy()
    return np.array([e["elj4"], e["ecor4"], e["ealphas"], e["esolid"]])


def term_fields(o, p, rho):
    """The fields of each energy term at rho, from the oracle's poten."""
    o.set_den(np.ascontiguousarray(rho))
    o.poten()
    delj4, dencg, sto1 = o.field("delj4"), o.field("dencg"), o.field("sto1")
    ualphas, pot4 = o.field("ualphas"), o.field("pot4")
    fcor = dencg ** 2 * (p.cp4 / 2.0 + p.cpp4 / 3.0 * dencg) + sto1
    fsol = pot4 - delj4 - fcor - ualphas
    return np.array([delj4, fcor, ualphas, fsol])


def directional_derivative(o, rho, drho, eps):
    """Richardson extrapolation of the central difference: error O(eps^4)."""
    d1 = (energy_terms(o, rho + eps * drho) - energy_terms(o, rho - eps * drho)) / (2 * eps)
    d2 = (energy_terms(o, rho + 2 * eps * drho) - energy_terms(o, rho - 2 * eps * drho)) / (4 * eps)
    return (4.0 * d1 - d2) / 3.0


@pytest.mark.parametrize("icon", [ep.ICON_ZERO, ep.ICON_PERIODIC])
@pytest.mark.parametrize("core", [ep.CORE_OT, ep.CORE_OTC])
def test_fields_are_functional_derivatives(icon, core):
    p, o = droplet(icon, core)
    rho = o.den()
    drho = smooth_perturbation(rho, seed=11 + icon + 3 * core)
    dV = float(np.prod(o.spacing()))
    fields = term_fields(o, p, rho)
    lhs = np.array([np.sum(f * drho) * dV for f in fields])
    rhs = directional_derivative(o, rho, drho, eps=1e-3)
    scale = np.abs(lhs[:3]).max()
    names = ["elj4", "ecor4", "ealphas", "esolid"]
    for k, name in enumerate(names):
        if core == ep.CORE_OT and name == "ealphas":
            assert lhs[k] == 0.0 and rhs[k] == 0.0
            continue
        # every term is compared on the scale of the largest one (the solid
        # penalty is ~1e-7 K at these densities, so it checks at rounding level)
        assert abs(lhs[k] - rhs[k]) <= 1e-9 * scale, (name, lhs[k], rhs[k])
    # the terms are not trivially small: each of the three is a sizeable share
    assert abs(lhs[0]) > 1e-2 * scale and abs(lhs[1]) > 1e-2 * scale
    if core == ep.CORE_OTC:
        assert abs(lhs[2]) > 1e-3 * scale
    o.close()


def test_alpha_s_identity_detects_planted_mistakes():
    """Rebuild ualphas from the oracle's alpha_s intermediates in numpy and
    plant the mistakes the identity must catch: a lost 1/rho0s, a flipped sign
    of the divergence part, the coarse-graining kernel in place of the
    Gaussian.  The correct rebuild matches the oracle's field; every planted
    one misses the energy derivative by far more than the identity's bar."""
    p, o = droplet(ep.ICON_ZERO, ep.CORE_OTC)
    rho = o.den()
    drho = smooth_perturbation(rho, seed=5)
    dV = float(np.prod(o.spacing()))
    o.set_den(np.ascontiguousarray(rho))
    o.poten()
    wcgk, _, kalfs = o.kernels()
    ints = [o.field(k) for k in ("intxalf", "intyalf", "intzalf")]
    grads = [o.field(k) for k in ("dxden", "dyden", "dzden")]
    falfs = 1.0 - o.field("denalf") / p.den0s
    c = p.h2o2m4 * p.alphas

    def conv(f, K):
        return o.fft_c2r(o.fft_r2c(f) * K)

    dot = sum(g * i for g, i in zip(grads, ints))
    div = sum(o.derivnd(1, a + 1, falfs * ints[a]) for a in range(3))
    good = c * (conv(dot, kalfs) / p.den0s + div)
    np.testing.assert_allclose(good, o.field("ualphas"), rtol=0, atol=1e-12 * np.abs(good).max())
    rhs = directional_derivative(o, rho, drho, eps=1e-3)[2]
    bar = 1e-9 * abs(rhs)
    assert abs(np.sum(good * drho) * dV - rhs) <= bar
    planted = {
        "lost 1/rho0s": c * (conv(dot, kalfs) + div),
        "divergence sign": c * (conv(dot, kalfs) / p.den0s - div),
        "sphere kernel": c * (conv(dot, wcgk) / p.den0s + div),
    }
    for name, wrong in planted.items():
        miss = abs(np.sum(wrong * drho) * dV - rhs)
        assert miss > 1e4 * bar, (name, miss, bar)
    o.close()


def test_correlation_identity_detects_planted_mistakes():
    """The same for the correlation field: dropping the convolution sto1 or
    convolving with the LJ kernel instead of the sphere misses d ecor4/de."""
    p, o = droplet(ep.ICON_ZERO, ep.CORE_OT)
    rho = o.den()
    drho = smooth_perturbation(rho, seed=9)
    dV = float(np.prod(o.spacing()))
    o.set_den(np.ascontiguousarray(rho))
    o.poten()
    wcgk, fvlj, _ = o.kernels()
    dencg = o.field("dencg")
    local = dencg ** 2 * (p.cp4 / 2.0 + p.cpp4 / 3.0 * dencg)
    src = rho * dencg * (p.cp4 + p.cpp4 * dencg)
    sto1 = o.fft_c2r(o.fft_r2c(src) * wcgk)
    np.testing.assert_allclose(sto1, o.field("sto1"), rtol=0, atol=1e-12 * np.abs(sto1).max())
    rhs = directional_derivative(o, rho, drho, eps=1e-3)[1]
    bar = 1e-9 * abs(rhs)
    assert abs(np.sum((local + sto1) * drho) * dV - rhs) <= bar
    for name, wrong in {"no sto1": local, "LJ kernel": local + o.fft_c2r(o.fft_r2c(src) * fvlj)}.items():
        assert abs(np.sum(wrong * drho) * dV - rhs) > 1e4 * bar, name
    o.close()
