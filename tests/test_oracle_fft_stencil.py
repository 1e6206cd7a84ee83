"""Pins of the oracle's FFT and 13-point stencils (CPU only).

FFT (fft.f90 semantics: r2c unnormalised, c2r scaled by 1/N): numpy.fft on
seeded inputs, Parseval, delta -> constant spectrum, round trip.
Stencils (derivnd, init_deriv_parallel.f90:368-642 with redef.f:133-170):
centred coefficients equal the reference's own Bickley table rows
(tests/golden/fd13_coeffs.json, from mates.f), polynomials of degree <= 12
differentiate exactly away from the boundary, icon 13 equals brute-force
zero padding and icon 8 equals brute-force periodic wrapping.
"""
import json
import os

import numpy as np
import pytest

from eff_pot_amd.abi import make_params, ICON_PERIODIC, ICON_ZERO
from oracle.oracle import Oracle, fd_coeffs

GOLD = os.path.join(os.path.dirname(__file__), "golden")


@pytest.mark.parametrize("n", [16, 24, 32, 48, 64])
def test_fft_vs_numpy(n):
    o = Oracle(make_params(n, xmax=5.0))
    rng = np.random.default_rng(100 + n)
    f = rng.standard_normal(o.shape)
    F = o.fft_r2c(f)
    ref = np.fft.rfftn(f)
    assert np.abs(F - ref).max() <= 1e-13 * np.abs(ref).max()
    g = o.fft_c2r(ref)
    assert np.abs(g - f).max() < 1e-13


def test_fft_nonsquare_and_properties():
    p = make_params(32, xmax=5.0, ny=24, nz=48, ymax=4.0, zmax=7.0)
    o = Oracle(p)
    rng = np.random.default_rng(3)
    f = rng.standard_normal(o.shape)
    F = o.fft_r2c(f)
    assert np.allclose(F, np.fft.rfftn(f), rtol=0, atol=1e-12)
    # Parseval over the half spectrum (weights 2 except the self-conjugate planes)
    w = np.full(F.shape, 2.0)
    w[..., 0] = 1.0
    w[..., -1] = 1.0
    assert np.isclose((w * np.abs(F) ** 2).sum() / f.size, (f ** 2).sum(), rtol=1e-12)
    d = np.zeros(o.shape)
    d[0, 0, 0] = 1.0
    assert np.allclose(o.fft_r2c(d), 1.0, atol=1e-15)


def test_coefficients_match_reference_table():
    ref = json.load(open(os.path.join(GOLD, "fd13_coeffs.json")))
    for k in (1, 2):
        c = fd_coeffs(k)
        r = np.array(ref["K%d" % k])
        # one correctly rounded division each: equal to the decimal table to 1 ulp
        assert np.allclose(c, r, rtol=2.3e-16, atol=0), (k, c - r)


def test_coefficient_closed_forms():
    c1, c2 = fd_coeffs(1), fd_coeffs(2)
    m = np.arange(-6, 7)
    # exact on polynomials: sum c_m m^j = j! delta_{j,id}
    for j in range(13):
        assert abs((c1 * m.astype(float) ** j).sum() - (1.0 if j == 1 else 0.0)) < 1e-9
        assert abs((c2 * m.astype(float) ** j).sum() - (2.0 if j == 2 else 0.0)) < 1e-9
    assert np.allclose(c1, -c1[::-1]) and np.allclose(c2, c2[::-1])


def _brute(f, id_, axis, h, periodic):
    c = fd_coeffs(id_) / h ** id_
    n = f.shape[axis]
    out = np.zeros_like(f)
    for mi, m in enumerate(range(-6, 7)):
        if periodic:
            sh = np.roll(f, -m, axis=axis)
        else:
            pad = [(0, 0)] * 3
            pad[axis] = (6, 6)
            fp = np.pad(f, pad)
            sl = [slice(None)] * 3
            sl[axis] = slice(6 + m, 6 + m + n)
            sh = fp[tuple(sl)]
        out += c[mi] * sh
    return out


@pytest.mark.parametrize("icon", [ICON_ZERO, ICON_PERIODIC])
def test_derivnd_brute_force(icon):
    p = make_params(16, xmax=4.0, ny=24, nz=32, ymax=5.0, zmax=7.0, icon=icon)
    o = Oracle(p)
    h = o.spacing()
    rng = np.random.default_rng(11)
    f = rng.standard_normal(o.shape)
    for id_ in (1, 2):
        for iv in (1, 2, 3):
            axis = 3 - iv  # array is (z, y, x)
            got = o.derivnd(id_, iv, f)
            ref = _brute(f, id_, axis, h[iv - 1], icon == ICON_PERIODIC)
            assert np.abs(got - ref).max() < 1e-11 * np.abs(ref).max()
    # complex field: real and imaginary parts independently
    g = f + 1j * rng.standard_normal(o.shape)
    d = o.derivnd(2, 3, g)
    assert np.allclose(d.real, o.derivnd(2, 3, g.real), atol=1e-12)
    assert np.allclose(d.imag, o.derivnd(2, 3, g.imag), atol=1e-12)


def test_derivnd_polynomial_exact_interior():
    p = make_params(32, xmax=4.0)
    o = Oracle(p)
    h = o.spacing()[0]
    x = -4.0 + h * np.arange(32)
    X = np.broadcast_to(x[None, None, :], o.shape)
    f = X ** 5 - 3 * X ** 2 + 1
    d1 = o.derivnd(1, 1, f)
    d2 = o.derivnd(2, 1, f)
    inner = slice(6, 26)
    assert np.allclose(d1[..., inner], (5 * X ** 4 - 6 * X)[..., inner], rtol=1e-9, atol=1e-9)
    assert np.allclose(d2[..., inner], (20 * X ** 3 - 6)[..., inner], rtol=1e-9, atol=1e-8)
