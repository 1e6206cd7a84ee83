"""Cross-check of the oracle against routines of the reference compiled from
its own sources (oracle/ref/build_ref.sh -> oracle/_ref/libref.so): Deriv3D_p
and Deriv3D_pc (init_deriv_parallel.f90:368-642, 923-1197) with the Bickley
tables (newder.f/redef.f/mates.f), fforma (fforma.f90:11-60) and S13ADF
(s13adf.f90).  Skipped where the reference-compiled library is absent."""
import ctypes as C
import os

import numpy as np
import pytest

from eff_pot_amd.abi import make_params, ICON_PERIODIC, ICON_ZERO
from oracle.oracle import Oracle, lib

REFLIB = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle", "_ref", "libref.so")
pytestmark = pytest.mark.skipif(not os.path.exists(REFLIB), reason="reference-compiled checker not built")

_ref = None


def ref():
    global _ref
    if _ref is None:
        _ref = C.CDLL(REFLIB)
        i = C.c_int
        _ref.init_deriv_p_(C.byref(i(13)), C.byref(i(2)), C.byref(i(1)))
    return _ref


def ref_deriv(id_, iv, h, f, icon, complex_=False):
    nz, ny, nx = f.shape
    nn = (C.c_int * 3)(nx, ny, nz)
    if complex_:
        src = np.stack([f.real, f.imag], axis=-1).astype(np.float64).copy()
        out = np.zeros_like(src)
        fn = ref().deriv3d_pc_
    else:
        src = np.ascontiguousarray(f, dtype=np.float64)
        out = np.zeros_like(src)
        fn = ref().deriv3d_p_
    fn(C.byref(C.c_int(id_)), nn, C.byref(C.c_double(h)), C.byref(C.c_int(iv)),
       src.ctypes.data_as(C.c_void_p), out.ctypes.data_as(C.c_void_p), C.byref(C.c_int(icon)))
    if complex_:
        return out[..., 0] + 1j * out[..., 1]
    return out


@pytest.mark.parametrize("icon", [ICON_ZERO, ICON_PERIODIC])
def test_deriv3d_matches_reference(icon):
    p = make_params(16, xmax=4.0, ny=24, nz=32, ymax=5.0, zmax=7.0, icon=icon)
    o = Oracle(p)
    h = o.spacing()
    rng = np.random.default_rng(5)
    f = rng.standard_normal(o.shape)
    g = f + 1j * rng.standard_normal(o.shape)
    for id_ in (1, 2):
        for iv in (1, 2, 3):
            r = ref_deriv(id_, iv, h[iv - 1], f, icon)
            mine = o.derivnd(id_, iv, f)
            assert np.abs(mine - r).max() <= 1e-13 * np.abs(r).max(), (id_, iv)
            rc = ref_deriv(id_, iv, h[iv - 1], g, icon, complex_=True)
            mc = o.derivnd(id_, iv, g)
            assert np.abs(mc - rc).max() <= 1e-13 * np.abs(rc).max(), (id_, iv)


def test_s13adf_matches_reference():
    # S13ADF's large-x Chebyshev branch (x > 16) is good to ~4e-14; the oracle's
    # Si is good to 1 ulp (checked against scipy elsewhere)
    ref().s13adf_.restype = C.c_double
    for x in (0.05, 0.7, 2.0, 5.5, 15.9, 16.1, 27.0, 80.0):
        v = ref().s13adf_(C.byref(C.c_double(x)), C.byref(C.c_int(0)))
        tol = 5e-16 if x <= 16 else 5e-14
        assert abs(v - lib().oracle_si(x)) < tol * max(1.0, abs(v)), x


def test_fforma_matches_reference():
    n = 32
    p = make_params(n, xmax=8.0, h_core=2.377)
    o = Oracle(p)
    w, v, k = o.kernels()
    h = 16.0 / n
    px = np.arange(n // 2 + 1) / (n * h)
    pf = np.array([i / (n * h) if i <= n // 2 else -1 / (2 * h) + (i - n // 2) / (n * h) for i in range(n)])
    pmod = np.sqrt(px[None, None, :] ** 2 + pf[None, :, None] ** 2 + pf[:, None, None] ** 2)
    pmod = np.ascontiguousarray(pmod)
    out = np.zeros_like(pmod)
    core = C.create_string_buffer(b"OT")
    i = C.c_int
    d = C.c_double
    ref().fforma_(core, C.byref(d(-718.99)), C.byref(d(10.22)), C.byref(d(2.556)), C.byref(d(2.377)),
                  C.byref(i(n)), C.byref(i(n)), C.byref(i(n)), pmod.ctypes.data_as(C.c_void_p),
                  out.ctypes.data_as(C.c_void_p), C.c_size_t(2))
    scale = np.abs(out).max()
    # both evaluate the same closed form; its cancellation at large p
    # (sinx11 loses ~(2 pi p h)^10/10! digits) magnifies S13ADF's 4e-14 error to
    # 1.5e-4 K A^3 at |p| = 1.73/A (the oracle there is 2.4e-6 from the exact
    # value, see test_fvlj_exact_mpmath); below |p| = 0.5/A they agree to 1e-13.
    assert np.abs(v - out).max() < 2e-4
    assert np.abs(v - out)[pmod < 0.5].max() < 1e-13 * scale
    assert np.abs(v - out)[pmod < 1.0].max() < 1e-10 * scale
    assert abs(v[0, 0, 0] - out[0, 0, 0]) < 1e-12 * scale
