"""The FP32 kernels of the mixed-precision form, field by field, against the
FP64 oracle: the FP32 transforms (x r2c / fused plane, strided y and z, x
c2r) against numpy, and every field of the FP32 poten (k_fft_plane_r2c,
k_fft_zconv, k_grad, k_div, k_pot4 in float) against the oracle's FP64 poten
of the same density, with the defect correction zero (a fresh state: no
energy check has set it), so nothing FP64 corrects the FP32 result.

Bars (relative L2 over the grid): the FP32 transforms 1e-6 (single rounding
of FP32 arithmetic, ~6e-8, times the passes and log n); the FP32 fields 2e-6
(density rounded to FP32, 14 FFTs, stencils and products; measured 1.2e-7 to
6.2e-7 at 32^3, 96^3 and 128^3, the first-derivative fields the largest); the
impurity field uimp, an FP64 field rounded once, 1e-7."""
import numpy as np
import pytest

import eff_pot_amd as ep
from eff_pot_amd.abi import make_params, make_morse, CORE_OTC, PREC_MIXED
from oracle.oracle import Oracle

pytestmark = pytest.mark.gpu


def rel(a, b):
    a = np.asarray(a)
    a = a.astype(np.complex128 if np.iscomplexobj(a) else np.float64)
    return np.linalg.norm((a - b).ravel()) / np.linalg.norm(np.asarray(b).ravel())


@pytest.mark.parametrize("n", [16, 24, 32, 48, 64, 96, 128, 192, 256])
def test_fft_fp32_r2c_c2r(n, gpu_available):
    print(f"[fft fp32] n={n}", flush=True)
    p = make_params(n, xmax=8.0, precision=PREC_MIXED, psi_complex=0)
    rng = np.random.default_rng(100 + n)
    f = rng.standard_normal((n, n, n)).astype(np.float32)
    ref = np.fft.rfftn(f.astype(np.float64))
    with ep.Context(p) as c:
        F = c.fft_r2c_fp32(f)
        assert rel(F, ref) < 1e-6, rel(F, ref)
        g = c.fft_c2r_fp32(ref.astype(np.complex64))
        assert rel(g, f) < 1e-6, rel(g, f)


# (grid, spacing, atoms, impurity z0 / R): the cfg0-like small case, the
# 96^3 scan shape with the impurity at the surface, the 128^3 headline shape
CASES = [(32, 0.75, 100.0, 0.3), (96, 0.6, 1000.0, 1.0), (128, 0.5, 1000.0, 0.5)]
FIELDS32 = ["pot4", "dencg", "delj4", "sto1", "denalf", "ualphas", "dxden", "dyden", "dzden", "intxalf",
            "intyalf", "intzalf"]


# every case with the 256-thread z convolution and the packed half-spectrum
# layout between the plane kernels (the default); the smallest and the largest
# also with the 512-thread z convolution (32 lines per group) and with the
# padded layout (rows of nx/2+1 padded to a multiple of 16)
ZCASES = ([c + ("256", "1") for c in CASES] + [CASES[0] + ("512", "1"), CASES[2] + ("512", "1")]
          + [CASES[0] + ("256", "0"), CASES[2] + ("256", "0")])


@pytest.mark.parametrize("n,h,natoms,zr,zth,packed", ZCASES)
def test_fp32_fields_match_oracle(n, h, natoms, zr, zth, packed, gpu_available, monkeypatch):
    monkeypatch.setenv("EFFPOT_ZCONV_TH", zth)
    monkeypatch.setenv("EFFPOT_PACKED", packed)
    print(f"[fp32 fields] {n}^3 h={h} N={natoms} z convolution {zth} threads, packed {packed}", flush=True)
    xmax = n * h / 2
    R = 2.22 * natoms ** (1 / 3)
    imp = make_morse(rimp=(0.0, 0.0, zr * R))
    p = make_params(n, xmax=xmax, n4=natoms, core=CORE_OTC, h_core=2.377, precision=PREC_MIXED, psi_complex=0)
    with ep.Context(p) as c:
        assert c.zconv_threads() == int(zth)
        c.init_fermi(0.0218, R, 0.6, 1e-6, 515 + n)
        psi = c.psi()
        c.set_impurity(imp)
        c.poten_fp32()
        got = {k: c.field_fp32(k) for k in FIELDS32 + ["uimp", "corr"]}
    o = Oracle(make_params(n, xmax=xmax, n4=natoms, core=CORE_OTC, h_core=2.377))
    o.set_psi(psi)
    o.set_impurity(imp)
    o.poten()
    assert not np.any(got["corr"]), "a fresh state carries no defect correction"
    assert rel(got["uimp"], o.field("uimp")) < 1e-7
    worst = {}
    for k in FIELDS32:
        worst[k] = rel(got[k], o.field(k))
    o.close()
    print({k: f"{v:.2e}" for k, v in worst.items()}, flush=True)
    for k, v in worst.items():
        assert v < 2e-6, (k, v)


@pytest.mark.parametrize("n,h,natoms,zr", [CASES[0], CASES[2]])
def test_div_gradient_matches_stored_gradient(n, h, natoms, zr, gpu_available):
    """With FP32 fields and the plane transforms the relaxation's poten does
    not store the density gradient: k_div forms it from the density with
    k_grad's 13-point sums in k_grad's order.  Its mean field agrees with the
    one poten_fp32 makes with the gradient stored by k_grad and read back by
    k_div to FP32 rounding: the compiler contracts the sums of the two kernels
    into FMAs differently (measured 8.3e-9 relative at 32^3).  Bar 1e-7
    relative L2 and 1e-6 relative pointwise, far inside the 2e-6 the FP32
    fields keep against the FP64 oracle."""
    print(f"[fp32 div gradient] {n}^3", flush=True)
    R = 2.22 * natoms ** (1 / 3)
    imp = make_morse(rimp=(0.0, 0.0, zr * R))
    p = make_params(n, xmax=n * h / 2, n4=natoms, core=CORE_OTC, h_core=2.377, precision=PREC_MIXED,
                    psi_complex=0)
    with ep.Context(p) as c:
        assert c.plane_fused(True) != 0, "the case must run the plane transforms"
        c.init_fermi(0.0218, R, 0.6, 1e-6, 616 + n)
        c.set_impurity(imp)
        c.poten_fp32()  # gradient stored and read back
        stored = {k: c.field_fp32(k) for k in ("pot4", "ualphas", "sto1")}
        c.kernel_time(2, reps=1)  # the relaxation's poten of the same density: gradient formed in k_div
        formed = {k: c.field_fp32(k) for k in ("pot4", "ualphas", "sto1")}
    for k in stored:
        assert np.all(np.isfinite(formed[k])), k
        r = rel(formed[k], stored[k])
        m = np.max(np.abs(formed[k].astype(np.float64) - stored[k])) / np.max(np.abs(stored[k]))
        print(k, f"{r:.2e} {m:.2e}", flush=True)
        assert r < 1e-7 and m < 1e-6, (k, r, m)
