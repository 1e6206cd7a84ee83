"""Host-side logic of the C ABI that runs without a GPU: argument
validation (the reference STOPs: tstgrid.f90:8-12, DFT4He3d.f90:355-363)
and the HBM footprint used for admission."""
import ctypes as C

import eff_pot_amd as ep
from eff_pot_amd import abi


def create(p):
    h = C.c_void_p()
    return ep.lib().effpot_ctx_create(C.byref(p), 0, C.byref(h))


def test_bad_grid_rejected():
    def msg():
        return ep.lib().effpot_last_error(None).decode()
    assert create(abi.make_params(100, xmax=10.0)) == abi.ERR_GRID   # allowed by tstgrid, no FFT schedule here
    assert "nx = 100" in msg() and "supported sizes" in msg()
    assert create(abi.make_params(14, xmax=10.0)) == abi.ERR_GRID
    assert "tstgrid" in msg()
    assert create(abi.make_params(7 * 16, xmax=10.0)) == abi.ERR_GRID
    assert create(abi.make_params(32, xmax=10.0, nz=80)) == abi.ERR_GRID
    assert "nz = 80" in msg()


def test_bad_core_and_stencil_rejected():
    p = abi.make_params(32, xmax=10.0)
    p.core = 7
    assert create(p) == abi.ERR_CORE
    p = abi.make_params(32, xmax=10.0)
    p.npd = 9
    assert create(p) == abi.ERR_ARG
    p = abi.make_params(32, xmax=10.0, icon=4)
    assert create(p) == abi.ERR_ARG


def test_footprint_scales_with_batch_and_precision():
    b1 = ep.footprint_bytes(abi.make_params(96, xmax=28.8, batch=1))
    b8 = ep.footprint_bytes(abi.make_params(96, xmax=28.8, batch=8))
    assert 7.5 * b1 < b8 < 8.1 * b1
    m = ep.footprint_bytes(abi.make_params(96, xmax=28.8, batch=8, precision=abi.PREC_MIXED))
    assert m > b8
