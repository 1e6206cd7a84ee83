"""Wavefunction snapshot I/O (readenc mode 3 / printoutc formats), host only.

Pins: files written by the Fortran runtime itself (flang list-directed and
unformatted sequential output of tests/fortran/snapwrite.f90, record for
record as printoutc.f90:69-114) read back value for value; the psi(ix,iy,iz)
ordering checked against the closed form that wrote it; list-directed forms a
different compiler writes (repeat counts "k*(re,im)", D exponents, values split
over lines, comma separators); exact round trips of this library's writer."""
import math
import os
import subprocess

import numpy as np
import pytest

import eff_pot_amd as ep

HERE = os.path.dirname(os.path.abspath(__file__))
FLANG = "/opt/rocm/llvm/bin/flang"


def _header(nx, ny, nz, ninvar=2):
    h = ep.Snapshot()
    h.nx, h.ny, h.nz = nx, ny, nz
    h.xmax, h.ymax, h.zmax = 0.5 * nx, 0.5 * ny, 0.5 * nz
    h.hx = h.hy = h.hz = 1.0
    h.rimp[:] = [0.25, -1.0, 3.5]
    h.vimp[:] = [0.0, 0.0, 1e-3]
    h.ninvar = ninvar
    for k in range(2 * ninvar):
        h.invar[k] = 0.25 * k - 0.5
    return h


@pytest.mark.parametrize("binary", [False, True])
def test_round_trip_is_exact(tmp_path, binary):
    rng = np.random.default_rng(7)
    psi = rng.standard_normal((5, 6, 7)) * 10.0 ** rng.integers(-20, 20, (5, 6, 7)) + 1j * rng.standard_normal((5, 6, 7))
    h = _header(7, 6, 5)
    path = str(tmp_path / "snap")
    ep.write_snapshot(path, h, psi, binary=binary)
    g, p = ep.read_snapshot(path)
    assert g.binary == int(binary)
    assert (g.nx, g.ny, g.nz, g.ninvar) == (7, 6, 5, 2)
    assert list(g.rimp) == list(h.rimp) and list(g.vimp) == list(h.vimp)
    assert list(g.invar)[:4] == list(h.invar)[:4]
    assert np.array_equal(p, psi)
    g2, none = ep.read_snapshot(path, header_only=True)
    assert none is None and g2.nx == 7


@pytest.mark.skipif(not os.path.exists(FLANG), reason="flang not installed")
def test_fortran_written_files(tmp_path):
    exe = str(tmp_path / "snapwrite")
    subprocess.run([FLANG, "-O0", "-o", exe, os.path.join(HERE, "fortran", "snapwrite.f90")], check=True,
                   cwd=str(tmp_path))
    txt, binf = str(tmp_path / "s.txt"), str(tmp_path / "s.bin")
    subprocess.run([exe, txt, binf], check=True, cwd=str(tmp_path))
    ht, pt = ep.read_snapshot(txt)
    hb, pb = ep.read_snapshot(binf)
    assert ht.binary == 0 and hb.binary == 1
    for h in (ht, hb):
        assert (h.nx, h.ny, h.nz, h.ninvar) == (8, 6, 5, 3)
        assert (h.xmax, h.ymax, h.zmax) == (4.0, 3.0, 2.5)
        assert list(h.rimp) == [0.125, -1.5, 2e-3]
        assert list(h.invar)[:6] == [1.0, 0.0, 0.0, -0.5, 3.25, 1e-7]
    assert np.array_equal(pt, pb)
    for ix, iy, iz in ((1, 1, 1), (2, 3, 1), (8, 6, 5), (4, 2, 3)):
        x = math.sin(0.37 * ix + 1.3 * iy - 0.71 * iz) * math.exp(-0.05 * (ix + iy + iz))
        want = 0j if iz == 3 else complex(x, 0.1 * x * x)
        assert abs(pb[iz - 1, iy - 1, ix - 1] - want) <= 1e-15 * max(1.0, abs(want))


def test_other_list_directed_forms(tmp_path):
    """Repeat counts, D exponents, a complex split over lines, commas."""
    text = ("# first comment\n#second\n"
            "  2.0D0, 2.0D0 2.0d0 1.0 1.0 1.0 2 2 2\n"
            " 0.0 0.0 -1.5D-01\n 3*0.0\n 1\n (0.5,-0.25)\n"
            " 3*(0.0,0.0) (1.25D+00,\n -2.5E-1) 2*(7.0,8.0)\n"
            " (1,2),(3,4)\n")
    path = str(tmp_path / "ifort.txt")
    with open(path, "w") as f:
        f.write(text)
    h, p = ep.read_snapshot(path)
    assert (h.nx, h.ny, h.nz) == (2, 2, 2)
    assert list(h.rimp) == [0.0, 0.0, -0.15] and list(h.vimp) == [0.0, 0.0, 0.0]
    assert list(h.invar)[:2] == [0.5, -0.25]
    flat = p.reshape(-1)
    assert np.array_equal(flat, np.array([0, 0, 0, 1.25 - 0.25j, 7 + 8j, 7 + 8j, 1 + 2j, 3 + 4j]))


def test_truncated_and_missing_files(tmp_path):
    h = _header(4, 4, 4)
    psi = np.ones((4, 4, 4), dtype=complex)
    for binary in (False, True):
        path = str(tmp_path / ("t.bin" if binary else "t.txt"))
        ep.write_snapshot(path, h, psi, binary=binary)
        data = open(path, "rb").read()
        with open(path, "wb") as f:
            f.write(data[: len(data) - 40])
        with pytest.raises(ep.EffpotError) as e:
            ep.read_snapshot(path)
        assert e.value.code == 10
    with pytest.raises(ep.EffpotError) as e:
        ep.read_snapshot(str(tmp_path / "absent"))
    assert e.value.code == 10 and "cannot read" in str(e.value)
