"""The C-ABI library loads and exports every entry point declared in
include/effpot.h (no compute calls: this runs without a GPU)."""
import ctypes
import os
import re

import numpy as np

import eff_pot_amd as ep

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared():
    txt = open(os.path.join(ROOT, "include", "effpot.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(effpot_[a-z0-9_]+)\s*\(", txt)))


def test_exports_all_declared_symbols():
    names = declared()
    assert "effpot_solve" in names and "effpot_poten" in names
    lib = ctypes.CDLL(ep.LIB_PATH)
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_struct_sizes_match_header():
    # 10 int32 + 16 doubles
    assert ctypes.sizeof(ep.Params) == 10 * 4 + 16 * 8
    assert ctypes.sizeof(ep.Impurity) == 8 + 9 * 8
    assert ctypes.sizeof(ep.Energies) == 18 * 8
    assert ctypes.sizeof(ep.SolveOpts) == 2 * 8 + 4 * 4 + 8 + 8


def test_footprint_is_host_only():
    p = ep.make_params(128, xmax=32.0, batch=4)
    b = ep.footprint_bytes(p)
    assert b > 4 * 128 ** 3 * 8 * 30
    p.precision = ep.PREC_MIXED
    assert ep.footprint_bytes(p) > b


def prototypes():
    """(name, number of parameters) of every function declared in effpot.h."""
    txt = open(os.path.join(ROOT, "include", "effpot.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    out = []
    for m in re.finditer(r"\b(effpot_[a-z0-9_]+)\s*\(([^;{}]*?)\)\s*;", txt, flags=re.S):
        args = m.group(2).strip()
        n = 0 if args in ("", "void") else args.count(",") + 1
        out.append((m.group(1), n))
    return out


def test_header_compiles_as_c11_and_links(tmp_path):
    """A C11 translation unit that calls every entry point compiles with
    implicit declarations as errors (so none is hidden from C) and links
    against libeffpot.so (so every one is exported).  Nothing is executed."""
    import subprocess
    protos = prototypes()
    assert len(protos) == len(declared())
    calls = "\n".join(f"  (void){name}({', '.join(['0'] * n)});" for name, n in protos)
    src = tmp_path / "abi_c11.c"
    src.write_text('#include "effpot.h"\n\nint effpot_abi_touch(void);\nint effpot_abi_touch(void) {\n'
                   f"{calls}\n  return 0;\n}}\n\nint main(void) {{ return 0; }}\n")
    exe = tmp_path / "abi_c11"
    cmd = ["gcc", "-std=c11", "-Wall", "-Werror=implicit-function-declaration", "-Werror=int-conversion",
           "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe),
           "-L", os.path.dirname(ep.LIB_PATH), "-leffpot", "-Wl,--allow-shlib-undefined"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stdout + r.stderr


def test_renormalize_noalloc_refuses_sizes_beyond_its_scratch():
    """The refusals happen on the host before any device call."""
    lib = ep.lib()
    n4 = np.full(65, 100.0)
    fake = ctypes.c_void_p(4096)  # never dereferenced: every call below is refused first
    cases = [(fake, 32 ** 3, 65, "nslots"), (fake, 256 ** 3 + 1, 1, "npoints"), (None, 32 ** 3, 1, "null"),
             (fake, 0, 1, "npoints"), (fake, 32 ** 3, 0, "nslots")]
    for psi, npts, ns, word in cases:
        rc = lib.effpot_renormalize_noalloc(psi, None, npts, ns, 1, 0.1, n4, None, None)
        assert rc == ep.ERR_ARG
        assert word in lib.effpot_last_error(None).decode()
    bad_n4 = np.array([100.0, -1.0])
    assert lib.effpot_renormalize_noalloc(fake, None, 8, 2, 1, 0.1, bad_n4, None, None) == ep.ERR_ARG
