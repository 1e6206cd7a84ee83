"""ctypes binding of the FP64 CPU oracle (liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, never by the eff_pot_amd package.
"""
import ctypes as C
import os
import subprocess

import numpy as np

from eff_pot_amd.abi import Params, Impurity, Plus, StateImp, Energies, SolveOpts, Result, OK

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = os.path.join(_HERE, "liboracle.so")
_lib = None

_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")


def build():
    subprocess.run(["make", "-s", "-C", _HERE, "liboracle.so"], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB):
            build()
        L = C.CDLL(_LIB)
        L.oracle_create.argtypes = [C.POINTER(Params), C.POINTER(C.c_void_p)]
        L.oracle_destroy.argtypes = [C.c_void_p]
        L.oracle_size.argtypes = [C.c_void_p, C.c_int]
        L.oracle_size.restype = C.c_size_t
        L.oracle_fd_coeffs.argtypes = [C.c_int, _dp]
        for f in ("oracle_si",):
            getattr(L, f).argtypes = [C.c_double]
            getattr(L, f).restype = C.c_double
        L.oracle_bforce.argtypes = [C.c_double] * 3
        L.oracle_bforce.restype = C.c_double
        L.oracle_fvlj.argtypes = [C.c_double] * 4
        L.oracle_fvlj.restype = C.c_double
        L.oracle_wcg.argtypes = [C.c_double] * 2
        L.oracle_wcg.restype = C.c_double
        L.oracle_fft_r2c.argtypes = [C.c_void_p, _dp, _dp]
        L.oracle_fft_c2r.argtypes = [C.c_void_p, _dp, _dp]
        L.oracle_derivnd.argtypes = [C.c_void_p, C.c_int, C.c_int, _dp, _dp]
        L.oracle_derivnd_c.argtypes = [C.c_void_p, C.c_int, C.c_int, _dp, _dp]
        L.oracle_set_impurity.argtypes = [C.c_void_p, C.POINTER(Impurity)]
        L.oracle_select_pot.argtypes = [C.POINTER(Impurity), C.c_double]
        L.oracle_select_pot.restype = C.c_double
        L.oracle_plus_pot.argtypes = [C.POINTER(Plus), C.c_double]
        L.oracle_plus_pot.restype = C.c_double
        L.oracle_set_plus.argtypes = [C.c_void_p, C.POINTER(Plus)]
        L.oracle_set_exclusion.argtypes = [C.c_void_p, C.c_double]
        L.oracle_set_force_filter.argtypes = [C.c_void_p, C.c_int]
        L.oracle_set_real_time.argtypes = [C.c_void_p, C.c_int]
        L.oracle_set_impurity_motion.argtypes = [C.c_void_p, _dp, C.c_double]
        L.oracle_get_impurity.argtypes = [C.c_void_p, _dp, _dp, _dp]
        L.oracle_state_moments.argtypes = [C.c_void_p, C.POINTER(StateImp), _dp]
        L.oracle_state_matrix.argtypes = [C.c_int, _dp, C.c_double, _dp]
        L.oracle_state_eso.argtypes = [C.c_int, _dp, C.c_double]
        L.oracle_state_eso.restype = C.c_double
        L.oracle_set_impurity_state.argtypes = [C.c_void_p, C.POINTER(StateImp)]
        L.oracle_set_state_dynamics.argtypes = [C.c_void_p, C.POINTER(StateImp), C.c_double]
        L.oracle_get_invar.argtypes = [C.c_void_p, _dp]
        L.oracle_update.argtypes = [C.c_double, C.c_int, C.c_double, _dp]
        L.oracle_update.restype = C.c_double
        L.oracle_poten.argtypes = [C.c_void_p]
        L.oracle_energy.argtypes = [C.c_void_p, C.POINTER(Energies)]
        L.oracle_step_rk.argtypes = [C.c_void_p, C.c_double]
        L.oracle_step_pc.argtypes = [C.c_void_p, C.c_double, C.POINTER(C.c_double)]
        L.oracle_renormalize.argtypes = [C.c_void_p]
        L.oracle_step_errors.argtypes = [C.c_void_p, _dp]
        L.oracle_uniform.argtypes = [C.c_uint64, C.c_uint64]
        L.oracle_uniform.restype = C.c_double
        L.oracle_init_fermi.argtypes = [C.c_void_p, C.c_double, C.c_double, C.c_double,
                                        C.c_double, C.c_uint64, _dp]
        L.oracle_set_psi.argtypes = [C.c_void_p, _dp]
        L.oracle_get_psi.argtypes = [C.c_void_p, _dp]
        L.oracle_get_den.argtypes = [C.c_void_p, _dp]
        L.oracle_set_den.argtypes = [C.c_void_p, _dp]
        L.oracle_get_field.argtypes = [C.c_void_p, C.c_int, _dp]
        L.oracle_get_kernels.argtypes = [C.c_void_p, _dp, _dp, _dp]
        L.oracle_grid.argtypes = [C.c_void_p, _dp]
        L.oracle_solve.argtypes = [C.c_void_p, C.POINTER(SolveOpts), C.POINTER(Result)]
        _lib = L
    return _lib


def plus_pot(plus, r):
    """Select_Pot of an X+ potential at r, summed in quad precision."""
    return lib().oracle_plus_pot(C.byref(plus), r)


def state_matrix(state, moments, als):
    """V + (Als/2) SO of instates from the density moments (complex n x n)."""
    n = 6 if state == 1 else 10
    m = np.zeros(31)
    m[:len(moments)] = moments
    H = np.zeros(2 * n * n)
    lib().oracle_state_matrix(state, m, als, H)
    return (H[0::2] + 1j * H[1::2]).reshape(n, n)


def state_eso(state, invar, als):
    v = np.zeros(20)
    v[0:2 * len(invar):2] = np.real(invar)
    v[1:2 * len(invar):2] = np.imag(invar)
    return lib().oracle_state_eso(state, v, als)


def fd_coeffs(id_):
    c = np.zeros(13)
    lib().oracle_fd_coeffs(id_, c)
    return c


class Oracle:
    """One FP64 CPU solve; arrays are numpy in Fortran (ix fastest) order,
    returned with shape (nz, ny, nx)."""

    def __init__(self, params: Params):
        self.p = params
        self.h = C.c_void_p()
        rc = lib().oracle_create(C.byref(params), C.byref(self.h))
        if rc != OK:
            raise ValueError(f"oracle_create failed: status {rc}")
        self.shape = (params.nz, params.ny, params.nx)
        self.hshape = (params.nz, params.ny, params.nx // 2 + 1)

    def close(self):
        if self.h:
            lib().oracle_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # --- transforms ---
    def fft_r2c(self, f):
        f = np.ascontiguousarray(f, dtype=np.float64).reshape(self.shape)
        out = np.zeros(self.hshape + (2,))
        lib().oracle_fft_r2c(self.h, f.ravel(), out.reshape(-1))
        return out[..., 0] + 1j * out[..., 1]

    def fft_c2r(self, F):
        F = np.ascontiguousarray(F, dtype=np.complex128).reshape(self.hshape)
        buf = np.stack([F.real, F.imag], axis=-1).copy()
        out = np.zeros(self.shape)
        lib().oracle_fft_c2r(self.h, buf.reshape(-1), out.reshape(-1))
        return out

    def derivnd(self, id_, iv, f):
        if np.iscomplexobj(f):
            buf = np.stack([f.real, f.imag], axis=-1).astype(np.float64).copy()
            out = np.zeros_like(buf)
            lib().oracle_derivnd_c(self.h, id_, iv, buf.reshape(-1), out.reshape(-1))
            return out[..., 0] + 1j * out[..., 1]
        f = np.ascontiguousarray(f, dtype=np.float64)
        out = np.zeros_like(f)
        lib().oracle_derivnd(self.h, id_, iv, f.reshape(-1), out.reshape(-1))
        return out

    def kernels(self):
        M = int(np.prod(self.hshape))
        w, v, k = np.zeros(M), np.zeros(M), np.zeros(M)
        lib().oracle_get_kernels(self.h, w, v, k)
        return w.reshape(self.hshape), v.reshape(self.hshape), k.reshape(self.hshape)

    def spacing(self):
        g = np.zeros(3)
        lib().oracle_grid(self.h, g)
        return g

    # --- state ---
    def set_psi(self, psi):
        psi = np.asarray(psi).reshape(self.shape)
        buf = np.stack([psi.real, psi.imag], axis=-1).astype(np.float64).copy()
        lib().oracle_set_psi(self.h, buf.reshape(-1))

    def psi(self):
        out = np.zeros(self.shape + (2,))
        lib().oracle_get_psi(self.h, out.reshape(-1))
        return out[..., 0] + 1j * out[..., 1]

    def psi_raw(self):
        out = np.zeros(self.shape + (2,))
        lib().oracle_get_psi(self.h, out.reshape(-1))
        return out

    def den(self):
        out = np.zeros(self.shape)
        lib().oracle_get_den(self.h, out.reshape(-1))
        return out

    def set_den(self, d):
        d = np.ascontiguousarray(d, dtype=np.float64)
        lib().oracle_set_den(self.h, d.reshape(-1))

    def field(self, which):
        from eff_pot_amd.abi import FIELDS
        out = np.zeros(self.shape)
        rc = lib().oracle_get_field(self.h, FIELDS.get(which, which), out.reshape(-1))
        assert rc == OK
        return out

    def init_fermi(self, rho0, R, afermi, noise, seed, centre=(0.0, 0.0, 0.0)):
        lib().oracle_init_fermi(self.h, rho0, R, afermi, noise, seed,
                                np.asarray(centre, dtype=np.float64))

    def set_impurity(self, imp: Impurity):
        rc = lib().oracle_set_impurity(self.h, C.byref(imp))
        if rc != OK:
            raise ValueError(f"oracle_set_impurity failed: {rc}")

    def set_real_time(self, on=True):
        lib().oracle_set_real_time(self.h, 1 if on else 0)

    def set_impurity_motion(self, v, mass_u):
        rc = lib().oracle_set_impurity_motion(self.h, np.asarray(v, dtype=np.float64), mass_u)
        assert rc == OK

    def impurity(self):
        r, v, a = np.zeros(3), np.zeros(3), np.zeros(3)
        lib().oracle_get_impurity(self.h, r, v, a)
        return r, v, a

    def state_moments(self, imp: StateImp):
        m = np.zeros(31)
        rc = lib().oracle_state_moments(self.h, C.byref(imp), m)
        assert rc == OK
        return m

    def set_impurity_state(self, imp: StateImp):
        """potenimp for imp.invar (the caller chooses it)."""
        rc = lib().oracle_set_impurity_state(self.h, C.byref(imp))
        assert rc == OK

    def set_state_dynamics(self, imp: StateImp, ev0=0.0):
        """invar evolves from imp.invar on (potinvar + the invar integrators)."""
        rc = lib().oracle_set_state_dynamics(self.h, C.byref(imp), ev0)
        assert rc == OK

    def invar(self):
        v = np.zeros(20)
        lib().oracle_get_invar(self.h, v)
        return v[0::2] + 1j * v[1::2]

    def set_plus(self, plus):
        lib().oracle_set_plus(self.h, C.byref(plus) if plus is not None else None)

    def set_exclusion(self, r):
        """z_exciplex_exclusion: points nearer than r to the impurity leave eHeX and uplus."""
        lib().oracle_set_exclusion(self.h, r)

    def set_force_filter(self, on):
        """Lfilter_exciplex_force: F_z only beyond the exclusion in z."""
        lib().oracle_set_force_filter(self.h, 1 if on else 0)

    # --- operators ---
    def poten(self):
        lib().oracle_poten(self.h)

    def energy(self):
        e = Energies()
        lib().oracle_energy(self.h, C.byref(e))
        return e.as_dict()

    def step_rk(self, dt):
        rc = lib().oracle_step_rk(self.h, dt)
        if rc != OK:
            raise FloatingPointError(f"oracle_step_rk: status {rc}")

    def step_pc(self, dt):
        err = C.c_double()
        rc = lib().oracle_step_pc(self.h, dt, C.byref(err))
        if rc != OK:
            raise FloatingPointError(f"oracle_step_pc: status {rc}")
        return err.value

    def step_errors(self):
        """(errimp, errvimp, erriv) of the last step_pc (steppc.f90:179,193,213-220)."""
        e = np.zeros(3)
        lib().oracle_step_errors(self.h, e)
        return tuple(e)

    def renormalize(self):
        lib().oracle_renormalize(self.h)

    def solve(self, opts: SolveOpts):
        r = Result()
        rc = lib().oracle_solve(self.h, C.byref(opts), C.byref(r))
        if rc != OK:
            raise RuntimeError(f"oracle_solve: status {rc}")
        return r
