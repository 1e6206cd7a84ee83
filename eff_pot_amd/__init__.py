"""eff_pot_amd: MI355X (gfx950) Orsay-Trento imaginary-time solver.

Python host binding of libeffpot.so (HIP kernels + C ABI, include/effpot.h).
The operators mirror the reference's zero-argument Fortran routines on module
state (poten.f90 `poten()`, energy.f90 `energy()`, steprk.f90 `steprk(dt)`,
steppc.f90 `steppc(dt, errHe, ...)`, init_deriv_parallel.f90 `derivnd`), as
methods of a context that owns the device state of `batch` independent solves.
There is no CPU fallback: without the built library every call raises.
"""
import ctypes as C
import os

import numpy as np

from .abi import (Params, Impurity, Plus, StateImp, STATE_P, STATE_D, Energies, SolveOpts, Result, Slot, Snapshot, OK, FIELDS,  # noqa: F401
                  SLOT_IDLE, SLOT_START, SLOT_RK, SLOT_PC, SLOT_DONE, SLOT_FAILED,
                  make_params, make_morse, CORE_OT, CORE_OTC, PREC_FP64, PREC_MIXED,
                  ICON_PERIODIC, ICON_ZERO, POT_MORSE, POT_NULL, POT_SERIES, POT_KOUTSELOS, H4OT, ERR_ARG,
                  Select, SELECT_PAIR, SELECT_JZ, SELECT_FIX, SELECT_AVERAGE, SELECT_MIXTURE)

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libeffpot.so")
_lib = None
_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_fp = np.ctypeslib.ndpointer(dtype=np.float32, flags="C_CONTIGUOUS")


class EffpotError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"status {code}: {msg}")
        self.code = code


def lib():
    """Load libeffpot.so; raises if it has not been built (no fallback)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is missing: build it with `python -m eff_pot_amd.build` "
                              "(the HIP path is the only implementation)")
        L = C.CDLL(LIB_PATH)
        vp = C.c_void_p
        L.effpot_ctx_create.argtypes = [C.POINTER(Params), C.c_int, C.POINTER(vp)]
        L.effpot_ctx_destroy.argtypes = [vp]
        L.effpot_last_error.argtypes = [vp]
        L.effpot_last_error.restype = C.c_char_p
        L.effpot_footprint_bytes.argtypes = [C.POINTER(Params)]
        L.effpot_footprint_bytes.restype = C.c_int64
        L.effpot_upload_psi.argtypes = [vp, C.c_int, _dp]
        L.effpot_download_psi.argtypes = [vp, C.c_int, _dp]
        L.effpot_download_den.argtypes = [vp, C.c_int, _dp]
        L.effpot_download_field.argtypes = [vp, C.c_int, C.c_int, _dp]
        L.effpot_poten_fp32.argtypes = [vp]
        L.effpot_download_field_fp32.argtypes = [vp, C.c_int, C.c_int, _fp]
        L.effpot_fft_r2c_fp32.argtypes = [vp, _fp, _fp]
        L.effpot_fft_c2r_fp32.argtypes = [vp, _fp, _fp]
        L.effpot_init_fermi.argtypes = [vp, C.c_int, C.c_double, C.c_double, C.c_double, C.c_double,
                                        C.c_uint64, _dp]
        L.effpot_set_impurity.argtypes = [vp, C.c_int, C.POINTER(Impurity)]
        L.effpot_set_natoms.argtypes = [vp, C.c_int, C.c_double]
        L.effpot_set_dt.argtypes = [vp, C.c_int, C.c_double]
        L.effpot_poten.argtypes = [vp]
        L.effpot_energy.argtypes = [vp, C.POINTER(Energies)]
        L.effpot_step_rk.argtypes = [vp, C.c_double]
        L.effpot_step_pc.argtypes = [vp, C.c_double] + [C.POINTER(C.c_double)] * 4
        L.effpot_renormalize.argtypes = [vp]
        L.effpot_renormalize_noalloc.argtypes = [vp, vp, C.c_int64, C.c_int, C.c_int, C.c_double, _dp, vp, vp]
        L.effpot_derivnd.argtypes = [vp, C.c_int, C.c_int, _dp, _dp]
        L.effpot_fft_r2c.argtypes = [vp, _dp, _dp]
        L.effpot_fft_c2r.argtypes = [vp, _dp, _dp]
        L.effpot_set_active.argtypes = [vp, C.c_int, C.c_int]
        L.effpot_solve.argtypes = [vp, C.POINTER(SolveOpts), C.POINTER(Result)]
        L.effpot_sync.argtypes = [vp]
        L.effpot_kernel_time.argtypes = [vp, C.c_int, C.c_int, C.POINTER(C.c_double)]
        L.effpot_plane_fused.argtypes = [vp, C.c_int]
        L.effpot_plane_inverse_fused.argtypes = [vp, C.c_int]
        L.effpot_zconv_threads.argtypes = [vp]
        L.effpot_packed_layout.argtypes = [vp, C.c_int]
        L.effpot_set_exclusion.argtypes = [vp, C.c_int, C.c_double]
        L.effpot_set_force_filter.argtypes = [vp, C.c_int, C.c_int]
        L.effpot_tick.argtypes = [vp, C.POINTER(SolveOpts), C.POINTER(Slot)]
        L.effpot_autotune.argtypes = [vp, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int)]
        L.effpot_set_tiles.argtypes = [vp, C.c_int, C.c_int]
        L.effpot_snapshot_read.argtypes = [C.c_char_p, C.POINTER(Snapshot), C.c_void_p]
        L.effpot_snapshot_write.argtypes = [C.c_char_p, C.POINTER(Snapshot), C.c_void_p, C.c_int]
        L.effpot_set_plus.argtypes = [vp, C.c_int, C.POINTER(Plus)]
        L.effpot_set_real_time.argtypes = [vp, C.c_int]
        L.effpot_set_impurity_motion.argtypes = [vp, C.c_int, _dp, C.c_double]
        L.effpot_get_impurity.argtypes = [vp, C.c_int, _dp, _dp, _dp]
        L.effpot_set_state_dynamics.argtypes = [vp, C.c_int, C.c_int, C.c_double]
        L.effpot_get_invar.argtypes = [vp, C.c_int, _dp]
        L.effpot_set_impurity_state.argtypes = [vp, C.c_int, C.POINTER(StateImp)]
        L.effpot_instates.argtypes = [C.c_int, _dp, C.c_double, C.c_int, _dp, _dp, C.POINTER(C.c_double)]
        L.effpot_instates_select.argtypes = [C.c_int, _dp, C.c_double, C.c_int, C.c_void_p, _dp, _dp,
                                             C.POINTER(C.c_double)]
        L.effpot_exciplex_moments.argtypes = [C.c_void_p, C.c_void_p, C.c_int, C.c_int, C.c_double, _dp]
        L.effpot_plus_preset.argtypes = [C.c_char_p, C.c_double, C.c_double, C.POINTER(Plus)]
        L.effpot_plus_eval.argtypes = [C.POINTER(Plus), C.c_double]
        L.effpot_plus_eval.restype = C.c_double
        _lib = L
    return _lib


def read_snapshot(path, header_only=False):
    """Read a wavefunction snapshot (readenc mode 3, formatted or unformatted).
    Returns (Snapshot header, psi complex array of shape (nz, ny, nx) = Fortran
    psi(nx,ny,nz), or None when header_only)."""
    h = Snapshot()
    rc = lib().effpot_snapshot_read(os.fsencode(path), C.byref(h), None)
    if rc != OK:
        raise EffpotError(rc, lib().effpot_last_error(None).decode())
    if header_only:
        return h, None
    buf = np.zeros((h.nz, h.ny, h.nx, 2))
    rc = lib().effpot_snapshot_read(os.fsencode(path), C.byref(h), buf.ctypes.data)
    if rc != OK:
        raise EffpotError(rc, lib().effpot_last_error(None).decode())
    return h, buf[..., 0] + 1j * buf[..., 1]


def write_snapshot(path, header: Snapshot, psi, binary=False):
    """Write a snapshot in printoutc's format (formatted, or unformatted when binary)."""
    psi = np.asarray(psi).reshape(header.nz, header.ny, header.nx)
    buf = np.ascontiguousarray(np.stack([psi.real, psi.imag], axis=-1), dtype=np.float64)
    rc = lib().effpot_snapshot_write(os.fsencode(path), C.byref(header), buf.ctypes.data, 1 if binary else 0)
    if rc != OK:
        raise EffpotError(rc, lib().effpot_last_error(None).decode())


# the reference's selec_plus / r_cutoff_plus / umax_plus defaults (V_impur.f90:14-16)
PLUS_DEFAULT = ("Ba_plus_2D2", 1.05, 6.260015e4)


def make_plus(selec=PLUS_DEFAULT[0], r_cutoff=PLUS_DEFAULT[1], umax=PLUS_DEFAULT[2]) -> Plus:
    """X+-He potential by Select_Pot name ('null', 'Ba_plus_gs', 'Ba_plus_2D0',
    'Ba_plus_2D1', 'Ba_plus_2D2', '<Na|K|Rb|Cs>_plus_Koutselos'); an unknown
    name raises (Select_Pot stops)."""
    p = Plus()
    rc = lib().effpot_plus_preset(selec.encode(), r_cutoff, umax, C.byref(p))
    if rc != OK:
        raise EffpotError(rc, f"{selec}: this potential is not defined (Select_Pot)")
    return p


def plus_eval(plus: Plus, r) -> float:
    """The X+ potential at distance r (A), in K, as the device evaluates it."""
    return lib().effpot_plus_eval(C.byref(plus), float(r))


def make_select(mode=SELECT_PAIR, ljz="", a_pi_3o2=0.94, a_sig_1o2=0.34117444, fixed_moments=False) -> Select:
    """Selection of invar: `ljz` is the reference's Ljz string ('-3/2', '+1/2',
    ... or '' for its default); the amplitudes default to modules.f90:305."""
    if ljz.strip():
        num, _, den = ljz.strip().partition("/")
        if den != "2" or not num.lstrip("+-").isdigit():
            raise EffpotError(ERR_ARG, f"wrong Ljz value: {ljz!r}")
        jz2 = int(num)
    else:
        jz2 = 0
    return Select(mode=mode, jz2=jz2, fixed_moments=int(bool(fixed_moments)), a_pi_3o2=a_pi_3o2,
                  a_sig_1o2=a_sig_1o2)


def make_state_imp(state, rimp=(0.0, 0.0, 0.0), instate=0, als=0.0, v_pi=None, v_sigma=None, v_delta=None,
                   invar=None, select=None, moments=None) -> StateImp:
    """A P or D internal-state impurity; the potentials default to the
    reference's selec_pi / selec_sigma / selec_delta (V_impur.f90:5-13).
    `select` (make_select) chooses invar; `moments` are kept as given when
    select.fixed_moments (an exciplex, exciplex_moments)."""
    imp = StateImp(state=state, instate=instate, als=als)
    if select is not None:
        imp.select = select
    if moments is not None:
        for i, m in enumerate(moments):
            imp.moments[i] = m
    for i in range(3):
        imp.rimp[i] = rimp[i]
    imp.v_pi = v_pi if v_pi is not None else make_plus("Ba_plus_2D1", 1.05, 5.216077e4)
    imp.v_sigma = v_sigma if v_sigma is not None else make_plus("Ba_plus_2D0", 1.05, 9.674836e4)
    imp.v_delta = v_delta if v_delta is not None else make_plus("Ba_plus_2D2", 1.05, 6.260015e4)
    if invar is not None:
        for i, c in enumerate(invar):
            imp.invar[2 * i], imp.invar[2 * i + 1] = c.real, c.imag
    return imp


def instates(state, moments, als, instate=0, select=None):
    """Host half of instates: (eigenvalues, invar, eso) of V + (Als/2) SO
    built from the density moments (effpot_instates_select; `select` from
    make_select, None for the Kramers-pair default)."""
    n = 6 if state == STATE_P else 10
    mom = np.zeros(31)
    mom[:len(moments)] = moments
    inv, eig, eso = np.zeros(20), np.zeros(10), C.c_double()
    rc = lib().effpot_instates_select(state, mom, als, instate, C.byref(select) if select is not None else None,
                                      inv, eig, C.byref(eso))
    if rc != OK:
        raise EffpotError(rc, "effpot_instates_select: bad arguments (instate, mode, Ljz or amplitudes)")
    return eig[:n], inv[:2 * n:2] + 1j * inv[1:2 * n:2], eso.value


def exciplex_moments(v_pi, v_sigma, exciplex="Ring", axis="Z", r_exc=3.9):
    """Lexciplex_state_fix: the P moments of a 'Ring' (6 He) or 'Linear' (2 He)
    exciplex at r_exc about `axis` (effpot_exciplex_moments; defaults
    modules.f90:288-289, 304)."""
    n_he = {"Ring": 6, "Linear": 2}.get(exciplex.strip())
    ax = {"X": 0, "Y": 1, "Z": 2}.get(axis.strip().upper())
    if n_he is None or ax is None:
        raise EffpotError(ERR_ARG, f"exciplex must be Ring or Linear and axis X, Y or Z: {exciplex!r} {axis!r}")
    m = np.zeros(7)
    rc = lib().effpot_exciplex_moments(C.byref(v_pi), C.byref(v_sigma), n_he, ax, r_exc, m)
    if rc != OK:
        raise EffpotError(rc, "effpot_exciplex_moments: bad arguments")
    return m


def renormalize_noalloc(psi_ptr, npoints, nslots, psi_complex, dxyz, n4, den_ptr=None, natoms_ptr=None,
                        stream=None):
    """effpot_renormalize_noalloc on device pointers (ints) owned by the
    caller; n4 is one value per slot.  Enqueues on `stream` (a hipStream_t
    handle, None: the null stream) and returns without waiting."""
    n4 = np.ascontiguousarray(np.broadcast_to(np.asarray(n4, dtype=np.float64), (nslots,)))
    rc = lib().effpot_renormalize_noalloc(psi_ptr, den_ptr, int(npoints), int(nslots), 1 if psi_complex else 0,
                                          float(dxyz), n4, natoms_ptr, stream)
    if rc != OK:
        raise EffpotError(rc, lib().effpot_last_error(None).decode())


def footprint_bytes(params: Params) -> int:
    return int(lib().effpot_footprint_bytes(C.byref(params)))


class Context:
    """Device state of `params.batch` solves on one GPU (opaque effpot_ctx)."""

    def __init__(self, params: Params, device: int = 0):
        self.p = params
        self.h = C.c_void_p()
        rc = lib().effpot_ctx_create(C.byref(params), device, C.byref(self.h))
        if rc != OK:
            raise EffpotError(rc, "effpot_ctx_create failed")
        self.B = params.batch
        self.shape = (params.nz, params.ny, params.nx)
        self.hshape = (params.nz, params.ny, params.nx // 2 + 1)

    def _chk(self, rc):
        if rc != OK:
            msg = lib().effpot_last_error(self.h)
            raise EffpotError(rc, msg.decode() if msg else "")

    def close(self):
        if self.h:
            lib().effpot_ctx_destroy(self.h)
            self.h = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ---- state -----------------------------------------------------------
    def upload_psi(self, psi, slot=0):
        psi = np.asarray(psi).reshape(self.shape)
        buf = np.stack([psi.real, psi.imag], axis=-1).astype(np.float64).copy()
        self._chk(lib().effpot_upload_psi(self.h, slot, buf.reshape(-1)))

    def psi(self, slot=0):
        out = np.zeros(self.shape + (2,))
        self._chk(lib().effpot_download_psi(self.h, slot, out.reshape(-1)))
        return out[..., 0] + 1j * out[..., 1]

    def psi_raw(self, slot=0):
        out = np.zeros(self.shape + (2,))
        self._chk(lib().effpot_download_psi(self.h, slot, out.reshape(-1)))
        return out

    def den(self, slot=0):
        out = np.zeros(self.shape)
        self._chk(lib().effpot_download_den(self.h, slot, out.reshape(-1)))
        return out

    def field(self, which, slot=0):
        out = np.zeros(self.shape)
        self._chk(lib().effpot_download_field(self.h, slot, FIELDS.get(which, which), out.reshape(-1)))
        return out

    def field_fp32(self, which, slot=0):
        """A field of the last poten_fp32 (mixed contexts); 'corr' is the defect correction."""
        out = np.zeros(self.shape, dtype=np.float32)
        w = 13 if which == "corr" else FIELDS.get(which, which)
        self._chk(lib().effpot_download_field_fp32(self.h, slot, w, out.reshape(-1)))
        return out

    def poten_fp32(self):
        """poten on the FP32 fields of a mixed context (from the live psi)."""
        self._chk(lib().effpot_poten_fp32(self.h))

    def fft_r2c_fp32(self, f):
        f = np.ascontiguousarray(f, dtype=np.float32).reshape(self.shape)
        out = np.zeros(self.hshape + (2,), dtype=np.float32)
        self._chk(lib().effpot_fft_r2c_fp32(self.h, f.reshape(-1), out.reshape(-1)))
        return out[..., 0] + 1j * out[..., 1]

    def fft_c2r_fp32(self, F):
        F = np.asarray(F).reshape(self.hshape)
        buf = np.ascontiguousarray(np.stack([F.real, F.imag], axis=-1), dtype=np.float32)
        out = np.zeros(self.shape, dtype=np.float32)
        self._chk(lib().effpot_fft_c2r_fp32(self.h, buf.reshape(-1), out.reshape(-1)))
        return out

    def init_fermi(self, rho0, R, afermi, noise, seed, centre=(0.0, 0.0, 0.0), slot=0):
        self._chk(lib().effpot_init_fermi(self.h, slot, rho0, R, afermi, noise, seed,
                                          np.asarray(centre, dtype=np.float64)))

    def set_natoms(self, n4, slot=0):
        self._chk(lib().effpot_set_natoms(self.h, slot, n4))

    def set_dt(self, dt, slot=0):
        self._chk(lib().effpot_set_dt(self.h, slot, dt))

    def set_impurity(self, imp: Impurity, slot=0):
        self._chk(lib().effpot_set_impurity(self.h, slot, C.byref(imp)))

    def set_real_time(self, on=True):
        """Real-time propagation for step_rk / step_pc (timec = i)."""
        self._chk(lib().effpot_set_real_time(self.h, 1 if on else 0))

    def set_impurity_motion(self, v, mass_u, slot=0):
        """The slot's isotropic impurity moves (velocity v in A per K^-1, mass in u)."""
        self._chk(lib().effpot_set_impurity_motion(self.h, slot, np.asarray(v, dtype=np.float64), mass_u))

    def impurity(self, slot=0):
        """(r, v, a) of the slot's impurity."""
        r, v, a = np.zeros(3), np.zeros(3), np.zeros(3)
        self._chk(lib().effpot_get_impurity(self.h, slot, r, v, a))
        return r, v, a

    def set_state_dynamics(self, on=True, ev0=0.0, slot=0):
        """The slot's P/D invar evolves in step_rk / step_pc (potinvar)."""
        self._chk(lib().effpot_set_state_dynamics(self.h, slot, 1 if on else 0, ev0))

    def invar(self, slot=0):
        v = np.zeros(20)
        self._chk(lib().effpot_get_invar(self.h, slot, v))
        return v[0::2] + 1j * v[1::2]

    def set_impurity_state(self, imp: StateImp, slot=0) -> StateImp:
        """P/D-state impurity (instates on the slot's current psi when
        imp.instate >= 0); fills imp.invar, eigen, moments and eso."""
        self._chk(lib().effpot_set_impurity_state(self.h, slot, C.byref(imp)))
        return imp

    def set_exclusion(self, r, slot=0):
        """z_exciplex_exclusion (energy.f90:155): grid points nearer than r to
        the impurity are left out of eHeX and uplus (0: none)."""
        self._chk(lib().effpot_set_exclusion(self.h, slot, r))

    def set_force_filter(self, on, slot=0):
        """Lfilter_exciplex_force (potenimp.f90:216-245): the z force of a
        moving impurity sums only the planes beyond the exclusion radius."""
        self._chk(lib().effpot_set_force_filter(self.h, slot, 1 if on else 0))

    def set_plus(self, plus, slot=0):
        """X+ potential of the uplus energy term for one slot (None: none)."""
        self._chk(lib().effpot_set_plus(self.h, slot, C.byref(plus) if plus is not None else None))

    def set_active(self, slot, active):
        self._chk(lib().effpot_set_active(self.h, slot, 1 if active else 0))

    # ---- operators (reference routine names) -------------------------------
    def poten(self):
        self._chk(lib().effpot_poten(self.h))

    def energy(self):
        e = (Energies * self.B)()
        self._chk(lib().effpot_energy(self.h, e))
        return [x.as_dict() for x in e]

    def step_rk(self, dt):
        self._chk(lib().effpot_step_rk(self.h, dt))

    def step_pc(self, dt, impurity_errors=False):
        """One predictor-corrector step; returns errHe per slot, or with
        impurity_errors the lists (errHe, errimp, errvimp, erriv) of
        steppc(deltat, errHe, errimp, errvimp, erriv)."""
        err = (C.c_double * self.B)()
        if not impurity_errors:
            self._chk(lib().effpot_step_pc(self.h, dt, err, None, None, None))
            return list(err)
        e_r, e_v, e_i = (C.c_double * self.B)(), (C.c_double * self.B)(), (C.c_double * self.B)()
        self._chk(lib().effpot_step_pc(self.h, dt, err, e_r, e_v, e_i))
        return list(err), list(e_r), list(e_v), list(e_i)

    def renormalize(self):
        self._chk(lib().effpot_renormalize(self.h))

    def derivnd(self, id_, iv, f):
        f = np.ascontiguousarray(f, dtype=np.float64).reshape(self.shape)
        out = np.zeros_like(f)
        self._chk(lib().effpot_derivnd(self.h, id_, iv, f.reshape(-1), out.reshape(-1)))
        return out

    def fft_r2c(self, f):
        f = np.ascontiguousarray(f, dtype=np.float64).reshape(self.shape)
        out = np.zeros(self.hshape + (2,))
        self._chk(lib().effpot_fft_r2c(self.h, f.reshape(-1), out.reshape(-1)))
        return out[..., 0] + 1j * out[..., 1]

    def fft_c2r(self, F):
        F = np.asarray(F).reshape(self.hshape)
        buf = np.stack([F.real, F.imag], axis=-1).astype(np.float64).copy()
        out = np.zeros(self.shape)
        self._chk(lib().effpot_fft_c2r(self.h, buf.reshape(-1), out.reshape(-1)))
        return out

    def solve(self, opts: SolveOpts):
        r = (Result * self.B)()
        self._chk(lib().effpot_solve(self.h, C.byref(opts), r))
        return list(r)

    def new_slots(self):
        return (Slot * self.B)()

    def tick(self, opts: SolveOpts, slots):
        """Advance every START/RK/PC slot by one step of its own phase."""
        self._chk(lib().effpot_tick(self.h, C.byref(opts), slots))

    def autotune(self, reps=10):
        tx, zc = C.c_int(), C.c_int()
        self._chk(lib().effpot_autotune(self.h, reps, C.byref(tx), C.byref(zc)))
        return tx.value, zc.value

    def set_tiles(self, tx, zc):
        self._chk(lib().effpot_set_tiles(self.h, tx, zc))

    def kernel_time(self, which, reps=20):
        ms = C.c_double()
        self._chk(lib().effpot_kernel_time(self.h, which, reps, C.byref(ms)))
        return ms.value

    def plane_fused(self, fp32_fields):
        """Schedule of poten()'s forward x and y transforms: 0 separate
        passes, 1 one plane launch with scratch, 3 one plane launch in place."""
        return int(lib().effpot_plane_fused(self.h, 1 if fp32_fields else 0))

    def plane_inverse_fused(self, fp32_fields):
        """1 when poten()'s inverse y and x transforms run as one in-place
        plane launch, 0 for separate passes."""
        return int(lib().effpot_plane_inverse_fused(self.h, 1 if fp32_fields else 0))

    def packed_layout(self, fp32_fields):
        """1 when poten()'s half spectra use the packed layout between the plane kernels."""
        return int(lib().effpot_packed_layout(self.h, 1 if fp32_fields else 0))

    def zconv_threads(self):
        """Threads per z-convolution work-group (256 or 512) this context uses."""
        return int(lib().effpot_zconv_threads(self.h))

    def sync(self):
        self._chk(lib().effpot_sync(self.h))


def default_dt(params: Params, cap=1000.0):
    """Largest stable step of the Hamming PC integrator with margin:
    0.25 / (h2o2m4 * 7.07 * sum 1/h^2 + cap + 50 K); 7.07 = max |symbol| of
    the 13-point second derivative."""
    hx = 2 * params.xmax / params.nx
    hy = 2 * params.ymax / params.ny
    hz = 2 * params.zmax / params.nz
    lam = params.h2o2m4 * 7.0729 * (1 / hx ** 2 + 1 / hy ** 2 + 1 / hz ** 2) + cap + 50.0
    return 0.25 / lam
