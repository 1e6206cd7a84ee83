"""ctypes mirror of include/effpot.h (structs and status codes).

The same structures are used by the HIP library (libeffpot.so) and, in the
tests only, by the CPU oracle; the layouts follow the C header field for field.
"""
import ctypes as C

OK = 0
ERR_ARG, ERR_GRID, ERR_CORE, ERR_POT, ERR_DEVICE, ERR_NONFINITE, ERR_STATE, ERR_NOMEM, ERR_DT = range(1, 10)
CORE_OT, CORE_OTC = 0, 1
PREC_FP64, PREC_MIXED = 0, 1
POT_NULL, POT_MORSE, POT_SERIES, POT_KOUTSELOS = 0, 1, 2, 3
SERIES_MAX = 16
ICON_PERIODIC, ICON_ZERO = 8, 13

FIELDS = {"pot4": 0, "dencg": 1, "delj4": 2, "denalf": 3, "ualphas": 4, "uimp": 5,
          "dxden": 6, "dyden": 7, "dzden": 8, "sto1": 9, "intxalf": 10, "intyalf": 11,
          "intzalf": 12}


class Params(C.Structure):
    _fields_ = [("nx", C.c_int32), ("ny", C.c_int32), ("nz", C.c_int32),
                ("core", C.c_int32), ("npd", C.c_int32), ("icon", C.c_int32),
                ("precision", C.c_int32), ("batch", C.c_int32),
                ("psi_complex", C.c_int32), ("reserved_", C.c_int32),
                ("xmax", C.c_double), ("ymax", C.c_double), ("zmax", C.c_double),
                ("n4", C.c_double), ("h_core", C.c_double),
                ("eps4", C.c_double), ("sigma4", C.c_double),
                ("cp4", C.c_double), ("cpp4", C.c_double),
                ("alphas", C.c_double), ("l", C.c_double), ("den0s", C.c_double),
                ("h2o2m4", C.c_double),
                ("C", C.c_double), ("beta", C.c_double), ("den_m", C.c_double)]


class Impurity(C.Structure):
    _fields_ = [("kind", C.c_int32), ("pad_", C.c_int32),
                ("D", C.c_double), ("re", C.c_double), ("a", C.c_double),
                ("r_cutoff", C.c_double), ("umax", C.c_double), ("cap", C.c_double),
                ("rimp", C.c_double * 3)]


class Plus(C.Structure):
    """X+-He potential of the uplus energy term (effpot_plus)."""
    _fields_ = [("kind", C.c_int32), ("nterms", C.c_int32), ("k0", C.c_int32), ("pad_", C.c_int32),
                ("r_cutoff", C.c_double), ("umax", C.c_double), ("rcutoff", C.c_double), ("fcutoff", C.c_double),
                ("r0", C.c_double), ("aa", C.c_double), ("bb", C.c_double), ("cc", C.c_double),
                ("pg", C.c_double * SERIES_MAX), ("pg_lo", C.c_double * SERIES_MAX)]


STATE_P, STATE_D = 1, 2
SELECT_PAIR, SELECT_JZ, SELECT_FIX, SELECT_AVERAGE, SELECT_MIXTURE = 0, 1, 2, 3, 4


class Select(C.Structure):
    """How instates picks invar (effpot_select)."""
    _fields_ = [("mode", C.c_int32), ("jz2", C.c_int32), ("fixed_moments", C.c_int32), ("pad_", C.c_int32),
                ("a_pi_3o2", C.c_double), ("a_sig_1o2", C.c_double)]


class StateImp(C.Structure):
    """Impurity in a P or D internal state (effpot_state_imp)."""
    _fields_ = [("state", C.c_int32), ("instate", C.c_int32), ("als", C.c_double), ("rimp", C.c_double * 3),
                ("v_pi", Plus), ("v_sigma", Plus), ("v_delta", Plus), ("invar", C.c_double * 20),
                ("eigen", C.c_double * 10), ("moments", C.c_double * 31), ("eso", C.c_double),
                ("select", Select)]

    def invar_c(self):
        n = 6 if self.state == STATE_P else 10
        return [complex(self.invar[2 * i], self.invar[2 * i + 1]) for i in range(n)]


class Energies(C.Structure):
    _fields_ = [(n, C.c_double) for n in
                ("natoms", "etot4", "ekin4", "elj4", "ealphas", "esolid", "ecor4", "eimpu",
                 "ekinx", "eHeX", "uplus", "eso", "etot", "mu", "xcm", "ycm", "zcm", "resid")]

    def as_dict(self):
        return {n: getattr(self, n) for n, _ in self._fields_}


class SolveOpts(C.Structure):
    _fields_ = [("dt", C.c_double), ("tol", C.c_double), ("n_rk", C.c_int32),
                ("pener", C.c_int32), ("max_steps", C.c_int32), ("polish_steps", C.c_int32),
                ("tol_fp32", C.c_double), ("check_finite", C.c_int32), ("graph_steps", C.c_int32)]


SLOT_IDLE, SLOT_START, SLOT_RK, SLOT_PC, SLOT_DONE, SLOT_FAILED = range(6)


class Slot(C.Structure):
    _fields_ = [("phase", C.c_int32), ("fp32", C.c_int32), ("nrk", C.c_int32), ("steps", C.c_int32),
                ("since", C.c_int32), ("converged", C.c_int32), ("eprev", C.c_double),
                ("rel_de", C.c_double), ("e", Energies)]


SNAPSHOT_MAX_INVAR = 16


class Snapshot(C.Structure):
    """Header of a wavefunction snapshot (printoutc.f90:69-114 / readenc.f90:151-168)."""
    _fields_ = [("nx", C.c_int32), ("ny", C.c_int32), ("nz", C.c_int32), ("ninvar", C.c_int32),
                ("xmax", C.c_double), ("ymax", C.c_double), ("zmax", C.c_double),
                ("hx", C.c_double), ("hy", C.c_double), ("hz", C.c_double),
                ("rimp", C.c_double * 3), ("vimp", C.c_double * 3),
                ("invar", C.c_double * (2 * SNAPSHOT_MAX_INVAR)), ("binary", C.c_int32), ("reserved_", C.c_int32)]


class Result(C.Structure):
    _fields_ = [("e", Energies), ("steps", C.c_int32), ("converged", C.c_int32),
                ("rel_de", C.c_double), ("err_he", C.c_double)]


# Orsay-Trento 4He functional constants (modules.f90:116-135, 187-199).
OT_DEFAULTS = dict(eps4=10.22, sigma4=2.556, cp4=-2.41186e4, cpp4=1.85850e6,
                   alphas=54.31, l=1.0, den0s=0.04, h2o2m4=6.05969638298,
                   C=3.1577504e4, beta=40.0, den_m=0.37)
H4OT = 2.190323   # modules.f90:192
H4OP = 2.359665   # modules.f90:191


def make_params(n, xmax, n4=300.0, core=CORE_OTC, h_core=2.377, icon=ICON_ZERO,
                precision=PREC_FP64, batch=1, ny=None, nz=None, ymax=None, zmax=None, psi_complex=1, **kw):
    d = dict(OT_DEFAULTS)
    d.update(kw)
    return Params(nx=n, ny=ny or n, nz=nz or n, core=core, npd=13, icon=icon,
                  precision=precision, batch=batch, psi_complex=psi_complex, xmax=xmax, ymax=ymax or xmax,
                  zmax=zmax or xmax, n4=n4, h_core=h_core, **d)


def make_morse(rimp=(0.0, 0.0, 0.0), D=30.0, re=4.0, a=1.0, r_cutoff=0.0, umax=0.0, cap=1000.0):
    imp = Impurity(kind=POT_MORSE, D=D, re=re, a=a, r_cutoff=r_cutoff, umax=umax, cap=cap)
    for i in range(3):
        imp.rimp[i] = rimp[i]
    return imp
