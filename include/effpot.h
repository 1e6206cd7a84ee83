/*
 * effpot.h -- C ABI of the MI355X (gfx950) Orsay-Trento imaginary-time solver.
 *
 * The reference (bcntls2016/eff_pot, Fortran) keeps all state in module
 * globals and exposes zero-argument subroutines.  This ABI replaces those
 * subroutines one-for-one with calls on an opaque context that owns the
 * device memory of `batch` independent solves (slots) sharing one grid and
 * one functional.  Every entry point returns EFFPOT_OK (0) or a non-zero
 * status; effpot_last_error() returns the message.  Where the reference
 * would STOP, this ABI returns a status and the Fortran shim turns it into
 * `error stop` with the same message.
 *
 * Arrays crossing the boundary are column-major with ix fastest, exactly as
 * the Fortran module arrays (dimen.f90:29-145): real fields are
 * double[nz][ny][nx]; the wavefunction is interleaved complex double
 * (re, im) per point, i.e. Fortran complex(8) psi(nx,ny,nz).
 * The library never keeps host pointers past the call that received them.
 */
#ifndef EFFPOT_H
#define EFFPOT_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ------------------------------------------------------ */
#define EFFPOT_OK 0
#define EFFPOT_ERR_ARG 1       /* bad argument / NULL pointer                     */
#define EFFPOT_ERR_GRID 2      /* grid size not supported: see effpot_params.nx   */
#define EFFPOT_ERR_CORE 3      /* unknown core4 (DFT4He3d.f90:355-363)            */
#define EFFPOT_ERR_POT 4       /* unknown impurity potential (V_impur.f90:928-931)*/
#define EFFPOT_ERR_DEVICE 5    /* HIP runtime failure                             */
#define EFFPOT_ERR_NONFINITE 6 /* non-positive / NaN density (steprk.f90:94-110)  */
#define EFFPOT_ERR_STATE 7     /* call out of order (e.g. step before impurity)   */
#define EFFPOT_ERR_NOMEM 8     /* device allocation failed                        */
#define EFFPOT_ERR_DT 9        /* inconsistent / non-positive time step           */
#define EFFPOT_ERR_IO 10       /* snapshot file missing, truncated or malformed   */

/* ---- enumerations ------------------------------------------------------ */
#define EFFPOT_CORE_OT 0  /* core4='OT '  (poten.f90:27-92, no alpha_s)      */
#define EFFPOT_CORE_OTC 1 /* core4='OTC'  (poten.f90:94-105, term_alfa.f90) */

#define EFFPOT_PREC_FP64 0  /* every array and operation in FP64                 */
#define EFFPOT_PREC_MIXED 1 /* FP32 bulk fields/FFTs/stencils, FP64 reductions   */

#define EFFPOT_POT_NULL 0  /* Select_Pot 'null' (V_impur.f90:171-176): U = 0, r_cutoff ignored */
#define EFFPOT_POT_MORSE 1 /* isotropic Morse D[e^{-2a(r-re)} - 2e^{-a(r-re)}] */
#define EFFPOT_POT_SERIES 2 /* inverse-power fit (X+-He potentials, see effpot_plus) */
#define EFFPOT_POT_KOUTSELOS 3 /* Koutselos alkali-ion form (X+-He, see effpot_plus) */

#define EFFPOT_ICON_PERIODIC 8 /* derivnd icon=8  (init_deriv_parallel.f90:529-634) */
#define EFFPOT_ICON_ZERO 13    /* derivnd icon=13 (redef.f:133-170): f=0 outside  */

/* ---- parameters -------------------------------------------------------- */
/* Grid, functional and run-shape parameters: the namelist subset of
 * DFT4He3d.f90:104-134 that the hot path reads, plus h_core (the LJ core
 * radius h4, chosen from core4 in DFT4He3d.f90:346-363, here explicit). */
typedef struct effpot_params {
  int32_t nx, ny, nz; /* grid points per axis: 16, 24, 32, 48, 64, 96, 128,
                         192 or 256 (the sizes of the register/LDS FFTs,
                         N = N1 N2 with N1 a power of two >= N2).  The
                         reference's tstgrid (tstgrid.f90:8-12) allows any
                         2^a 3^b 5^c <= 256; the other sizes (e.g. 40, 80,
                         100, 120, 144, 160, 180, 240) get EFFPOT_ERR_GRID
                         from effpot_ctx_create, and effpot_last_error(NULL)
                         says which axis                                    */
  int32_t core;       /* EFFPOT_CORE_OT | EFFPOT_CORE_OTC                   */
  int32_t npd;        /* stencil points; only 13 is supported (modules.f90:15) */
  int32_t icon;       /* EFFPOT_ICON_ZERO (13, default) | EFFPOT_ICON_PERIODIC */
  int32_t precision;  /* EFFPOT_PREC_FP64 | EFFPOT_PREC_MIXED               */
  int32_t batch;      /* number of independent solves (slots), >= 1         */
  int32_t psi_complex;/* 1: complex psi (e.g. a real-time snapshot); 0: real psi
                         (imaginary time from a real start keeps Im psi = 0
                         exactly, so the real path does the same arithmetic on
                         half the bytes, equal up to FMA-contraction rounding);
                         uploads with Im psi != 0 are then rejected           */
  int32_t reserved_;
  double xmax, ymax, zmax; /* box half widths; h = 2*xmax/nx (DFT4He3d.f90:205-207) */
  double n4;          /* number of He atoms the density is normalised to   */
  double h_core;      /* LJ core radius h4 (A)                             */
  double eps4, sigma4;          /* LJ parameters (modules.f90:193-194)     */
  double cp4, cpp4;             /* correlation c2, c3 (modules.f90:122-123)*/
  double alphas, l, den0s;      /* alpha_s term (modules.f90:125-127)      */
  double h2o2m4;                /* hbar^2/2m_4 (modules.f90:128)           */
  double C, beta, den_m;        /* solid penalty (modules.f90:129-131)     */
} effpot_params;

/* Impurity: the potential kind and its parameters, the Select_Pot cutoff
 * (r < r_cutoff -> umax, V_impur.f90:87-932 pattern), the cap applied to the
 * interpolated field (potenimp.f90:184-186, 1000 K) and the position. */
typedef struct effpot_impurity {
  int32_t kind;      /* EFFPOT_POT_MORSE | EFFPOT_POT_NULL               */
  int32_t pad_;
  double D, re, a;   /* Morse: depth (K), minimum (A), range (1/A)        */
  double r_cutoff;   /* r < r_cutoff -> umax                              */
  double umax;
  double cap;        /* uimp = min(uimp, cap)                             */
  double rimp[3];    /* impurity position (A)                             */
} effpot_impurity;

/* X+-He potential of the uplus term (energy.f90:146-163): the radial
 * potential Select_Pot(selec_plus, r, r_cutoff_plus, umax_plus) evaluated at
 * |rimp - r| (V_impur.f90:87).  kind EFFPOT_POT_NULL is 0 everywhere, as
 * 'null' is (V_impur.f90:171-176); kind EFFPOT_POT_SERIES is the
 * form of the reference's inverse-power fits (V_Ba_plus_*, V_impur.f90:1218-1396):
 *   r <  r_cutoff            umax
 *   r <= rcutoff             fcutoff
 *   r <= r0                  |aa r^2 + bb r + cc|
 *   otherwise                sum_{k=1..nterms} pg[k-1] r^-(k+k0)
 * The fits' coefficients carry 33 digits; pg + pg_lo holds them as a
 * double-double pair and the device sums the series in double-double, so the
 * cancellation between terms of 1e10 K near r0 costs no precision.
 * Kind EFFPOT_POT_KOUTSELOS is the form of V_Koutselos_*_plus_gs
 * (V_impur.f90:1704-1883): r < r_cutoff -> umax, otherwise, with
 * x = r / a_bohr, E_h = Hartree * eV_K,
 *   [v0 (A e^{-alpha x/rho} - B e^{-beta x/rho})
 *    + (-CC4/x^4 - CC6/x^6 - CC8/x^8) h(x)] E_h,
 *   h = 1 for x >= 1.28 R_m, else exp(-(1.28 R_m/x - 1)^2),
 * its parameters in pg[EFFPOT_KOUT_*] (nterms, k0 and the plateau and
 * parabola fields unused). */
#define EFFPOT_SERIES_MAX 16
#define EFFPOT_KOUT_V0 0
#define EFFPOT_KOUT_RHO 1
#define EFFPOT_KOUT_RM 2
#define EFFPOT_KOUT_CC4 3
#define EFFPOT_KOUT_CC6 4
#define EFFPOT_KOUT_CC8 5
#define EFFPOT_KOUT_A 6
#define EFFPOT_KOUT_ALPHA 7
#define EFFPOT_KOUT_B 8
#define EFFPOT_KOUT_BETA 9
#define EFFPOT_KOUT_ABOHR 10   /* a_bohr in A                              */
#define EFFPOT_KOUT_HARTREE 11 /* Hartree in eV                            */
#define EFFPOT_KOUT_EVK 12     /* eV in K                                  */
typedef struct effpot_plus {
  int32_t kind;       /* EFFPOT_POT_NULL | EFFPOT_POT_SERIES | EFFPOT_POT_KOUTSELOS */
  int32_t nterms;     /* series terms, 1..EFFPOT_SERIES_MAX                */
  int32_t k0;         /* power offset, >= 0                                */
  int32_t pad_;
  double r_cutoff, umax;   /* Select_Pot's cutoff                          */
  double rcutoff, fcutoff; /* inner plateau                                */
  double r0, aa, bb, cc;   /* parabola between rcutoff and r0              */
  double pg[EFFPOT_SERIES_MAX];    /* coefficients (high parts)            */
  double pg_lo[EFFPOT_SERIES_MAX]; /* low parts (0 for plain doubles)      */
} effpot_plus;

/* Impurity in a P or D internal state (Lstate 'P' / 'D', modules.f90:303):
 * the He field of potenimp (potenimp.f90:34-180) for the internal state
 * invar, chosen by instates (instates.f90:1-800) from the current helium
 * density.  The radial potentials are Select_Pot entries (V_pi, V_sigma,
 * V_delta, V_impur.f90:66-85); their umax also gives updatepoten's values at
 * r <= 1e-3 A (potenimp.f90:536-563).  Fields marked "out" are written. */
#define EFFPOT_STATE_P 1   /* 6 components: (x,y,z) orbital x (up,down) spin  */
#define EFFPOT_STATE_D 2   /* 10 components: d1..d5 x (up,down)               */

/* How instates picks invar among the eigenvectors of V + SO (the branches
 * of instates.f90:228-364 for P and :613-726 for D, in the reference's order
 * of precedence Ldiag_jz > Lexcite_state_fix > Laverage_P_value > Lmixture).
 * FIX, AVERAGE and MIXTURE exist for P only. */
#define EFFPOT_SELECT_PAIR 0    /* (v_2k+1 + v_2k+2)/sqrt2, k = instate (default) */
#define EFFPOT_SELECT_JZ 1      /* Ldiag_jz: the eigenvector of Jz = Lz + Sz within
                                   the j manifold of `instate` (P: 0 -> j=1/2,
                                   1,2 -> j=3/2; D: 0,1 -> j=3/2, 2..4 -> j=5/2)
                                   whose eigenvalue is the jz2/2-th (Ljz)       */
#define EFFPOT_SELECT_FIX 2     /* Lexcite_state_fix: |x up>, (|x up>+|y up>)/sqrt2
                                   or (|x up>+|y up>+|z up>)/sqrt3 for instate
                                   0, 1, >= 2; no diagonalisation              */
#define EFFPOT_SELECT_AVERAGE 3 /* Laverage_P_value: sum of the six eigenvectors
                                   over sqrt6                                  */
#define EFFPOT_SELECT_MIXTURE 4 /* Lmixture: sqrt(a_pi_3o2) pair 1 + sqrt(a_sig_1o2)
                                   pair 2 + sqrt(max(1-a_pi_3o2-a_sig_1o2,0))
                                   pair 0, each pair as in PAIR                */
typedef struct effpot_select {
  int32_t mode;        /* EFFPOT_SELECT_*; 0 (a zeroed struct) is the default */
  int32_t jz2;         /* JZ: twice Ljz (-5..5, odd); 0 = the reference's
                          default for `instate` (modules.f90:310-318)      */
  int32_t fixed_moments; /* 1: `moments` is input, not summed over the
                          density (Lexciplex_state_fix, instates.f90:54-113,
                          with effpot_exciplex_moments)                     */
  int32_t pad_;
  double a_pi_3o2, a_sig_1o2; /* MIXTURE amplitudes (>= 0)                 */
} effpot_select;

typedef struct effpot_state_imp {
  int32_t state;       /* EFFPOT_STATE_P | EFFPOT_STATE_D                   */
  int32_t instate;     /* >= 0: the level (Kramers pair, 0 = lowest) or j
                          manifold that `select` picks from (instates.f90:
                          360-364, :722-726); -1: invar as given (readenc
                          instate 5)                                       */
  double als;          /* spin-orbit splitting Als_P / Als_D (K)            */
  double rimp[3];      /* impurity position (A)                            */
  effpot_plus v_pi, v_sigma, v_delta; /* v_delta is unused for P           */
  double invar[20];    /* in (instate -1) / out: complex invar(1..6|10)     */
  double eigen[10];    /* out: eigenvalues of V + SO, ascending (K)         */
  double moments[31];  /* out: density moments of instates (P: A, Vx, Vy, Vz,
                          Vxy, Vxz, Vyz; D: MP(i<=j), MS(i<=j) row-major,
                          then the Delta sum), times dxyz; input when
                          select.fixed_moments                              */
  double eso;          /* out: spin-orbit energy (energy.f90:170-192)       */
  effpot_select select; /* how invar is chosen (zeroed: PAIR)               */
} effpot_state_imp;

/* Energies as printed by DFT4He3d.f90:460-463 (formats 6050/6060) and
 * computed by energy.f90:12-202, plus the chemical potential and the
 * helium centre of mass (r_cm.f90:22-57). */
typedef struct effpot_energies {
  double natoms;   /* sum(den)*dxyz                                       */
  double etot4;    /* helium total (energy.f90:101-109)                   */
  double ekin4;    /* h2o2m4 sum |grad psi|^2 dV                          */
  double elj4;     /* 1/2 sum delj4*den dV                                */
  double ealphas;  /* alpha_s term (OTC)                                  */
  double esolid;   /* solid penalty energy, reported, not added (:96)     */
  double ecor4;    /* correlation                                         */
  double eimpu;    /* impurity energy E_imp = ekinx + eHeX + eso (:195)   */
  double ekinx;    /* classical impurity kinetic energy (0: fixed)        */
  double eHeX;     /* sum den*uimp dV (:146-165)                          */
  double uplus;    /* X+-He interaction (0: no selec_plus table)          */
  double eso;      /* spin-orbit (0 for an isotropic impurity)            */
  double etot;     /* etot4 + eimpu (:197)                                */
  double mu;       /* chemical potential <psi|H|psi>/N                    */
  double xcm, ycm, zcm; /* helium centre of mass, normalised by n4          */
  double resid;    /* |H psi - mu psi| / (|mu| |psi|): how far psi is from an
                      eigenstate of its own mean field (the residual norm
                      varmu.f90:11-29 keeps); 0 at convergence, independent
                      of the time step and of the trajectory               */
} effpot_energies;

/* Options of the relaxation loop (effpot_solve). */
typedef struct effpot_solve_opts {
  double dt;        /* imaginary time step (1/K)                          */
  double tol;       /* converged when |dE/E| per step < tol               */
  int32_t n_rk;     /* Runge-Kutta-Gill starter steps (3)                  */
  int32_t pener;    /* steps between energy evaluations (DFT4He3d pener)   */
  int32_t max_steps;/* hard cap on steps (not counting polish)             */
  int32_t polish_steps; /* MIXED only: cap on FP64 steps after the FP32 phase */
  double tol_fp32;  /* MIXED only: per-step |dE/E| that ends the FP32 phase
                       (>= tol)                                            */
  int32_t check_finite; /* 1: scan the density for NaN/<=0 at every energy */
  int32_t graph_steps;  /* predictor-corrector steps replayed per graph launch
                           while no slot of a phase is due an energy check,
                           a phase change or its step cap (0 or 1: one step
                           per effpot_tick; results are the same bits)      */
} effpot_solve_opts;

typedef struct effpot_result {
  effpot_energies e; /* energies of the final state                       */
  int32_t steps;     /* steps taken (RK + PC [+ polish])                   */
  int32_t converged; /* 1 if the tol criterion was met                     */
  double rel_de;     /* last |dE/E| per step                               */
  double err_he;     /* last predictor-corrector error estimate errHe      */
} effpot_result;

/* Per-slot progress of the continuous-batching engine (effpot_tick). */
#define EFFPOT_SLOT_IDLE 0    /* no solve in this slot                           */
#define EFFPOT_SLOT_START 1   /* state uploaded: next tick zeroes the counters, runs poten + RK start */
#define EFFPOT_SLOT_RK 2      /* in the Runge-Kutta-Gill starter steps            */
#define EFFPOT_SLOT_PC 3      /* in the predictor-corrector phase                 */
#define EFFPOT_SLOT_DONE 4    /* converged (or out of steps): results valid        */
#define EFFPOT_SLOT_FAILED 5  /* non-finite density/energy                        */
typedef struct effpot_slot {
  int32_t phase;     /* EFFPOT_SLOT_*                                         */
  int32_t fp32;      /* 1 while in the FP32 phase of a MIXED context          */
  int32_t nrk;       /* RK steps done in the current phase                    */
  int32_t steps;     /* steps taken by this solve (all phases)                */
  int32_t since;     /* PC steps since the last energy evaluation             */
  int32_t converged; /* 1 if the FP64 tolerance was met                       */
  double eprev;      /* energy at the last evaluation                         */
  double rel_de;     /* last |dE/E| per step                                  */
  effpot_energies e; /* energies at the last evaluation                       */
} effpot_slot;

/* Header of a wavefunction snapshot (printoutc.f90:69-114 / readenc.f90:151-168):
 * the grid record, the impurity position and velocity, the impurity's internal
 * variables invar(ninvar) (complex) and, after it, psi(nx,ny,nz) complex(8). */
#define EFFPOT_SNAPSHOT_MAX_INVAR 16
typedef struct effpot_snapshot {
  int32_t nx, ny, nz, ninvar;
  double xmax, ymax, zmax, hx, hy, hz;
  double rimp[3], vimp[3];
  double invar[2 * EFFPOT_SNAPSHOT_MAX_INVAR]; /* interleaved complex, ninvar used */
  int32_t binary;   /* read: 1 if the file was unformatted (printoutc iprint < 0) */
  int32_t reserved_;
} effpot_snapshot;

typedef struct effpot_ctx effpot_ctx;

/* ---- context ----------------------------------------------------------- */
/* Replaces dimen() (dimen.f90:8), Init_deriv_p (init_deriv_parallel.f90:51-82),
 * fftini (fft.f90:6-55) and the one-time k-space kernels fforma/initcg/kalfs
 * (fforma.f90:11-60, initcg.f90:5-51, DFT4He3d.f90:396-400).  Selects HIP
 * device `device`. */
int effpot_ctx_create(const effpot_params* p, int device, effpot_ctx** out);
int effpot_ctx_destroy(effpot_ctx* ctx);
const char* effpot_last_error(const effpot_ctx* ctx); /* NULL: the calling thread's
                                                        last error of a call without
                                                        a context (snapshot I/O,
                                                        effpot_renormalize_noalloc) */
/* Device bytes held by a context with these parameters (admission control). */
int64_t effpot_footprint_bytes(const effpot_params* p);

/* ---- state I/O (readenc.f90:151-168 / printoutc.f90:69-114 payloads) --- */
int effpot_upload_psi(effpot_ctx* ctx, int slot, const double* psi);
int effpot_download_psi(effpot_ctx* ctx, int slot, double* psi);
int effpot_download_den(effpot_ctx* ctx, int slot, double* den);
/* Download a field of the last poten(): 0 pot4, 1 dencg, 2 delj4, 3 denalf,
 * 4 ualphas, 5 uimp, 6 dxden, 7 dyden, 8 dzden, 9 sto1 (correlation
 * convolution), 10 intxalf, 11 intyalf, 12 intzalf. */
int effpot_download_field(effpot_ctx* ctx, int slot, int which, double* out);
/* The FP32 mean field of a mixed-precision context (no reference
 * counterpart: the reference is FP64 throughout): effpot_poten_fp32 runs
 * poten (poten.f90:5-131, term_alfa.f90:5-109) on the FP32 fields from the
 * live psi of every active slot, as the FP32 phase of a mixed solve does;
 * its pot4 includes the slot's defect correction, which is zero until an
 * energy check of effpot_tick has set it.  effpot_download_field_fp32 reads
 * a field of it: `which` as in effpot_download_field, plus 13 = the defect
 * correction itself.  Other contexts get EFFPOT_ERR_ARG. */
int effpot_poten_fp32(effpot_ctx* ctx);
int effpot_download_field_fp32(effpot_ctx* ctx, int slot, int which, float* out);
/* Fermi droplet rho0/(1+exp((r-R)/afermi)) + noise*u(seed,i), psi=sqrt(rho)
 * normalised to n4, centred at (xc,yc,zc).  The reference promises this
 * start (readenc.f90:15-17) but never builds it.  Computed on the host so
 * that the CPU and GPU start from bit-identical input. */
int effpot_init_fermi(effpot_ctx* ctx, int slot, double rho0, double R,
                      double afermi, double noise, uint64_t seed,
                      const double centre[3]);

/* Atoms slot `slot` is normalised to (namelist n4, per slot; default
 * params.n4).  Set it before effpot_init_fermi for that slot. */
int effpot_set_natoms(effpot_ctx* ctx, int slot, double n4);

/* Time step of one slot for effpot_tick / effpot_solve (0: the options' dt).
 * Lets a server advance heterogeneous requests, and retry a request whose
 * mean field became too stiff for its step, without touching the others. */
int effpot_set_dt(effpot_ctx* ctx, int slot, double dt);

/* ---- impurity field (potenimpini.f90 / updatepoten+update+potenimp) ---- */
int effpot_set_impurity(effpot_ctx* ctx, int slot, const effpot_impurity* imp);

/* Impurity of one slot in a P or D internal state: tabulates the radial
 * potentials, runs instates on the slot's current density (instate >= 0),
 * writes the capped field uimp and keeps eso for the slot's energies
 * (eimpu = ekinx + eHeX + eso).  Replaces potenimpini + updatepoten +
 * instates + potenimp for Lstate 'P' / 'D'; effpot_set_impurity (isotropic)
 * resets eso to 0. */
int effpot_set_impurity_state(effpot_ctx* ctx, int slot, effpot_state_imp* imp);

/* The host half of instates: V + (Als/2) SO from the density moments (as in
 * effpot_state_imp.moments), its eigenvalues (ascending) and, for
 * instate >= 0, the selected invar and its eso.  Host only. */
int effpot_instates(int state, const double* moments, double als, int instate, double* invar, double* eigen,
                    double* eso);

/* effpot_instates with the selection `sel` (NULL: PAIR).  JZ returns an
 * invar fixed up to a global phase (the reference's eigensolver fixes that
 * phase its own way; no observable depends on it).  AVERAGE and MIXTURE add
 * eigenvectors of different levels, so their invar and the field built from
 * it depend on the eigensolver's phase and, within a Kramers pair, basis
 * conventions: this library's are those of its Jacobi solver.  Host only. */
int effpot_instates_select(int state, const double* moments, double als, int instate, const effpot_select* sel,
                           double* invar, double* eigen, double* eso);

/* Lexciplex_state_fix (instates.f90:54-113): the P-state moments of n_he He
 * atoms at fixed positions about the impurity instead of the density:
 * n_he = 2 ('Linear'): at +-r_exc on axis `axis` (0 x, 1 y, 2 z);
 * n_he = 6 ('Ring'): on the circle of radius r_exc normal to `axis`, the
 * first atom on the lower remaining axis.  Writes moments[0..6] (moment 0,
 * the atom count, is 0 as in the reference) for effpot_state_imp with
 * select.fixed_moments = 1.  Host only. */
int effpot_exciplex_moments(const effpot_plus* v_pi, const effpot_plus* v_sigma, int n_he, int axis, double r_exc,
                            double* moments);

/* ---- real-time propagation and a moving impurity ------------------------
 * Real time (timec = i; steppc.f90:53, steprk.f90:304-305): effpot_step_rk /
 * effpot_step_pc integrate d psi/dt = i (h2o2m4 lap - pot4 - uimp) psi with
 * the same integrators and no renormalisation.  Needs psi_complex = 1 and
 * EFFPOT_PREC_FP64; effpot_solve and effpot_tick refuse it (they relax). */
int effpot_set_real_time(effpot_ctx* ctx, int on);

/* The slot's impurity moves classically (an isotropic one, or a P/D state
 * whose invar evolves, effpot_set_state_dynamics): r and v advance with the
 * helium integrators (steprk.f90:117-150, steppc.f90:68-203) under forceimp,
 * F = -sum grad(den) uimp dV (potenimp.f90:193-258), a = F / m with
 * m = mass_u * mp_u (mp_u = 0.0207569277 K^-1 A^-2, modules.f90:333; v in
 * A per K^-1 of time), and uimp follows r.  Sets the initial acceleration from
 * the slot's current psi (DFT4He3d.f90:437-453).  mass_u <= 0 stops the
 * motion; effpot_set_impurity / effpot_set_impurity_state also do.  Energies
 * then hold ekinx = m |v|^2 / 2 (energy.f90:112) in eimpu.  FP64 contexts
 * only; relaxation (effpot_solve / effpot_tick) refuses it. */
int effpot_set_impurity_motion(effpot_ctx* ctx, int slot, const double* vimp, double mass_u);

/* Current r (A), v and a of the slot's impurity (v = a = 0 if it does not move). */
int effpot_get_impurity(effpot_ctx* ctx, int slot, double* r, double* v, double* a);

/* The invar of the slot's P/D-state impurity evolves in effpot_step_rk /
 * effpot_step_pc (the reference's default Lfix_lambda = .false.):
 * d invar/dt = -i (H(rho) invar - Ev0 invar), H(rho) the density-weighted
 * V + (Als/2) SO of instates recomputed after every poten (potinvar,
 * potenimp.f90:261-492), with the same RK-Gill / Hamming integrators
 * (steprk.f90:152-167, steppc.f90:84-102,205-226), and uimp and eso follow
 * invar.  ev0 = 0 takes Ev0 = eHeX + eso of the current state
 * (DFT4He3d.f90:444).  on = 0 freezes invar again.  FP64 contexts only;
 * relaxation (effpot_solve / effpot_tick) refuses it. */
int effpot_set_state_dynamics(effpot_ctx* ctx, int slot, int on, double ev0);

/* Current invar (complex(10), interleaved) of the slot's P/D-state impurity. */
int effpot_get_invar(effpot_ctx* ctx, int slot, double* invar);

/* X+-He potential of one slot for the uplus energy term, evaluated at the
 * slot's impurity position (energy.f90:146-163).  A context starts with no
 * X+ potential (uplus = 0); NULL removes it again. */
int effpot_set_plus(effpot_ctx* ctx, int slot, const effpot_plus* plus);

/* Fills `out` for a Select_Pot name (trailing blanks ignored): 'null',
 * 'Ba_plus_gs', 'Ba_plus_2D0', 'Ba_plus_2D1', 'Ba_plus_2D2' (the fits of
 * V_impur.f90:1218-1396; the reference's default selec_plus is 'Ba_plus_2D2'
 * with r_cutoff_plus 1.05, umax_plus 6.260015e4, V_impur.f90:14-16),
 * 'Na_plus_Koutselos', 'K_plus_Koutselos', 'Rb_plus_Koutselos' (the one
 * prep.sh:5-7 selects, with r_cutoff_plus 2.0, umax_plus 6.1105922888e3) and
 * 'Cs_plus_Koutselos' (V_impur.f90:232-292, 1704-1883).  An
 * unknown name gives EFFPOT_ERR_POT, where Select_Pot stops
 * (V_impur.f90:927-931).  Host only; needs no context. */
int effpot_plus_preset(const char* selec, double r_cutoff, double umax, effpot_plus* out);

/* z_exciplex_exclusion (energy.f90:149-157): grid points nearer than r to the
 * slot's impurity are left out of eHeX and uplus (so of E_imp); r = 0, the
 * reference's default and every slot's initial value, keeps every point.
 * The sphere follows a moving impurity. */
int effpot_set_exclusion(effpot_ctx* ctx, int slot, double r);

/* Lfilter_exciplex_force (potenimp.f90:216-245): the z component of a moving
 * impurity's force sums only the planes beyond the slot's exclusion radius on
 * the far side of the impurity (iz >= int((rz - z(1) + r)/hz + 1.5) for rz <
 * 0, else iz <= int((rz - z(1) - r)/hz + 1.5), 1-based), so an exciplex bound
 * next to the impurity does not push it.  The x and y components are
 * unchanged.  The reference's He-He correction term is commented out there
 * and is not built. */
int effpot_set_force_filter(effpot_ctx* ctx, int slot, int on);

/* The potential itself at distance r (A), in K, as the device evaluates it
 * (Select_Pot's role when Lcontrol prints it).  Host only. */
double effpot_plus_eval(const effpot_plus* plus, double r);

/* ---- operators on all active slots ------------------------------------ */
int effpot_poten(effpot_ctx* ctx);                         /* poten.f90:5   */
int effpot_energy(effpot_ctx* ctx, effpot_energies* out);  /* energy.f90:12, out[batch] */
int effpot_step_rk(effpot_ctx* ctx, double dt);            /* steprk.f90:1  */
/* STEPPC(deltat, errHe, errimp, errvimp, erriv) (steppc.f90:1): each output
 * is [batch] or NULL.  errHe = 9/121 sum|pc| / nxyz (:149-150); errimp and
 * errvimp = sum|9/121 pcr| / 3 and sum|9/121 pcv| / 3 of a moving impurity's
 * position and velocity (:179, :193); erriv = sum|9/121 pciv| / 6 (P) or / 10
 * (D) of an evolving invar (:213-218).  A fixed impurity or a frozen invar
 * gives 0, as Lfix_lambda does for erriv (:220). */
int effpot_step_pc(effpot_ctx* ctx, double dt, double* err_he, double* err_imp, double* err_vimp, double* err_iv);
int effpot_renormalize(effpot_ctx* ctx);                   /* psi *= sqrt(n4/N) */
/* effpot_renormalize on caller-owned device memory that allocates nothing.
 * No counterpart in the reference: it is for callers that cannot allocate
 * device memory (code under HIP graph capture, serving loops, harnesses that
 * forbid hipMalloc).  For each of `nslots` slots of `npoints` points
 * (slot-major; psi interleaved complex when psi_complex, else real):
 * psi *= sqrt(n4[s] / (sum |psi|^2 dxyz)) in place, den = |psi|^2 (den may be
 * NULL), natoms_out[s] = sum |psi|^2 dxyz before the scaling (NULL: not
 * written).  psi, den and natoms_out are device pointers; n4 is host memory
 * read during the call; `stream` is a hipStream_t (NULL: the null stream) and
 * the call only enqueues kernels on it.  Limits: nslots <= 64, npoints <=
 * 256^3; beyond them EFFPOT_ERR_ARG (effpot_last_error(NULL) says why).  The
 * norm's partial sums live in a static 264 KiB device array per GPU, held until
 * the process ends and shared by every call on that GPU: calls that overlap on
 * one GPU race.  The first call on a device looks up that array's address;
 * make it outside graph capture. */
int effpot_renormalize_noalloc(double* psi, double* den, int64_t npoints, int nslots, int psi_complex, double dxyz,
                               const double* n4, double* natoms_out, void* stream);
/* One 13-point derivative (derivnd, init_deriv_parallel.f90:1-46) of a real
 * host field f along axis iv (1..3), order id (1..2), into df (host). */
int effpot_derivnd(effpot_ctx* ctx, int id, int iv, const double* f, double* df);
/* 3-D transforms of fft.f90 on host buffers: r2c unnormalised into the
 * half spectrum (nx/2+1,ny,nz) interleaved; c2r scaled by 1/(nx ny nz). */
int effpot_fft_r2c(effpot_ctx* ctx, const double* in, double* out);
int effpot_fft_c2r(effpot_ctx* ctx, const double* in, double* out);
/* The same transforms with the FP32 kernels of a mixed-precision context
 * (the ones its FP32 poten runs); float buffers, half spectrum interleaved. */
int effpot_fft_r2c_fp32(effpot_ctx* ctx, const float* in, float* out);
int effpot_fft_c2r_fp32(effpot_ctx* ctx, const float* in, float* out);
/* Slots taking part in the operators above (1 active, 0 frozen). */
int effpot_set_active(effpot_ctx* ctx, int slot, int active);

/* ---- whole relaxation ---------------------------------------------------- */
/* RK-Gill starter + Hamming predictor-corrector in imaginary time with
 * renormalisation to n4 after every step, energy every `pener` steps, until
 * |dE/E|/pener < tol for every active slot.  out[batch]. */
int effpot_solve(effpot_ctx* ctx, const effpot_solve_opts* opt, effpot_result* out);

/* Benchmark hook (no reference counterpart): average duration in ms of one
 * launch of a hot-path kernel on the context's stream, measured with HIP
 * events over `reps` back-to-back launches on the current (live-precision)
 * state: 0 = H psi + predictor/modifier, 1 = H psi + corrector, 2 = the whole
 * poten() kernel sequence, 3 = the corrector timed inside `reps` whole
 * predictor-corrector steps (caches as in a relaxation), 4 = those whole
 * steps (predictor, poten, corrector, norm, renormalisation), 100 + k = launch k
 * of the poten() sequence alone (0-4 stage A: x r2c, y, z convolution, y,
 * x c2r; 5 derden; 6-10 stage B; 11 alpha_s divergence; 12-16 stage C;
 * 17 field assembly).  The launches
 * overwrite the integrator state: call it after the solves that matter. */
int effpot_kernel_time(effpot_ctx* ctx, int which, int reps, double* ms_per_launch);

/* Benchmark hook (no reference counterpart): the schedule of the forward x
 * and y transforms of poten() with FP32 (fp32_fields = 1) or FP64 fields:
 * 0 separate passes, 1 one plane launch with per-sub-group scratch, 3 one
 * plane launch in place (launches 1, 7 and 13 are then empty). */
int effpot_plane_fused(effpot_ctx* ctx, int fp32_fields);

/* Benchmark hook (no reference counterpart): 1 when poten() and
 * effpot_fft_c2r run the inverse y and x transforms as one in-place plane
 * launch per z plane (launches 3, 9 and 15 of the sequence are then empty),
 * 0 for separate passes; with FP32 (fp32_fields = 1) or FP64 fields. */
int effpot_plane_inverse_fused(effpot_ctx* ctx, int fp32_fields);

/* Benchmark hook (no reference counterpart): 1 when poten() keeps its half
 * spectra in the packed layout between the plane kernels (rows of nx/2 and
 * W = S(kx=0) + i S(kx=nx/2) in a block of its own), 0 for rows of nx/2+1
 * padded to a multiple of 16; with FP32 (fp32_fields = 1) or FP64 fields.
 * On by default where both in-place plane kernels run; the autotuner times
 * both layouts and EFFPOT_PACKED fixes it. */
int effpot_packed_layout(effpot_ctx* ctx, int fp32_fields);

/* Benchmark hook (no reference counterpart): threads per work-group of the
 * z convolutions of poten(), 256 (16 lines of the half spectrum per group)
 * or 512 (32 lines), as the autotuner or EFFPOT_ZCONV_TH left it. */
int effpot_zconv_threads(effpot_ctx* ctx);

/* Continuous batching: advance every slot whose phase is START, RK or PC by
 * one step of its own phase (a START slot gets poten + its first RK step).
 * Energies are FP64 in every mode; a slot is DONE at the first energy check
 * with FP64 fields where |dE/E| per step < tol, or after max_steps (plus
 * polish_steps for MIXED).  A MIXED slot advances with FP32 mean fields
 * (defect-corrected from the FP64 field at every check) and switches to FP64
 * fields (the polish) when |dE/E| per step < tol_fp32, when it is below
 * 2 tol and its trend would reach tol within 10 energy checks, or after
 * max_steps.
 * `slots[batch]` is host memory owned by the caller;
 * set a slot to START after effpot_init_fermi / effpot_upload_psi and
 * effpot_set_impurity.  No reference counterpart (the reference has no
 * driver loop); the solve of one slot is the same whatever the other slots do. */
int effpot_tick(effpot_ctx* ctx, const effpot_solve_opts* opt, effpot_slot* slots);

/* Autotuner (no reference counterpart): times the H psi + predictor,
 * corrector and poten kernels for every stencil tile (x width 64/32/16/8
 * dividing nx, z chunk 4/8/16/32) on the live state and keeps the fastest,
 * then times poten under each forward-transform schedule that applies
 * (effpot_plane_fused: 0, 1, 3; the environment variable EFFPOT_PLANE fixes
 * it instead) and keeps the fastest, and the same for the inverse schedule
 * (effpot_plane_inverse_fused; EFFPOT_IPLANE), the half-spectrum layout
 * (effpot_packed_layout; EFFPOT_PACKED) and the z-convolution group size
 * (effpot_zconv_threads; EFFPOT_ZCONV_TH).  Overwrites the
 * integrator state like effpot_kernel_time: call it before uploading states. */
int effpot_autotune(effpot_ctx* ctx, int reps, int* tx, int* zc);
/* Sets the stencil tile explicitly (x width, z chunk). */
int effpot_set_tiles(effpot_ctx* ctx, int tx, int zc);

/* Synchronise the context's stream. */
int effpot_sync(effpot_ctx* ctx);

/* ---- wavefunction snapshots (host I/O, no context) ------------------------
 * effpot_snapshot_read replaces readenc mode 3 (readenc.f90:151-168 with
 * titols.f90:12-37): it reads the formatted file (leading `#` lines skipped,
 * list-directed values: "(re,im)" complex, "k*" repeat counts, D exponents) or
 * the unformatted one (printoutc.f90:69-77, sequential records with 4-byte
 * markers), detected from the content.  psi (interleaved complex, nx*ny*nz)
 * may be NULL to read the header only.  effpot_snapshot_write replaces
 * printoutc's write (binary = 0 formatted, 1 unformatted); formatted values
 * are written with 17 significant digits, so a round trip is exact. */
int effpot_snapshot_read(const char* path, effpot_snapshot* hdr, double* psi);
int effpot_snapshot_write(const char* path, const effpot_snapshot* hdr, const double* psi, int binary);

#ifdef __cplusplus
}
#endif
#endif /* EFFPOT_H */
