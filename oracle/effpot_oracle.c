/*
 * effpot_oracle.c -- FP64 CPU restatement of the eff_pot hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the correctness oracle for the
 * HIP library (eff_pot_amd/csrc).  Only tests/, __graft_entry__.smoke() and
 * bench.py's cpu_baseline leg may load it, and only as the checker or the
 * timed CPU baseline; the product path never calls it.
 *
 * Every routine restates one reference routine (cited file:line, paths
 * relative to the reference's src/).  Deliberate differences from the
 * reference, all shared with the HIP library:
 *   - imaginary time: H psi = h2o2m4*lap(psi) - (pot4 + uimp)*psi, i.e.
 *     steppc.f90:53 / steprk.f90:304-305 with timec=1 and the -i*uimp term
 *     made real (timec is never initialised by the reference, dimen.f90:145);
 *   - psi is renormalised to n4 after every full step (the reference's
 *     integrators have no renormalisation), and the mean field inside an
 *     imaginary-time step is that of the stage or predicted density scaled
 *     to n4 atoms (poten_of_step);
 *   - den = re^2 + im^2 (the reference squares abs());
 *   - the FFT is this file's own mixed-radix (2,3,4,5) Stockham transform
 *     (the reference calls FFTW/MKL, fft.f90:41-55);
 *   - Si(x) is this file's own series / continued fraction (the reference
 *     uses a NAG-style Chebyshev S13ADF, s13adf.f90:2-183);
 *   - energy() zeroes eHeX/uplus before accumulating (energy.f90:146-163
 *     accumulates into stale values, safe only for one call).
 *
 * Pins (tests/test_oracle_*.py): numpy.fft, Fornberg closed forms and the
 * reference's own 13-point tables (mates.f:786-879), scipy.special.sici and
 * scipy quadrature of the LJ transform, the uniform-density known answer,
 * analytic Gaussian kinetic energy, RK-Gill / Hamming amplification factors,
 * and the variational identity sum(field * drho) dV = d/de E(rho + e drho)
 * term by term (elj4/delj4, ecor4/correlation + sto1, ealphas/ualphas,
 * esolid/penalty) on non-uniform droplets, which pins term_alfa and the
 * correlation convolution (tests/test_oracle_variational.py), and the
 * dt-independence of the relaxation's trajectory in imaginary time, which
 * pins the density scaling of poten_of_step (tests/test_oracle_relaxation.py).
 */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "../include/effpot.h"

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

typedef struct { double re, im; } cplx;

/* ------------------------------------------------------------------------ */
/* 1-D mixed radix complex FFT (Stockham autosort, radices 4,2,3,5).          */
/* ------------------------------------------------------------------------ */
typedef struct {
  int n, nf, fac[16];
  cplx* tw; /* tw[k] = exp(-2 pi i k / n) */
} fftplan;

static int plan_init(fftplan* pl, int n) {
  pl->n = n; pl->nf = 0;
  int m = n;
  while (m % 4 == 0) { pl->fac[pl->nf++] = 4; m /= 4; }
  while (m % 2 == 0) { pl->fac[pl->nf++] = 2; m /= 2; }
  while (m % 3 == 0) { pl->fac[pl->nf++] = 3; m /= 3; }
  while (m % 5 == 0) { pl->fac[pl->nf++] = 5; m /= 5; }
  if (m != 1) return -1;
  pl->tw = (cplx*)malloc(sizeof(cplx) * n);
  for (int k = 0; k < n; k++) {
    double a = -2.0 * M_PI * (double)k / (double)n;
    pl->tw[k].re = cos(a); pl->tw[k].im = sin(a);
  }
  return 0;
}

/* In-place transform of x[0..n-1] using scratch w[0..n-1].
 * sign = -1 forward, +1 backward (unnormalised).  Each pass with radix r and
 * running product p: butterfly i (0 <= i < n/r) reads x[i + j n/r], applies
 * the twiddle w_{p r}^{j k} (k = i mod p), a length-r DFT, and writes
 * y[(i - k) r + k + t p]. */
static void fft1d(const fftplan* pl, cplx* x, cplx* w, int sign) {
  const int n = pl->n;
  cplx *src = x, *dst = w;
  int p = 1;
  for (int f = 0; f < pl->nf; f++) {
    const int r = pl->fac[f], m = n / r, stride = n / (p * r);
    for (int i = 0; i < m; i++) {
      const int k = i % p;
      cplx v[5];
      for (int j = 0; j < r; j++) {
        cplx a = src[i + j * m];
        int t = (j * k * stride) % n;
        double c = pl->tw[t].re, s = sign < 0 ? pl->tw[t].im : -pl->tw[t].im;
        v[j].re = a.re * c - a.im * s;
        v[j].im = a.re * s + a.im * c;
      }
      const int base = (i - k) * r + k;
      for (int t = 0; t < r; t++) {
        double sr = 0.0, si = 0.0;
        for (int j = 0; j < r; j++) {
          int e = ((j * t) % r) * (n / r);
          double c = pl->tw[e].re, s = sign < 0 ? pl->tw[e].im : -pl->tw[e].im;
          sr += v[j].re * c - v[j].im * s;
          si += v[j].re * s + v[j].im * c;
        }
        dst[base + t * p].re = sr;
        dst[base + t * p].im = si;
      }
    }
    p *= r;
    cplx* tmp = src; src = dst; dst = tmp;
  }
  if (src != x) memcpy(x, src, sizeof(cplx) * n);
}

/* ------------------------------------------------------------------------ */
/* Oracle state: the module arrays of modules.f90 / dimen.f90 it needs.     */
/* ------------------------------------------------------------------------ */
typedef struct oracle {
  effpot_params p;
  int nx, ny, nz, nxh;
  size_t N, M;
  double hx, hy, hz, dxyz, renor;
  double *x, *y, *z;
  double c1[13], c2[13]; /* centred 13-point coefficients, unscaled */
  fftplan plx, ply, plz;
  double *wcgk, *fvlj4, *kalfs; /* k-space kernels, size M */
  /* fields */
  double *psi, *den, *pot4, *uimp, *dencg, *delj4, *denalf, *ualphas;
  double *dxden, *dyden, *dzden, *falfs, *intx, *inty, *intz, *sto1;
  double *rw1, *rw2, *rw3, *rw4, *rw5, *rw6; /* real scratch */
  cplx *fden, *wk1, *wk2;                    /* half spectra */
  double *psiold[3], *hpsiold[2], *pc, *q, *lap, *hpsi; /* complex (2N) */
  int ioldp[3], ioldh[2];
  /* impurity table (potenimpini.f90:11-50) */
  int npot; double delinter; double* vtab; /* vtab[0..npot+1] */
  effpot_impurity imp; int have_imp;
  effpot_plus plus; /* X+ potential of uplus (kind NULL: none) */
  double excl;      /* z_exciplex_exclusion (energy.f90:155) */
  int ffilter;      /* Lfilter_exciplex_force (potenimp.f90:216-245) */
  double eso;       /* spin-orbit energy of a P/D-state impurity */
  int rt;           /* real time: the step integrates i H psi */
  /* moving impurity (classicimp): r = imp.rimp, v, a, histories, pc, q */
  int moving; double mimp, vimp[3], aimp[3];
  double rimpold[3][3], vimpold[3][3], aimpold[2][3], pcr[3], pcv[3], qr[3], qv[3];
  int ioldr[3], ioldv[3], iolda[2];
  /* evolving invar of a P/D state (potinvar + the invar integrators) */
  int dyn; double ev0, hiv[20], ivold[3][20], hivold[2][20], pciv[20], qiv[20];
  int iiv[3], ihiv[2];
  effpot_state_imp simp; /* the state's set-up; simp.invar is the current invar */
  double *stab, *sf[3]; /* P/D state: potential tables [3][npot+2], updatepoten fields */
} oracle_t;

static int tstgrid(int n) { /* tstgrid.f90:8-12: n = 2^a 3^b 5^c <= 256 */
  if (n < 16 || n > 256) return 1;
  while (n % 2 == 0) n /= 2;
  while (n % 3 == 0) n /= 3;
  while (n % 5 == 0) n /= 5;
  return n != 1;
}

#define IDX(o, ix, iy, iz) ((size_t)(ix) + (size_t)(o)->nx * ((size_t)(iy) + (size_t)(o)->ny * (size_t)(iz)))

/* Centred 13-point coefficients of the first and second derivative.  Each is
 * one division of exactly representable integers (Fornberg's closed form), so
 * it is the correctly rounded value of the rational in the reference's
 * Bickley table (mates.f:1117-1125 for K=1, mates.f:1202-1210 for K=2). */
void oracle_fd_coeffs(int id, double* c) {
  const double p2 = 518400.0; /* (6!)^2 */
  static const double fact[13] = {1, 1, 2, 6, 24, 120, 720, 5040, 40320, 362880,
                                  3628800, 39916800, 479001600};
  for (int m = -6; m <= 6; m++) {
    int am = m < 0 ? -m : m;
    if (m == 0) {
      c[6] = id == 1 ? 0.0 : -5369.0 / 1800.0; /* -2 sum_{m=1}^6 1/m^2 */
      continue;
    }
    double den = fact[6 - am] * fact[6 + am];
    double sgn = (am % 2 == 1) ? 1.0 : -1.0;
    if (id == 1) {
      double v = p2 / ((double)am * den);
      c[m + 6] = (m > 0 ? sgn : -sgn) * v;
    } else {
      double v = 2.0 * p2 / ((double)(am * am) * den);
      c[m + 6] = sgn * v;
    }
  }
}

/* ------------------------------------------------------------------------ */
/* Sine integral Si(x): power series for |x| <= 4, otherwise pi/2 + Im E1(ix) */
/* with E1 from its continued fraction (modified Lentz).                      */
/* ------------------------------------------------------------------------ */
double oracle_si(double x) {
  double ax = fabs(x);
  if (ax == 0.0) return 0.0;
  double res;
  if (ax <= 4.0) {
    double term = ax, sum = ax, x2 = ax * ax;
    for (int k = 1; k < 60; k++) {
      term *= -x2 / ((2.0 * k) * (2.0 * k + 1.0));
      double add = term / (2.0 * k + 1.0);
      sum += add;
      if (fabs(add) < 1e-18 * fabs(sum)) break;
    }
    res = sum;
  } else {
    /* E1(z) = exp(-z) / (z + 1 - 1/(z + 3 - 4/(z + 5 - ...))), z = i ax */
    double br = 1.0, bi = ax;           /* b = z + 1 */
    double cr = 1e300, ci = 0.0;        /* c = 1/tiny */
    double dr, di;                      /* d = 1/b */
    double dd = br * br + bi * bi;
    dr = br / dd; di = -bi / dd;
    double hr = dr, hi = di;
    for (int i = 1; i < 1000; i++) {
      double an = -(double)i * (double)i;
      br += 2.0; /* b += 2 */
      /* d = 1/(an*d + b) */
      double tr = an * dr + br, ti = an * di + bi;
      dd = tr * tr + ti * ti;
      dr = tr / dd; di = -ti / dd;
      /* c = b + an/c */
      double cc = cr * cr + ci * ci;
      double ir = cr / cc, ii = -ci / cc;
      cr = br + an * ir; ci = bi + an * ii;
      /* del = c*d ; h *= del */
      double delr = cr * dr - ci * di, deli = cr * di + ci * dr;
      double nr = hr * delr - hi * deli, ni = hr * deli + hi * delr;
      hr = nr; hi = ni;
      if (fabs(delr - 1.0) + fabs(deli) < 1e-16) break;
    }
    /* E1 = (cos ax - i sin ax) * h */
    double c = cos(ax), s = sin(ax);
    double ei = c * hi - s * hr;
    res = 0.5 * M_PI + ei;
  }
  return x < 0 ? -res : res;
}

/* fforma.f90:73-102 (sinx11) and :108-137 (sinx5), with sinint (:143-196). */
static double sinint(double a, double x) {
  double s = oracle_si(a * x);
  return a > 0.0 ? 0.5 * M_PI - s : 0.5 * M_PI + s;
}
static double sinx11(double a, double x) {
  double y = a * x, y2 = y * y;
  double d[5] = {1.0, y2, y2 * y2, y2 * y2 * y2, y2 * y2 * y2 * y2};
  static const double b[5] = {40320.0, -720.0, 24.0, -2.0, 1.0};
  static const double c[5] = {362880.0, -5040.0, 120.0, -6.0, 1.0};
  double bd = 0.0, cd = 0.0;
  for (int i = 0; i < 5; i++) { bd += b[i] * d[i]; cd += c[i] * d[i]; }
  double x9 = pow(x, 9.0);
  double res1 = (a * bd * cos(y) + cd * sin(y) / x) / (3628800.0 * x9);
  double res2 = -(pow(a, 10.0) * sinint(a, x)) / 3628800.0;
  return res1 + res2;
}
static double sinx5(double a, double x) {
  double y = a * x, y2 = y * y;
  double bd = -2.0 + 1.0 * y2, cd = -6.0 + 1.0 * y2;
  double res1 = -(a * bd * cos(y) + cd * sin(y) / x) / (24.0 * x * x * x);
  double res2 = (pow(a, 4.0) * sinint(a, x)) / 24.0;
  return res1 + res2;
}
/* p = 0 limit of the OT-core LJ transform: 4 pi 4 eps (s^12/9h^9 - s^6/3h^3)
 * (fforma.f90:256-280, core OT). */
double oracle_bforce(double eps, double sigma, double hc) {
  double pi4 = 4.0 * M_PI, eps4 = 4.0 * eps;
  double s6 = pow(sigma, 6.0), s12 = s6 * s6;
  return pi4 * eps4 * s12 / (9.0 * pow(hc, 9.0)) - pi4 * eps4 * s6 / (3.0 * hc * hc * hc);
}
/* fforma.f90:40-60: transform of the OT-screened LJ at |p| (1/A). */
double oracle_fvlj(double p, double eps, double sigma, double hc) {
  if (fabs(p) < 1e-10) return oracle_bforce(eps, sigma, hc);
  double s6 = pow(sigma, 6.0);
  double a = 2.0 * M_PI * p;
  double t1 = sinx11(a, hc), t2 = sinx5(a, hc);
  return (8.0 * eps * s6 * (s6 * t1 - t2)) / p;
}
/* initcg.f90:32-49 */
double oracle_wcg(double p, double h) {
  if (!(p > 0.0)) return 1.0;
  double arg = 2.0 * M_PI * p * h;
  double aux = sin(arg) - arg * cos(arg);
  if (fabs(aux) < 1e-10) return 1.0;
  return 3.0 * aux / (arg * arg * arg);
}

/* Frequency of index i on an axis of n points with spacing h
 * (DFT4He3d.f90:269-290). */
static double pfreq(int i, int n, double h) {
  double hp = 1.0 / (n * h), pmax = 1.0 / (2.0 * h);
  if (i <= n / 2) return hp * i;
  return -pmax + hp * (i - n / 2);
}

/* ------------------------------------------------------------------------ */
/* Creation / destruction                                                     */
/* ------------------------------------------------------------------------ */
static double* dalloc(size_t n) { return (double*)calloc(n ? n : 1, sizeof(double)); }

int oracle_create(const effpot_params* p, oracle_t** out) {
  if (!p || !out) return EFFPOT_ERR_ARG;
  if (tstgrid(p->nx) || tstgrid(p->ny) || tstgrid(p->nz)) return EFFPOT_ERR_GRID;
  if (p->core != EFFPOT_CORE_OT && p->core != EFFPOT_CORE_OTC) return EFFPOT_ERR_CORE;
  if (p->npd != 13) return EFFPOT_ERR_ARG;
  if (p->icon != 13 && p->icon != 8) return EFFPOT_ERR_ARG;
  oracle_t* o = (oracle_t*)calloc(1, sizeof(oracle_t));
  o->p = *p;
  o->nx = p->nx; o->ny = p->ny; o->nz = p->nz; o->nxh = p->nx / 2 + 1;
  o->N = (size_t)o->nx * o->ny * o->nz;
  o->M = (size_t)o->nxh * o->ny * o->nz;
  o->hx = 2.0 * fabs(p->xmax) / o->nx;
  o->hy = 2.0 * fabs(p->ymax) / o->ny;
  o->hz = 2.0 * fabs(p->zmax) / o->nz;
  o->dxyz = o->hx * o->hy * o->hz;
  o->renor = 1.0 / ((double)o->nx * o->ny * o->nz);
  o->x = dalloc(o->nx); o->y = dalloc(o->ny); o->z = dalloc(o->nz);
  for (int i = 0; i < o->nx; i++) o->x[i] = -p->xmax + o->hx * i;
  for (int i = 0; i < o->ny; i++) o->y[i] = -p->ymax + o->hy * i;
  for (int i = 0; i < o->nz; i++) o->z[i] = -p->zmax + o->hz * i;
  oracle_fd_coeffs(1, o->c1);
  oracle_fd_coeffs(2, o->c2);
  plan_init(&o->plx, o->nx); plan_init(&o->ply, o->ny); plan_init(&o->plz, o->nz);
  size_t N = o->N, M = o->M;
  o->wcgk = dalloc(M); o->fvlj4 = dalloc(M); o->kalfs = dalloc(M);
  /* k-space kernels: fforma (DFT4He3d.f90:385), initcg (:393), kalfs (:396-400) */
  for (int iz = 0; iz < o->nz; iz++) {
    double pz = pfreq(iz, o->nz, o->hz);
    for (int iy = 0; iy < o->ny; iy++) {
      double py = pfreq(iy, o->ny, o->hy);
      for (int ix = 0; ix < o->nxh; ix++) {
        double px = o->hx > 0 ? (1.0 / (o->nx * o->hx)) * ix : 0.0;
        double pm = sqrt(px * px + py * py + pz * pz);
        size_t k = (size_t)ix + (size_t)o->nxh * ((size_t)iy + (size_t)o->ny * iz);
        o->fvlj4[k] = oracle_fvlj(pm, p->eps4, p->sigma4, p->h_core);
        o->wcgk[k] = oracle_wcg(pm, p->h_core);
        double t = M_PI * p->l * pm;
        o->kalfs[k] = exp(-(t * t));
      }
    }
  }
  o->psi = dalloc(2 * N); o->den = dalloc(N); o->pot4 = dalloc(N); o->uimp = dalloc(N);
  o->dencg = dalloc(N); o->delj4 = dalloc(N); o->denalf = dalloc(N); o->ualphas = dalloc(N);
  o->dxden = dalloc(N); o->dyden = dalloc(N); o->dzden = dalloc(N); o->falfs = dalloc(N);
  o->intx = dalloc(N); o->inty = dalloc(N); o->intz = dalloc(N); o->sto1 = dalloc(N);
  o->rw1 = dalloc(N); o->rw2 = dalloc(N); o->rw3 = dalloc(N);
  o->rw4 = dalloc(N); o->rw5 = dalloc(N); o->rw6 = dalloc(N);
  o->fden = (cplx*)dalloc(2 * M); o->wk1 = (cplx*)dalloc(2 * M); o->wk2 = (cplx*)dalloc(2 * M);
  for (int i = 0; i < 3; i++) o->psiold[i] = dalloc(2 * N);
  for (int i = 0; i < 2; i++) o->hpsiold[i] = dalloc(2 * N);
  o->pc = dalloc(2 * N); o->q = dalloc(2 * N); o->lap = dalloc(2 * N); o->hpsi = dalloc(2 * N);
  for (int i = 0; i < 3; i++) o->ioldp[i] = i;
  for (int i = 0; i < 2; i++) o->ioldh[i] = i;
  /* potenimpini.f90:11-14 */
  int nmax = o->nx > o->ny ? o->nx : o->ny; if (o->nz > nmax) nmax = o->nz;
  double xm = p->xmax > p->ymax ? p->xmax : p->ymax; if (p->zmax > xm) xm = p->zmax;
  o->npot = 100 * nmax + 1;
  o->delinter = 4.0 * xm / (double)(o->npot - 1);
  o->vtab = dalloc(o->npot + 2);
  *out = o;
  return EFFPOT_OK;
}

void oracle_destroy(oracle_t* o) {
  if (!o) return;
  double* arrs[] = {o->x, o->y, o->z, o->wcgk, o->fvlj4, o->kalfs, o->psi, o->den, o->pot4,
                    o->uimp, o->dencg, o->delj4, o->denalf, o->ualphas, o->dxden, o->dyden,
                    o->dzden, o->falfs, o->intx, o->inty, o->intz, o->sto1, o->rw1, o->rw2,
                    o->rw3, o->rw4, o->rw5, o->rw6, (double*)o->fden, (double*)o->wk1,
                    (double*)o->wk2, o->psiold[0], o->psiold[1], o->psiold[2],
                    o->hpsiold[0], o->hpsiold[1], o->pc, o->q, o->lap, o->hpsi, o->vtab,
                    o->stab, o->sf[0], o->sf[1], o->sf[2]};
  for (size_t i = 0; i < sizeof(arrs) / sizeof(arrs[0]); i++) free(arrs[i]);
  free(o->plx.tw); free(o->ply.tw); free(o->plz.tw);
  free(o);
}

size_t oracle_size(const oracle_t* o, int half) { return half ? o->M : o->N; }

/* ------------------------------------------------------------------------ */
/* 3-D FFTs (fft.f90): r2c unnormalised, c2r unnormalised (callers scale).  */
/* ------------------------------------------------------------------------ */
static void fft_axis_yz(oracle_t* o, cplx* a, int axis, int sign) {
  const int nxh = o->nxh, ny = o->ny, nz = o->nz;
  const fftplan* pl = axis == 1 ? &o->ply : &o->plz;
  const int n = pl->n;
  const size_t stride = axis == 1 ? (size_t)nxh : (size_t)nxh * ny;
  const int nouter = axis == 1 ? nz : ny;
#pragma omp parallel
  {
    cplx* b = (cplx*)malloc(sizeof(cplx) * n);
    cplx* w = (cplx*)malloc(sizeof(cplx) * n);
#pragma omp for collapse(2) schedule(static)
    for (int outer = 0; outer < nouter; outer++)
      for (int ix = 0; ix < nxh; ix++) {
        size_t base = axis == 1 ? (size_t)ix + (size_t)nxh * ny * outer
                                : (size_t)ix + (size_t)nxh * outer;
        for (int i = 0; i < n; i++) b[i] = a[base + stride * i];
        fft1d(pl, b, w, sign);
        for (int i = 0; i < n; i++) a[base + stride * i] = b[i];
      }
    free(b); free(w);
  }
}

void oracle_fft_r2c(oracle_t* o, const double* in, double* outd) {
  cplx* out = (cplx*)outd;
  const int nx = o->nx, nxh = o->nxh, ny = o->ny, nz = o->nz;
#pragma omp parallel
  {
    cplx* b = (cplx*)malloc(sizeof(cplx) * nx);
    cplx* w = (cplx*)malloc(sizeof(cplx) * nx);
#pragma omp for collapse(2) schedule(static)
    for (int iz = 0; iz < nz; iz++)
      for (int iy = 0; iy < ny; iy++) {
        const double* row = in + IDX(o, 0, iy, iz);
        for (int i = 0; i < nx; i++) { b[i].re = row[i]; b[i].im = 0.0; }
        fft1d(&o->plx, b, w, -1);
        cplx* orow = out + (size_t)nxh * ((size_t)iy + (size_t)ny * iz);
        for (int i = 0; i < nxh; i++) orow[i] = b[i];
      }
    free(b); free(w);
  }
  fft_axis_yz(o, out, 1, -1);
  fft_axis_yz(o, out, 2, -1);
}

/* c2r: the half spectrum is destroyed (like FFTW's c2r). */
static void fft_c2r_inplace(oracle_t* o, cplx* in, double* out) {
  const int nx = o->nx, nxh = o->nxh, ny = o->ny, nz = o->nz;
  fft_axis_yz(o, in, 2, +1);
  fft_axis_yz(o, in, 1, +1);
#pragma omp parallel
  {
    cplx* b = (cplx*)malloc(sizeof(cplx) * nx);
    cplx* w = (cplx*)malloc(sizeof(cplx) * nx);
#pragma omp for collapse(2) schedule(static)
    for (int iz = 0; iz < nz; iz++)
      for (int iy = 0; iy < ny; iy++) {
        const cplx* irow = in + (size_t)nxh * ((size_t)iy + (size_t)ny * iz);
        for (int i = 0; i < nxh; i++) b[i] = irow[i];
        /* self-conjugate bins: their imaginary parts are dropped, as c2r does */
        b[0].im = 0.0;
        if (nx % 2 == 0) b[nx / 2].im = 0.0;
        for (int i = nxh; i < nx; i++) { b[i].re = irow[nx - i].re; b[i].im = -irow[nx - i].im; }
        fft1d(&o->plx, b, w, +1);
        double* orow = out + IDX(o, 0, iy, iz);
        for (int i = 0; i < nx; i++) orow[i] = b[i].re;
      }
    free(b); free(w);
  }
}

void oracle_fft_c2r(oracle_t* o, const double* ind, double* out) {
  memcpy(o->wk2, ind, sizeof(cplx) * o->M);
  fft_c2r_inplace(o, o->wk2, out);
  for (size_t i = 0; i < o->N; i++) out[i] *= o->renor;
}

/* IFFT(fden*K) * renor  (fft.f90:94-144 fftbk_* after a k-space product). */
static void conv_back(oracle_t* o, const cplx* F, const double* K, double* out) {
  cplx* w = o->wk1;
  for (size_t k = 0; k < o->M; k++) { w[k].re = F[k].re * K[k]; w[k].im = F[k].im * K[k]; }
  fft_c2r_inplace(o, w, out);
  for (size_t i = 0; i < o->N; i++) out[i] *= o->renor;
}

/* ------------------------------------------------------------------------ */
/* 13-point derivatives (derivnd -> Deriv3D_p / Deriv3D_pc).                 */
/* ------------------------------------------------------------------------ */
/* ncomp = 1 real, 2 interleaved complex.  icon 13: centred stencil with the
 * field taken as zero outside [1,n] (redef.f:133-170 boundary rows);
 * icon 8: periodic wrap (init_deriv_parallel.f90:529-634). */
static void deriv(const oracle_t* o, int id, int iv, const double* f, double* df, int ncomp) {
  const int nx = o->nx, ny = o->ny, nz = o->nz;
  const double h = iv == 1 ? o->hx : (iv == 2 ? o->hy : o->hz);
  const double sto = id == 1 ? 1.0 / h : 1.0 / (h * h);
  const double* c0 = id == 1 ? o->c1 : o->c2;
  double cd[13];
  for (int m = 0; m < 13; m++) cd[m] = c0[m] * sto;
  const int n = iv == 1 ? nx : (iv == 2 ? ny : nz);
  const size_t s = iv == 1 ? 1 : (iv == 2 ? (size_t)nx : (size_t)nx * ny);
  const int periodic = o->p.icon == 8;
#pragma omp parallel for collapse(2) schedule(static)
  for (int iz = 0; iz < nz; iz++)
    for (int iy = 0; iy < ny; iy++)
      for (int ix = 0; ix < nx; ix++) {
        size_t c = IDX(o, ix, iy, iz);
        int i = iv == 1 ? ix : (iv == 2 ? iy : iz);
        for (int comp = 0; comp < ncomp; comp++) {
          double acc = 0.0;
          for (int m = -6; m <= 6; m++) {
            int j = i + m;
            if (j < 0 || j >= n) {
              if (!periodic) continue;
              j = (j + n) % n;
            }
            acc += cd[m + 6] * f[(c + (size_t)((long)(j - i) * (long)s)) * ncomp + comp];
          }
          df[c * ncomp + comp] = acc;
        }
      }
}

void oracle_derivnd(oracle_t* o, int id, int iv, const double* f, double* df) { deriv(o, id, iv, f, df, 1); }
void oracle_derivnd_c(oracle_t* o, int id, int iv, const double* f, double* df) { deriv(o, id, iv, f, df, 2); }

/* ------------------------------------------------------------------------ */
/* Impurity: Select_Pot, potenimpini, updatepoten/update, potenimp.          */
/* ------------------------------------------------------------------------ */
double oracle_select_pot(const effpot_impurity* imp, double r) {
  if (imp->kind == EFFPOT_POT_NULL) return 0.0; /* V_null on both sides of r_cutoff, V_impur.f90:171-176 */
  if (r < imp->r_cutoff) return imp->umax;
  double e = exp(-imp->a * (r - imp->re));
  return imp->D * (e * e - 2.0 * e);
}

/* update() with Lparab (potenimp.f90:588-632): the parabola through the
 * nearest table node and its two neighbours. */
double oracle_update(double r, int npot, double rmax, const double* f) {
  double h = rmax / (double)(npot - 1);
  double c1oh = 1.0 / h, c1o2h2 = c1oh * c1oh * 0.5;
  int i0 = (int)(1.5 + r / h);
  if (i0 > npot) i0 = npot; /* r beyond the table: clamp (the reference reads past it) */
  double r0 = (double)(i0 - 1) * h;
  double f0 = f[i0];
  int i1, i2;
  if (r < r0) { i1 = i0 - 1; i2 = i0 + 1; } else { i1 = i0 + 1; i2 = i0 - 1; }
  double f1 = f[i1], f2 = f[i2];
  double aa = (f2 - 2.0 * f0 + f1) * c1o2h2;
  double bb = (r < r0) ? (f2 - f0) * c1oh - aa * (2.0 * r0 + h) : (f1 - f0) * c1oh - aa * (2.0 * r0 + h);
  double cc = f0 - r0 * (bb + r0 * aa);
  return r * (aa * r + bb) + cc;
}

int oracle_set_impurity(oracle_t* o, const effpot_impurity* imp) {
  if (imp->kind != EFFPOT_POT_MORSE && imp->kind != EFFPOT_POT_NULL) return EFFPOT_ERR_POT;
  o->imp = *imp; o->have_imp = 1; o->eso = 0.0; o->moving = 0; o->dyn = 0;
  /* potenimpini.f90:35-58 (isotropic: V_pi = V, Delta = 0) */
  for (int i = 1; i <= o->npot; i++) o->vtab[i] = oracle_select_pot(imp, (double)(i - 1) * o->delinter);
  o->vtab[0] = o->vtab[1];
  o->vtab[o->npot + 1] = o->vtab[o->npot];
  double rmax = (double)(o->npot - 1) * o->delinter;
  const double* rr = imp->rimp;
#pragma omp parallel for schedule(static)
  for (int iz = 0; iz < o->nz; iz++) {
    double zt = (o->z[iz] - rr[2]) * (o->z[iz] - rr[2]);
    for (int iy = 0; iy < o->ny; iy++) {
      double yt = (o->y[iy] - rr[1]) * (o->y[iy] - rr[1]) + zt;
      for (int ix = 0; ix < o->nx; ix++) {
        double r2 = (o->x[ix] - rr[0]) * (o->x[ix] - rr[0]) + yt;
        double r = sqrt(r2);
        double u = oracle_update(r, o->npot, rmax, o->vtab);
        o->uimp[IDX(o, ix, iy, iz)] = u < imp->cap ? u : imp->cap; /* potenimp.f90:184-186 */
      }
    }
  }
  return EFFPOT_OK;
}

/* Select_Pot(selec_plus, r, r_cutoff_plus, umax_plus) for the uplus term
 * (V_impur.f90:87): 'null' (:171-176), the Koutselos ion form (:258-262,
 * :1796-1839, FP64 as the reference evaluates it) or an inverse-power fit of the
 * V_Ba_plus_* form (:1238-1261): plateau, parabola, then the series summed
 * as the reference does, Aux = xq^(k0+1), Sto += pg(k)*Aux, Aux *= xq, in
 * REAL*16 (__float128 here; the coefficient is pg + pg_lo). */
static double koutselos(const double* c, double xx) {
  /* V_Koutselos_*_plus_gs (V_impur.f90:1796-1839 for Rb+): atomic units inside */
  double x = xx / c[EFFPOT_KOUT_ABOHR];
  double rs = x / c[EFFPOT_KOUT_RHO];
  double v_ex = c[EFFPOT_KOUT_V0] * (c[EFFPOT_KOUT_A] * exp(-c[EFFPOT_KOUT_ALPHA] * rs) -
                                     c[EFFPOT_KOUT_B] * exp(-c[EFFPOT_KOUT_BETA] * rs));
  double aux = c[EFFPOT_KOUT_RM] * 1.28, h_r;
  if (x >= aux) {
    h_r = 1.0;
  } else {
    h_r = 0.0;
    if (x != 0.0) h_r = exp(-pow(aux / x - 1.0, 2));
  }
  double sto = v_ex;
  if (x != 0.0) {
    double x2 = x * x, x4 = x2 * x2, x6 = x4 * x2, x8 = x6 * x2;
    sto = sto + (-c[EFFPOT_KOUT_CC4] / x4 - c[EFFPOT_KOUT_CC6] / x6 - c[EFFPOT_KOUT_CC8] / x8) * h_r;
  }
  return sto * c[EFFPOT_KOUT_HARTREE] * c[EFFPOT_KOUT_EVK];
}

double oracle_plus_pot(const effpot_plus* p, double r) {
  if (p->kind == EFFPOT_POT_KOUTSELOS) return r < p->r_cutoff ? p->umax : koutselos(p->pg, r);
  if (p->kind != EFFPOT_POT_SERIES) return 0.0;
  if (r < p->r_cutoff) return p->umax;
  if (r <= p->rcutoff) return p->fcutoff;
  if (r <= p->r0) return fabs(p->aa * (r * r) + p->bb * r + p->cc); /* aa*x**2+bb*x+cc */
  __float128 xq = (__float128)1.0 / (__float128)r, aux = 1, sto = 0;
  for (int k = 0; k <= p->k0; k++) aux *= xq;
  for (int k = 0; k < p->nterms; k++) {
    sto += ((__float128)p->pg[k] + (__float128)p->pg_lo[k]) * aux;
    aux *= xq;
  }
  return (double)sto;
}

void oracle_set_plus(oracle_t* o, const effpot_plus* p) {
  if (p) o->plus = *p;
  else memset(&o->plus, 0, sizeof(o->plus));
}

/* z_exciplex_exclusion of energy.f90:155 (0: every point) */
void oracle_set_exclusion(oracle_t* o, double r) { o->excl = r; }
void oracle_set_force_filter(oracle_t* o, int on) { o->ffilter = on; }

/* ------------------------------------------------------------------------ */
/* P/D internal state: potenimpini (potenimpini.f90:14-66), updatepoten     */
/* (potenimp.f90:531-565), the grid sums and matrices of instates           */
/* (instates.f90:136-195 P, :448-565 D), potenimp (potenimp.f90:34-186) and */
/* the spin-orbit energy (energy.f90:168-192).  The eigenvectors are left   */
/* to the caller (the tests diagonalise the matrix with LAPACK).            */
/* ------------------------------------------------------------------------ */
static void state_prepare(oracle_t* o, const effpot_state_imp* s) {
  const int L = o->npot + 2;
  if (!o->stab) {
    o->stab = dalloc(3 * (size_t)L);
    for (int k = 0; k < 3; k++) o->sf[k] = dalloc(o->N);
  }
  double* t0 = o->stab; double* t1 = o->stab + L; double* t2 = o->stab + 2 * L;
  for (int i = 1; i <= o->npot; i++) {
    double r = (double)(i - 1) * o->delinter;
    if (s->state == EFFPOT_STATE_P) {
      double aux = oracle_plus_pot(&s->v_pi, r);
      t0[i] = aux;                                   /* potpi  */
      t1[i] = oracle_plus_pot(&s->v_sigma, r) - aux; /* potdel */
    } else {
      double aux = oracle_plus_pot(&s->v_delta, r);
      t0[i] = aux;                                   /* potdel    */
      t1[i] = oracle_plus_pot(&s->v_pi, r) - aux;    /* potpidel  */
      t2[i] = oracle_plus_pot(&s->v_sigma, r) - aux; /* potsigdel */
    }
  }
  double* ts[3] = {t0, t1, t2};
  for (int k = 0; k < 3; k++) { ts[k][0] = ts[k][1]; ts[k][o->npot + 1] = ts[k][o->npot]; }
  const double rmax = (double)(o->npot - 1) * o->delinter, zeror = 1e-3;
  for (int iz = 0; iz < o->nz; iz++) {
    double zt = (o->z[iz] - s->rimp[2]) * (o->z[iz] - s->rimp[2]);
    for (int iy = 0; iy < o->ny; iy++) {
      double yt = (o->y[iy] - s->rimp[1]) * (o->y[iy] - s->rimp[1]) + zt;
      for (int ix = 0; ix < o->nx; ix++) {
        size_t i = IDX(o, ix, iy, iz);
        double r2 = (o->x[ix] - s->rimp[0]) * (o->x[ix] - s->rimp[0]) + yt, r = sqrt(r2);
        if (s->state == EFFPOT_STATE_P) {
          o->sf[0][i] = oracle_update(r, o->npot, rmax, t0);                                      /* Vpi   */
          o->sf[1][i] = r > zeror ? oracle_update(r, o->npot, rmax, t1) / r2 : s->v_sigma.umax;   /* Delta */
        } else {
          double r4 = r2 * r2;
          o->sf[0][i] = oracle_update(r, o->npot, rmax, t0);                                      /* Delta   */
          o->sf[1][i] = r > zeror ? oracle_update(r, o->npot, rmax, t1) / r4 : s->v_pi.umax;      /* Pi_Del  */
          o->sf[2][i] = r > zeror ? oracle_update(r, o->npot, rmax, t2) / r4 : s->v_sigma.umax;   /* Sig_Del */
        }
      }
    }
  }
}

/* armonicd and the MP entries of instates/potenimp for one point. */
static void state_dforms(double xx, double yy, double zz, double* arm, double mp[5][5]) {
  const double sqr3 = sqrt(3.0), sqr3o2 = sqrt(3.0) / 2.0, insqr3 = 1.0 / sqrt(3.0), in3 = 1.0 / 3.0;
  double d0 = xx * xx + yy * yy + zz * zz, d1 = sqr3 * xx * yy, d2 = sqr3 * yy * zz;
  double d3 = (-xx * xx - yy * yy + 2 * zz * zz) / 2.0, d4 = sqr3 * xx * zz, d5 = sqr3o2 * (xx * xx - yy * yy);
  double sd1 = d1 * d1, sd2 = d2 * d2, sd4 = d4 * d4, sd5 = d5 * d5;
  double sumd3d5 = d3 + insqr3 * d5, difd3d5 = d3 - insqr3 * d5;
  arm[0] = d1; arm[1] = d2; arm[2] = d3; arm[3] = d4; arm[4] = d5;
  mp[0][0] = in3 * (sd2 + sd4 + 4.0 * sd5);
  mp[1][1] = in3 * (sd1 + sd4) + sumd3d5 * sumd3d5;
  mp[2][2] = sd4 + sd2;
  mp[3][3] = in3 * (sd1 + sd2) + difd3d5 * difd3d5;
  mp[4][4] = in3 * (sd2 + sd4 + 4.0 * sd1);
  mp[0][1] = -4.0 * in3 * d1 * d2 + insqr3 * d4 * d0;
  mp[0][2] = -2.0 * insqr3 * d2 * d4;
  mp[0][3] = -4.0 * in3 * d1 * d4 + insqr3 * d2 * d0;
  mp[0][4] = -4.0 * in3 * d1 * d5;
  mp[1][2] = insqr3 * d1 * d4 - d2 * sumd3d5;
  mp[1][3] = -4.0 * in3 * d2 * d4 + insqr3 * d1 * d0;
  mp[1][4] = -d1 * d4 - insqr3 * d2 * sumd3d5;
  mp[2][3] = insqr3 * d1 * d2 - d4 * difd3d5;
  mp[2][4] = insqr3 * (sd2 - sd4);
  mp[3][4] = d1 * d2 + insqr3 * d4 * difd3d5;
  for (int i = 0; i < 5; i++)
    for (int j = 0; j < i; j++) mp[i][j] = mp[j][i];
}

/* instates' grid sums (times dxyz) in the library's moment order. */
int oracle_state_moments(oracle_t* o, const effpot_state_imp* s, double* mom) {
  if (s->state != EFFPOT_STATE_P && s->state != EFFPOT_STATE_D) return EFFPOT_ERR_ARG;
  state_prepare(o, s);
  for (int k = 0; k < 31; k++) mom[k] = 0.0;
  for (int iz = 0; iz < o->nz; iz++)
    for (int iy = 0; iy < o->ny; iy++)
      for (int ix = 0; ix < o->nx; ix++) {
        size_t i = IDX(o, ix, iy, iz);
        double xx = o->x[ix] - s->rimp[0], yy = o->y[iy] - s->rimp[1], zz = o->z[iz] - s->rimp[2];
        double den = o->den[i];
        if (s->state == EFFPOT_STATE_P) { /* instates.f90:145-153 */
          double vpi = o->sf[0][i], dv = o->sf[1][i];
          mom[0] += den;
          mom[1] += den * (vpi + dv * xx * xx);
          mom[2] += den * (vpi + dv * yy * yy);
          mom[3] += den * (vpi + dv * zz * zz);
          mom[4] += den * dv * xx * yy;
          mom[5] += den * dv * xx * zz;
          mom[6] += den * dv * yy * zz;
        } else { /* instates.f90:460-517 */
          double arm[5], mp[5][5];
          state_dforms(xx, yy, zz, arm, mp);
          double dvdel = o->sf[0][i] * den, dv1 = o->sf[1][i] * den, dv2 = o->sf[2][i] * den;
          int k = 0;
          for (int p = 0; p < 5; p++)
            for (int q = p; q < 5; q++) mom[k++] += mp[p][q] * dv2;
          for (int p = 0; p < 5; p++)
            for (int q = p; q < 5; q++) mom[k++] += (arm[p] * arm[q]) * dv1;
          mom[30] += dvdel;
        }
      }
  for (int k = 0; k < 31; k++) mom[k] *= o->dxyz;
  return EFFPOT_OK;
}

/* V + (Als/2) SO as instates builds it (P :159-195, D :520-565), n x n
 * row-major, re/im interleaved. */
int oracle_state_matrix(int state, const double* mom, double als, double* H) {
  const int n = state == EFFPOT_STATE_P ? 6 : 10;
  double vr[10][10] = {{0}}, sr[10][10] = {{0}}, si[10][10] = {{0}};
  if (state == EFFPOT_STATE_P) {
    double Vx = mom[1], Vy = mom[2], Vz = mom[3], Vxy = mom[4], Vxz = mom[5], Vyz = mom[6];
    vr[0][0] = vr[1][1] = Vx; vr[2][2] = vr[3][3] = Vy; vr[4][4] = vr[5][5] = Vz;
    vr[0][2] = vr[1][3] = Vxy; vr[0][4] = vr[1][5] = Vxz; vr[2][4] = vr[3][5] = Vyz;
    for (int i = 0; i < n; i++)
      for (int j = 0; j < i; j++) vr[i][j] = vr[j][i];
    /* SO(1,3)=-i SO(2,4)=i SO(2,5)=-1 SO(1,6)=1 SO(4,5)=-i SO(3,6)=-i, Hermitian */
    si[0][2] = -1; si[1][3] = 1; sr[1][4] = -1; sr[0][5] = 1; si[3][4] = -1; si[2][5] = -1;
  } else {
    double mp[5][5], ms[5][5];
    int k = 0;
    for (int p = 0; p < 5; p++)
      for (int q = p; q < 5; q++) { mp[p][q] = mp[q][p] = mom[k++]; }
    for (int p = 0; p < 5; p++)
      for (int q = p; q < 5; q++) { ms[p][q] = ms[q][p] = mom[k++]; }
    for (int p = 0; p < 5; p++)
      for (int q = 0; q < 5; q++) {
        double v = (p == q ? mom[30] : 0.0) + ms[p][q] + mp[p][q];
        vr[2 * p][2 * q] = v; vr[2 * p + 1][2 * q + 1] = v;
      }
    const double s3 = sqrt(3.0);
    sr[0][3] = 1; si[0][7] = -1; si[0][8] = 2; sr[1][2] = -1; si[1][6] = -1; si[1][9] = -2;
    si[2][5] = -s3; si[2][6] = 1; si[2][9] = -1; si[3][4] = -s3; si[3][7] = -1; si[3][8] = -1;
    sr[4][7] = -s3; sr[5][6] = s3; sr[6][9] = -1; sr[7][8] = 1;
  }
  for (int i = 0; i < n; i++)
    for (int j = 0; j < i; j++) { sr[i][j] = sr[j][i]; si[i][j] = -si[j][i]; }
  for (int i = 0; i < n; i++)
    for (int j = 0; j < n; j++) {
      H[2 * (i * n + j)] = vr[i][j] + 0.5 * als * sr[i][j];
      H[2 * (i * n + j) + 1] = 0.5 * als * si[i][j];
    }
  return n;
}

/* eso of invar: energy.f90:170-172 (P, closed form) and :177-190 (D). */
double oracle_state_eso(int state, const double* iv, double als) {
  if (state == EFFPOT_STATE_P) {
    /* c = conj(invar): 2 Re(c1 v6 - c2 v5) + 2 Im(c1 v3 + c4 v2 + c3 v6 + c4 v5) */
#define CRE(a, b) (iv[2 * (a)] * iv[2 * (b)] + iv[2 * (a) + 1] * iv[2 * (b) + 1])
#define CIM(a, b) (iv[2 * (a)] * iv[2 * (b) + 1] - iv[2 * (a) + 1] * iv[2 * (b)])
    double t = 2.0 * (CRE(0, 5) - CRE(1, 4)) + 2.0 * (CIM(0, 2) + CIM(3, 1) + CIM(2, 5) + CIM(3, 4));
#undef CRE
#undef CIM
    return 0.5 * als * t;
  }
  double zero[31] = {0}, H[200];
  oracle_state_matrix(EFFPOT_STATE_D, zero, 2.0, H); /* H = SOD */
  double re = 0.0;
  for (int i = 0; i < 10; i++)
    for (int j = 0; j < 10; j++) {
      double hr = H[2 * (i * 10 + j)], hi = H[2 * (i * 10 + j) + 1];
      double vr = iv[2 * j], vi = iv[2 * j + 1], cr = iv[2 * i], ci = -iv[2 * i + 1];
      re += cr * (hr * vr - hi * vi) - ci * (hr * vi + hi * vr);
    }
  return 0.5 * re * als;
}

/* potenimp for a given invar (P :34-45, D :54-158), cap (:184-186); keeps
 * eso for energy(). */
int oracle_set_impurity_state(oracle_t* o, const effpot_state_imp* s) {
  if (s->state != EFFPOT_STATE_P && s->state != EFFPOT_STATE_D) return EFFPOT_ERR_ARG;
  state_prepare(o, s);
  const double* iv = s->invar;
  for (int iz = 0; iz < o->nz; iz++)
    for (int iy = 0; iy < o->ny; iy++)
      for (int ix = 0; ix < o->nx; ix++) {
        size_t i = IDX(o, ix, iy, iz);
        double xx = o->x[ix] - s->rimp[0], yy = o->y[iy] - s->rimp[1], zz = o->z[iz] - s->rimp[2];
        double u;
        if (s->state == EFFPOT_STATE_P) {
          double odr = iv[0] * xx + iv[4] * yy + iv[8] * zz, odi = iv[1] * xx + iv[5] * yy + iv[9] * zz;
          double evr = iv[2] * xx + iv[6] * yy + iv[10] * zz, evi = iv[3] * xx + iv[7] * yy + iv[11] * zz;
          u = o->sf[0][i] + o->sf[1][i] * (odr * odr + odi * odi + evr * evr + evi * evi);
        } else {
          double arm[5], mp[5][5], m[5][5];
          state_dforms(xx, yy, zz, arm, mp);
          for (int p = 0; p < 5; p++)
            for (int q = 0; q < 5; q++)
              m[p][q] = (p == q ? o->sf[0][i] : 0.0) + arm[p] * arm[q] * o->sf[1][i] + mp[p][q] * o->sf[2][i];
          /* sum_ij MatD_ij conj(iv_i) iv_j, MatD = m (x) 1_2 */
          u = 0.0;
          for (int a = 0; a < 10; a++)
            for (int b = 0; b < 10; b++) {
              if ((a & 1) != (b & 1)) continue;
              double md = m[a / 2][b / 2];
              u += md * (iv[2 * a] * iv[2 * b] + iv[2 * a + 1] * iv[2 * b + 1]);
            }
        }
        o->uimp[i] = u < 1000.0 ? u : 1000.0;
      }
  o->eso = oracle_state_eso(s->state, iv, s->als);
  memcpy(o->imp.rimp, s->rimp, sizeof(s->rimp));
  o->have_imp = 1;
  return EFFPOT_OK;
}

void oracle_poten(oracle_t* o);
void oracle_energy(oracle_t* o, effpot_energies* e);

/* potinvar (potenimp.f90:261-492): Hinvar = -i (H(rho) invar - Ev0 invar). */
static void potinvar(oracle_t* o) {
  double mom[31], H[200];
  oracle_state_moments(o, &o->simp, mom);
  const int n = oracle_state_matrix(o->simp.state, mom, o->simp.als, H);
  const double* iv = o->simp.invar;
  for (int i = 0; i < n; i++) {
    double re = 0.0, im = 0.0;
    for (int j = 0; j < n; j++) {
      double hr = H[2 * (i * n + j)], hi = H[2 * (i * n + j) + 1];
      re += hr * iv[2 * j] - hi * iv[2 * j + 1];
      im += hr * iv[2 * j + 1] + hi * iv[2 * j];
    }
    re -= o->ev0 * iv[2 * i];
    im -= o->ev0 * iv[2 * i + 1];
    o->hiv[2 * i] = im; /* -i (re + i im) */
    o->hiv[2 * i + 1] = -re;
  }
}

/* invar evolves from now on (Lfix_lambda = .false.); ev0 = 0: Ev0 = eHeX + eso
 * of the current state (DFT4He3d.f90:444).  s is the P/D set-up of
 * oracle_set_impurity_state with the starting invar. */
int oracle_set_state_dynamics(oracle_t* o, const effpot_state_imp* s, double ev0) {
  o->simp = *s;
  oracle_set_impurity_state(o, &o->simp);
  oracle_poten(o);
  if (ev0 == 0.0) {
    effpot_energies e;
    oracle_energy(o, &e);
    ev0 = e.eHeX + e.eso;
  }
  o->ev0 = ev0;
  memset(o->hiv, 0, sizeof(o->hiv)); memset(o->ivold, 0, sizeof(o->ivold));
  memset(o->hivold, 0, sizeof(o->hivold)); memset(o->pciv, 0, sizeof(o->pciv)); memset(o->qiv, 0, sizeof(o->qiv));
  for (int k = 0; k < 3; k++) o->iiv[k] = k;
  o->ihiv[0] = 0; o->ihiv[1] = 1;
  o->dyn = 1;
  potinvar(o);
  return EFFPOT_OK;
}

void oracle_get_invar(const oracle_t* o, double* iv) { memcpy(iv, o->simp.invar, sizeof(o->simp.invar)); }

/* ------------------------------------------------------------------------ */
/* poten (poten.f90:5-131) and term_alfa (term_alfa.f90:5-109).              */
/* ------------------------------------------------------------------------ */
static void term_alfa(oracle_t* o) {
  const size_t N = o->N;
  const effpot_params* p = &o->p;
  for (size_t i = 0; i < N; i++) o->falfs[i] = 1.0 - o->denalf[i] / p->den0s;
  double* srcs[3] = {o->rw1, o->rw2, o->rw3};
  double* grads[3] = {o->dxden, o->dyden, o->dzden};
  double* ints[3] = {o->intx, o->inty, o->intz};
  for (int a = 0; a < 3; a++) {
    for (size_t i = 0; i < N; i++) srcs[a][i] = o->falfs[i] * grads[a][i];
    oracle_fft_r2c(o, srcs[a], (double*)o->wk2);
    conv_back(o, o->wk2, o->kalfs, ints[a]);
  }
  for (size_t i = 0; i < N; i++)
    o->rw1[i] = (o->dxden[i] * o->intx[i]) + (o->dyden[i] * o->inty[i]) + (o->dzden[i] * o->intz[i]);
  oracle_fft_r2c(o, o->rw1, (double*)o->wk2);
  conv_back(o, o->wk2, o->kalfs, o->ualphas);
  for (size_t i = 0; i < N; i++) {
    o->ualphas[i] = o->ualphas[i] / p->den0s;
    o->rw1[i] = o->falfs[i] * o->intx[i];
    o->rw2[i] = o->falfs[i] * o->inty[i];
    o->rw3[i] = o->falfs[i] * o->intz[i];
  }
  deriv(o, 1, 1, o->rw1, o->rw4, 1);
  deriv(o, 1, 2, o->rw2, o->rw5, 1);
  deriv(o, 1, 3, o->rw3, o->rw6, 1);
  for (size_t i = 0; i < N; i++)
    o->ualphas[i] = (o->ualphas[i] + o->rw4[i] + o->rw5[i] + o->rw6[i]) * p->h2o2m4 * p->alphas;
}

void oracle_poten(oracle_t* o) {
  const size_t N = o->N;
  const effpot_params* p = &o->p;
  oracle_fft_r2c(o, o->den, (double*)o->fden);           /* fftfw_den */
  conv_back(o, o->fden, o->wcgk, o->dencg);               /* fftbk_cg  */
  deriv(o, 1, 1, o->den, o->dxden, 1);                    /* derden    */
  deriv(o, 1, 2, o->den, o->dyden, 1);
  deriv(o, 1, 3, o->den, o->dzden, 1);
  conv_back(o, o->fden, o->fvlj4, o->delj4);              /* fftbk_lj  */
  const double a0 = p->cp4 / 2.0, a1 = p->cpp4 / 3.0;
  for (size_t i = 0; i < N; i++) o->sto1[i] = o->den[i] * o->dencg[i] * (p->cp4 + p->cpp4 * o->dencg[i]);
  oracle_fft_r2c(o, o->sto1, (double*)o->wk2);            /* fftfw_1   */
  conv_back(o, o->wk2, o->wcgk, o->sto1);                 /* fftbk_1   */
  for (size_t i = 0; i < N; i++)
    o->pot4[i] = o->delj4[i] + o->dencg[i] * o->dencg[i] * (a0 + a1 * o->dencg[i]) + o->sto1[i];
  if (p->core == EFFPOT_CORE_OTC) {
    conv_back(o, o->fden, o->kalfs, o->denalf);           /* fftbk_as  */
    term_alfa(o);
    for (size_t i = 0; i < N; i++) o->pot4[i] = o->pot4[i] + o->ualphas[i];
  } else {
    memset(o->ualphas, 0, sizeof(double) * N);
  }
  /* solid penalty (poten.f90:80-81, 122-126) */
  for (size_t i = 0; i < N; i++) {
    double t = tanh(p->beta * (o->den[i] - p->den_m));
    double pen = p->C * (1.0 + t + p->beta * o->den[i] * (1.0 - t * t));
    o->pot4[i] = o->pot4[i] + pen;
  }
}

/* ------------------------------------------------------------------------ */
/* energy (energy.f90:12-202) + r_cm (r_cm.f90:22-57) + mu                    */
/* ------------------------------------------------------------------------ */
void oracle_energy(oracle_t* o, effpot_energies* e) {
  const size_t N = o->N;
  const effpot_params* p = &o->p;
  double* g[3] = {o->lap, o->hpsi, NULL};
  double* gz = dalloc(2 * N);
  g[2] = gz;
  deriv(o, 1, 1, o->psi, g[0], 2);
  deriv(o, 1, 2, o->psi, g[1], 2);
  deriv(o, 1, 3, o->psi, g[2], 2);
  const double aux1 = 0.5 * p->cp4, aux2 = (1.0 / 3.0) * p->cpp4;
  double ekin = 0.0, elj = 0.0, ecor = 0.0, eal = 0.0, esol = 0.0, ehex = 0.0, nat = 0.0, upl = 0.0;
  const double* ri = o->imp.rimp;
  double xc = 0.0, yc = 0.0, zc = 0.0;
  for (int iz = 0; iz < o->nz; iz++)
    for (int iy = 0; iy < o->ny; iy++)
      for (int ix = 0; ix < o->nx; ix++) {
        size_t i = IDX(o, ix, iy, iz);
        double a3 = o->den[i], a4 = o->dencg[i];
        double a5 = a3 * a4 * a4 * (aux1 + aux2 * a4);
        ekin = ekin + (g[0][2 * i] * g[0][2 * i] + g[0][2 * i + 1] * g[0][2 * i + 1]) +
               (g[1][2 * i] * g[1][2 * i] + g[1][2 * i + 1] * g[1][2 * i + 1]) +
               (g[2][2 * i] * g[2][2 * i] + g[2][2 * i + 1] * g[2][2 * i + 1]);
        elj += o->delj4[i] * a3;
        ecor += a5;
        if (p->core == EFFPOT_CORE_OTC)
          eal += o->falfs[i] * (o->dxden[i] * o->intx[i] + o->dyden[i] * o->inty[i] + o->dzden[i] * o->intz[i]);
        esol += a3 * (1.0 + tanh(p->beta * (a3 - p->den_m)));
        { /* energy.f90:149-157: rXHe = rimp - r; points with r < z_exciplex_exclusion are left out */
          double rx = ri[0] - o->x[ix], ry = ri[1] - o->y[iy], rz = ri[2] - o->z[iz];
          double r = sqrt(rx * rx + ry * ry + rz * rz);
          if (r >= o->excl) {
            ehex += a3 * o->uimp[i];
            if (o->plus.kind != EFFPOT_POT_NULL) upl += a3 * oracle_plus_pot(&o->plus, r);
          }
        }
        nat += a3;
        xc += a3 * o->x[ix]; yc += a3 * o->y[iy]; zc += a3 * o->z[iz];
      }
  free(gz);
  memset(e, 0, sizeof(*e));
  e->natoms = nat * o->dxyz;
  e->ekin4 = p->h2o2m4 * ekin * o->dxyz;
  e->elj4 = 0.5 * elj * o->dxyz;
  e->ecor4 = ecor * o->dxyz;
  e->esolid = p->C * esol * o->dxyz;
  e->etot4 = e->ekin4 + e->elj4 + e->ecor4;
  if (p->core == EFFPOT_CORE_OTC) {
    e->ealphas = -p->h2o2m4 * 0.5 * p->alphas * eal * o->dxyz;
    e->etot4 += e->ealphas;
  }
  e->eHeX = ehex * o->dxyz;
  e->ekinx = o->moving ? 0.5 * o->mimp * (o->vimp[0] * o->vimp[0] + o->vimp[1] * o->vimp[1] + o->vimp[2] * o->vimp[2])
                       : 0.0; /* energy.f90:112 */
  e->eso = o->eso;
  e->uplus = upl * o->dxyz;
  e->eimpu = e->ekinx + e->eHeX + e->eso;
  e->etot = e->etot4 + e->eimpu;
  e->xcm = xc * o->dxyz / p->n4; e->ycm = yc * o->dxyz / p->n4; e->zcm = zc * o->dxyz / p->n4;
  /* mu = <psi|(-h2o2m4 lap + pot4 + uimp)|psi> / N */
  deriv(o, 2, 1, o->psi, o->lap, 2);
  deriv(o, 2, 2, o->psi, o->hpsi, 2);
  double* l3 = dalloc(2 * N);
  deriv(o, 2, 3, o->psi, l3, 2);
  double mu = 0.0;
  for (size_t i = 0; i < N; i++) {
    for (int c = 0; c < 2; c++) {
      double lp = (o->lap[2 * i + c] + o->hpsi[2 * i + c]) + l3[2 * i + c];
      double h = lp * p->h2o2m4 - o->pot4[i] * o->psi[2 * i + c] - o->uimp[i] * o->psi[2 * i + c];
      mu -= o->psi[2 * i + c] * h;
    }
  }
  e->mu = mu * o->dxyz / e->natoms;
  /* residual |(-h2o2m4 lap + pot4 + uimp) psi - mu psi| / (|mu| |psi|), summed
   * directly from the residual at every point (the device forms it from
   * |H psi|^2 - mu^2 N) */
  double r2 = 0.0;
  for (size_t i = 0; i < N; i++) {
    for (int c = 0; c < 2; c++) {
      double lp = (o->lap[2 * i + c] + o->hpsi[2 * i + c]) + l3[2 * i + c];
      double h = lp * p->h2o2m4 - o->pot4[i] * o->psi[2 * i + c] - o->uimp[i] * o->psi[2 * i + c];
      double r = -h - e->mu * o->psi[2 * i + c];
      r2 += r * r;
    }
  }
  free(l3);
  e->resid = e->mu != 0.0 ? sqrt(r2 * o->dxyz) / (fabs(e->mu) * sqrt(e->natoms)) : 0.0;
}

/* ------------------------------------------------------------------------ */
/* H psi and the integrators (steprk.f90, steppc.f90)                        */
/* ------------------------------------------------------------------------ */
/* hpsi = (lap(psi)*h2o2m4 - pot4*psi) - uimp*psi, lap = (dxx + dyy) + dzz. */
static void compute_hpsi(oracle_t* o, double* hp) {
  const size_t N = o->N;
  double* d1 = o->lap;
  double* d2 = hp;
  double* d3 = dalloc(2 * N);
  deriv(o, 2, 1, o->psi, d1, 2);
  deriv(o, 2, 2, o->psi, d2, 2);
  deriv(o, 2, 3, o->psi, d3, 2);
  const double h2 = o->p.h2o2m4;
#pragma omp parallel for schedule(static)
  for (size_t i = 0; i < N; i++)
    for (int c = 0; c < 2; c++) {
      size_t k = 2 * i + c;
      double lp = (d1[k] + d2[k]) + d3[k];
      hp[k] = (lp * h2 - o->pot4[i] * o->psi[k]) - o->uimp[i] * o->psi[k];
    }
  if (o->rt) /* timec = i (steppc.f90:53): i*H psi */
    for (size_t i = 0; i < N; i++) { double re = hp[2 * i]; hp[2 * i] = -hp[2 * i + 1]; hp[2 * i + 1] = re; }
  free(d3);
}

/* The impurity field at r from the table of oracle_set_impurity (updatepoten
 * with Delta = 0, cap 1000 K or the impurity's cap). */
static void imp_field_at(oracle_t* o, const double* rr) {
  double rmax = (double)(o->npot - 1) * o->delinter;
  for (int iz = 0; iz < o->nz; iz++) {
    double zt = (o->z[iz] - rr[2]) * (o->z[iz] - rr[2]);
    for (int iy = 0; iy < o->ny; iy++) {
      double yt = (o->y[iy] - rr[1]) * (o->y[iy] - rr[1]) + zt;
      for (int ix = 0; ix < o->nx; ix++) {
        double r = sqrt((o->x[ix] - rr[0]) * (o->x[ix] - rr[0]) + yt);
        double u = oracle_update(r, o->npot, rmax, o->vtab);
        o->uimp[IDX(o, ix, iy, iz)] = u < o->imp.cap ? u : o->imp.cap;
      }
    }
  }
}

/* forceimp (potenimp.f90:213-254 without the exciplex filter): a = F / m. */
static void force_imp(oracle_t* o) {
  double f[3] = {0.0, 0.0, 0.0};
  /* Lfilter_exciplex_force: F(3) over the 1-based planes iz >= iz0 (rimp(3)
   * < 0) or iz <= nz0, both from z_exciplex_exclusion (potenimp.f90:216-245) */
  int lo = 1, hi = o->nz;
  if (o->ffilter) {
    const double rz = o->imp.rimp[2];
    if (rz < 0.0) {
      int iz0 = (int)((rz - o->z[0] + o->excl) / o->hz + 1.5);
      lo = iz0 > 1 ? iz0 : 1;
    } else {
      int nz0 = (int)((rz - o->z[0] - o->excl) / o->hz + 1.5);
      hi = nz0 < o->nz ? nz0 : o->nz;
    }
  }
  const size_t plane = (size_t)o->nx * o->ny;
  for (size_t i = 0; i < o->N; i++) {
    const int iz = (int)(i / plane) + 1;
    f[0] += o->dxden[i] * o->uimp[i];
    f[1] += o->dyden[i] * o->uimp[i];
    if (iz >= lo && iz <= hi) f[2] += o->dzden[i] * o->uimp[i];
  }
  for (int c = 0; c < 3; c++) o->aimp[c] = (-f[c] * o->dxyz) / o->mimp;
}

void oracle_set_real_time(oracle_t* o, int on) { o->rt = on; }

/* The impurity moves with velocity v and mass mass_u (u); its acceleration
 * from the current density (DFT4He3d.f90:437-453: poten, forceimp). */
int oracle_set_impurity_motion(oracle_t* o, const double* v, double mass_u) {
  if (!o->have_imp) return EFFPOT_ERR_STATE;
  o->moving = mass_u > 0;
  if (!o->moving) return EFFPOT_OK;
  o->mimp = mass_u * 0.0207569277; /* mp_u, modules.f90:333 */
  memset(o->rimpold, 0, sizeof(o->rimpold)); memset(o->vimpold, 0, sizeof(o->vimpold));
  memset(o->aimpold, 0, sizeof(o->aimpold));
  for (int c = 0; c < 3; c++) { o->vimp[c] = v[c]; o->pcr[c] = o->pcv[c] = o->qr[c] = o->qv[c] = 0.0; }
  for (int k = 0; k < 3; k++) o->ioldr[k] = o->ioldv[k] = k;
  o->iolda[0] = 0; o->iolda[1] = 1;
  oracle_poten(o);
  force_imp(o);
  return EFFPOT_OK;
}

void oracle_get_impurity(const oracle_t* o, double* r, double* v, double* a) {
  for (int c = 0; c < 3; c++) {
    r[c] = o->imp.rimp[c];
    v[c] = o->moving ? o->vimp[c] : 0.0;
    a[c] = o->moving ? o->aimp[c] : 0.0;
  }
}

static void update_den(oracle_t* o) {
  for (size_t i = 0; i < o->N; i++) o->den[i] = o->psi[2 * i] * o->psi[2 * i] + o->psi[2 * i + 1] * o->psi[2 * i + 1];
}

static int den_ok(const oracle_t* o) {
  for (size_t i = 0; i < o->N; i++) if (!(o->den[i] > 0.0)) return 0;
  return 1;
}

/* steprk.f90:261-420 with the imaginary-time H. */

/* The mean field inside an imaginary-time step is evaluated on the density of
 * the stage or predicted state scaled to n4 atoms.  Each step multiplies psi
 * by about 1 - mu dt (the renormalisation after the step restores n4), so
 * a field of the unscaled density would be that of ~(1 - 2 mu dt) n4 atoms:
 * the iteration's fixed point would then solve H[(1 + O(dt)) rho] psi = mu psi,
 * not the eigenvalue equation, and the converged energies would depend on dt.
 * With the scaled density the fixed point is the stationary state of the
 * functional.  Real time (timec = i) keeps the reference's den = |psi|^2. */
static void poten_of_step(oracle_t* o) {
  if (!o->rt) {
    double s = 0.0;
    for (size_t i = 0; i < o->N; i++) s += o->den[i];
    const double f = o->p.n4 / (s * o->dxyz);
    for (size_t i = 0; i < o->N; i++) o->den[i] *= f;
  }
  oracle_poten(o);
}

int oracle_step_rk(oracle_t* o, double dt) {
  const double s2 = sqrt(2.0);
  const double arun[4] = {0.5, 1.0 - 1.0 / s2, 1.0 + 1.0 / s2, 1.0 / 6.0};
  const double brun[4] = {2.0, 1.0, 1.0, 2.0};
  const double crun[4] = {0.5, 1.0 - 1.0 / s2, 1.0 + 1.0 / s2, 0.5};
  const size_t N2 = 2 * o->N;
  double* hp = o->hpsi;
  for (int j = 0; j < 4; j++) {
    compute_hpsi(o, hp);
    if (j == 0) {
      /* shift the history once per step: hpsiold(2)=hpsiold(1), hpsiold(1)=Hpsi,
       * psiold(3)=psiold(2), psiold(2)=psiold(1), psiold(1)=psi */
      memcpy(o->hpsiold[1], o->hpsiold[0], sizeof(double) * N2);
      memcpy(o->hpsiold[0], hp, sizeof(double) * N2);
      memcpy(o->psiold[2], o->psiold[1], sizeof(double) * N2);
      memcpy(o->psiold[1], o->psiold[0], sizeof(double) * N2);
      memcpy(o->psiold[0], o->psi, sizeof(double) * N2);
    }
    for (size_t k = 0; k < N2; k++) {
      double a1 = hp[k];
      double a2 = arun[j] * (a1 - brun[j] * o->q[k]);
      o->q[k] = o->q[k] + 3.0 * a2 - crun[j] * a1;
      o->psi[k] = o->psi[k] + dt * a2;
    }
    update_den(o);
    if (!den_ok(o)) return EFFPOT_ERR_NONFINITE;
    if (o->moving) { /* steprk.f90:117-150 */
      double* r = o->imp.rimp;
      for (int c = 0; c < 3; c++) {
        double a2 = arun[j] * (o->vimp[c] - brun[j] * o->qr[c]);
        o->qr[c] = o->qr[c] + 3.0 * a2 - crun[j] * o->vimp[c];
        if (j == 0) { o->rimpold[2][c] = o->rimpold[1][c]; o->rimpold[1][c] = o->rimpold[0][c]; o->rimpold[0][c] = r[c]; }
        r[c] = r[c] + dt * a2;
      }
      for (int c = 0; c < 3; c++) {
        double a2 = arun[j] * (o->aimp[c] - brun[j] * o->qv[c]);
        o->qv[c] = o->qv[c] + 3.0 * a2 - crun[j] * o->aimp[c];
        if (j == 0) {
          o->aimpold[1][c] = o->aimpold[0][c]; o->aimpold[0][c] = o->aimp[c];
          o->vimpold[2][c] = o->vimpold[1][c]; o->vimpold[1][c] = o->vimpold[0][c]; o->vimpold[0][c] = o->vimp[c];
        }
        o->vimp[c] = o->vimp[c] + dt * a2;
      }
      if (j <= 2 && !o->dyn) imp_field_at(o, r);
    }
    if (o->dyn) { /* steprk.f90:152-167 */
      if (o->moving) memcpy(o->simp.rimp, o->imp.rimp, sizeof(o->simp.rimp));
      double* iv = o->simp.invar;
      for (int k = 0; k < 20; k++) {
        double a2 = arun[j] * (o->hiv[k] - brun[j] * o->qiv[k]);
        o->qiv[k] = o->qiv[k] + 3.0 * a2 - crun[j] * o->hiv[k];
        if (j == 0) {
          o->hivold[1][k] = o->hivold[0][k]; o->hivold[0][k] = o->hiv[k];
          o->ivold[2][k] = o->ivold[1][k]; o->ivold[1][k] = o->ivold[0][k]; o->ivold[0][k] = iv[k];
        }
        iv[k] = iv[k] + dt * a2;
      }
      if (j <= 2) oracle_set_impurity_state(o, &o->simp); /* potenimp(rimp, invar) */
    }
    if (j <= 2) poten_of_step(o);
    if (o->moving && j <= 2) force_imp(o);
    if (o->dyn && j <= 2) potinvar(o);
  }
  for (int i = 0; i < 3; i++) o->iiv[i] = i;
  o->ihiv[0] = 0; o->ihiv[1] = 1;
  for (int i = 0; i < 3; i++) o->ioldp[i] = i;
  for (int i = 0; i < 2; i++) o->ioldh[i] = i;
  for (int i = 0; i < 3; i++) o->ioldr[i] = o->ioldv[i] = i;
  o->iolda[0] = 0; o->iolda[1] = 1;
  return EFFPOT_OK;
}

/* steppc.f90:35-159 (helium part) with the imaginary-time H. */
int oracle_step_pc(oracle_t* o, double dt, double* err_he) {
  const double c112 = 112.0 / 121.0, c9 = 9.0 / 121.0, c4o3 = 4.0 / 3.0;
  const size_t N2 = 2 * o->N;
  double* hp = o->hpsi;
  compute_hpsi(o, hp);
  double* p3 = o->psiold[o->ioldp[2]];
  double* h1 = o->hpsiold[o->ioldh[0]];
  double* h2 = o->hpsiold[o->ioldh[1]];
  for (size_t k = 0; k < N2; k++) {
    double pred = p3[k] + c4o3 * dt * (2.0 * hp[k] - h1[k] + 2.0 * h2[k]);
    p3[k] = o->psi[k];
    o->psi[k] = pred - c112 * o->pc[k];
    o->pc[k] = pred;
    h2[k] = hp[k];
  }
  update_den(o);
  { int t = o->ioldh[1]; o->ioldh[1] = o->ioldh[0]; o->ioldh[0] = t; }
  const double c4o3m = 4.0 / 3.0;
  if (o->moving) { /* steppc.f90:68-103 */
    double* r = o->imp.rimp;
    int* ir = o->ioldr; int* iv = o->ioldv; int* ia = o->iolda;
    for (int c = 0; c < 3; c++) {
      double aux = o->rimpold[ir[2]][c] + c4o3m * dt * (2.0 * o->vimp[c] - o->vimpold[iv[0]][c] + 2.0 * o->vimpold[iv[1]][c]);
      o->rimpold[ir[2]][c] = r[c];
      r[c] = aux - c112 * o->pcr[c];
      o->pcr[c] = aux;
    }
    for (int c = 0; c < 3; c++) {
      double aux = o->vimpold[iv[2]][c] + c4o3m * dt * (2.0 * o->aimp[c] - o->aimpold[ia[0]][c] + 2.0 * o->aimpold[ia[1]][c]);
      o->vimpold[iv[2]][c] = aux - c112 * o->pcv[c];
      o->pcv[c] = aux;
    }
    for (int c = 0; c < 3; c++) o->aimpold[ia[1]][c] = o->aimp[c];
    { int t = ia[1]; ia[1] = ia[0]; ia[0] = t; }
    if (!o->dyn) imp_field_at(o, r); /* potenimp(rimp) */
  }
  if (o->dyn) { /* steppc.f90:84-102 */
    if (o->moving) memcpy(o->simp.rimp, o->imp.rimp, sizeof(o->simp.rimp));
    double* iv = o->simp.invar;
    int* ii = o->iiv; int* ih = o->ihiv;
    for (int k = 0; k < 20; k++) {
      double aux = o->ivold[ii[2]][k] + c4o3m * dt * (2.0 * o->hiv[k] - o->hivold[ih[0]][k] + 2.0 * o->hivold[ih[1]][k]);
      o->ivold[ii[2]][k] = iv[k];
      iv[k] = aux - c112 * o->pciv[k];
      o->pciv[k] = aux;
    }
    for (int k = 0; k < 20; k++) o->hivold[ih[1]][k] = o->hiv[k];
    { int t = ih[1]; ih[1] = ih[0]; ih[0] = t; }
    oracle_set_impurity_state(o, &o->simp); /* potenimp(rimp, invar) */
  }
  poten_of_step(o);
  if (o->moving) force_imp(o);
  if (o->dyn) potinvar(o);
  compute_hpsi(o, hp);
  p3 = o->psiold[o->ioldp[2]];
  double* p2 = o->psiold[o->ioldp[1]];
  h1 = o->hpsiold[o->ioldh[0]];
  h2 = o->hpsiold[o->ioldh[1]];
  double err = 0.0;
  for (size_t i = 0; i < o->N; i++) {
    for (int c = 0; c < 2; c++) {
      size_t k = 2 * i + c;
      double corr = 0.125 * (9.0 * p3[k] - p2[k] + 3.0 * dt * (hp[k] + 2.0 * h1[k] - h2[k]));
      o->pc[k] = o->pc[k] - corr;
      o->psi[k] = corr + c9 * o->pc[k];
    }
    err += c9 * sqrt(o->pc[2 * i] * o->pc[2 * i] + o->pc[2 * i + 1] * o->pc[2 * i + 1]);
  }
  update_den(o);
  { int t = o->ioldp[2]; o->ioldp[2] = o->ioldp[1]; o->ioldp[1] = o->ioldp[0]; o->ioldp[0] = t; }
  if (o->moving) { /* steppc.f90:160-203 */
    double* r = o->imp.rimp;
    int* ir = o->ioldr; int* iv = o->ioldv; int* ia = o->iolda;
    for (int c = 0; c < 3; c++) {
      double aux = 0.125 * (9.0 * o->rimpold[ir[2]][c] - o->rimpold[ir[1]][c] +
                            3.0 * dt * (o->vimpold[iv[2]][c] + 2.0 * o->vimp[c] - o->vimpold[iv[0]][c]));
      o->pcr[c] = o->pcr[c] - aux;
      r[c] = aux + c9 * o->pcr[c];
    }
    { int t = ir[2]; ir[2] = ir[1]; ir[1] = ir[0]; ir[0] = t; }
    for (int c = 0; c < 3; c++) {
      double aux = 0.125 * (9.0 * o->vimp[c] - o->vimpold[iv[1]][c] +
                            3.0 * dt * (o->aimp[c] + 2.0 * o->aimpold[ia[0]][c] - o->aimpold[ia[1]][c]));
      o->pcv[c] = o->pcv[c] - aux;
      o->vimpold[iv[2]][c] = o->vimp[c];
      o->vimp[c] = aux + c9 * o->pcv[c];
    }
    { int t = iv[2]; iv[2] = iv[1]; iv[1] = iv[0]; iv[0] = t; }
  }
  if (o->dyn) { /* steppc.f90:205-226; a P state keeps components 7-10 at zero */
    double* iv = o->simp.invar;
    int* ii = o->iiv; int* ih = o->ihiv;
    int nn = o->simp.state == EFFPOT_STATE_P ? 12 : 20;
    for (int k = 0; k < 20; k++) {
      double aux = 0.125 * (9.0 * o->ivold[ii[2]][k] - o->ivold[ii[1]][k] +
                            3.0 * dt * (o->hiv[k] + 2.0 * o->hivold[ih[0]][k] - o->hivold[ih[1]][k]));
      if (k < nn) { o->pciv[k] = o->pciv[k] - aux; iv[k] = aux + c9 * o->pciv[k]; }
      else iv[k] = 0.0;
    }
    { int t = ii[2]; ii[2] = ii[1]; ii[1] = ii[0]; ii[0] = t; }
    o->eso = oracle_state_eso(o->simp.state, iv, o->simp.als);
  }
  if (err_he) *err_he = err / (double)o->N;
  if (!den_ok(o)) return EFFPOT_ERR_NONFINITE;
  return EFFPOT_OK;
}

/* The impurity's error estimates of the last steppc (steppc.f90:179,193,
 * 213-220): errimp = sum |c9 pcr| * 0.3333333333, errvimp = sum |c9 pcv| *
 * 0.3333333333, erriv = sum |c9 pciv| * 0.1666666666666 (P, components 1-6)
 * or * 0.1 (D); 0 for a fixed impurity or a frozen invar (Lfix_lambda). */
void oracle_step_errors(const oracle_t* o, double* out3) {
  const double c9 = 9.0 / 121.0;
  out3[0] = out3[1] = out3[2] = 0.0;
  if (o->moving) {
    for (int c = 0; c < 3; c++) {
      out3[0] += fabs(c9 * o->pcr[c]);
      out3[1] += fabs(c9 * o->pcv[c]);
    }
    out3[0] *= 0.3333333333;
    out3[1] *= 0.3333333333;
  }
  if (o->dyn) {
    const int n = o->simp.state == EFFPOT_STATE_P ? 6 : 10;
    for (int k = 0; k < n; k++) out3[2] += c9 * sqrt(o->pciv[2 * k] * o->pciv[2 * k] + o->pciv[2 * k + 1] * o->pciv[2 * k + 1]);
    out3[2] *= o->simp.state == EFFPOT_STATE_P ? 0.1666666666666 : 0.1;
  }
}

/* psi *= sqrt(n4 / (sum |psi|^2 dxyz)); den = |psi|^2. */
void oracle_renormalize(oracle_t* o) {
  double s = 0.0;
  for (size_t i = 0; i < o->N; i++) s += o->psi[2 * i] * o->psi[2 * i] + o->psi[2 * i + 1] * o->psi[2 * i + 1];
  double f = sqrt(o->p.n4 / (s * o->dxyz));
  for (size_t k = 0; k < 2 * o->N; k++) o->psi[k] *= f;
  update_den(o);
}

/* ------------------------------------------------------------------------ */
/* Initial state                                                              */
/* ------------------------------------------------------------------------ */
/* Counter-based uniform in [0,1): splitmix64 finaliser of seed and index. */
static uint64_t mix64(uint64_t z) {
  z += 0x9E3779B97F4A7C15ULL;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}
double oracle_uniform(uint64_t seed, uint64_t i) {
  return (double)(mix64(seed ^ mix64(i)) >> 11) * 0x1.0p-53;
}

void oracle_init_fermi(oracle_t* o, double rho0, double R, double afermi, double noise,
                       uint64_t seed, const double* centre) {
  double sum = 0.0;
  for (int iz = 0; iz < o->nz; iz++)
    for (int iy = 0; iy < o->ny; iy++)
      for (int ix = 0; ix < o->nx; ix++) {
        size_t i = IDX(o, ix, iy, iz);
        double dx = o->x[ix] - centre[0], dy = o->y[iy] - centre[1], dz = o->z[iz] - centre[2];
        double r = sqrt(dx * dx + dy * dy + dz * dz);
        double rho = rho0 / (1.0 + exp((r - R) / afermi)) + noise * oracle_uniform(seed, i);
        o->den[i] = rho;
        sum += rho;
      }
  double f = sqrt(o->p.n4 / (sum * o->dxyz));
  for (size_t i = 0; i < o->N; i++) {
    o->psi[2 * i] = sqrt(o->den[i]) * f;
    o->psi[2 * i + 1] = 0.0;
  }
  update_den(o);
}

/* ------------------------------------------------------------------------ */
/* Accessors                                                                  */
/* ------------------------------------------------------------------------ */
void oracle_set_psi(oracle_t* o, const double* psi) {
  memcpy(o->psi, psi, sizeof(double) * 2 * o->N);
  update_den(o);
}
void oracle_get_psi(const oracle_t* o, double* psi) { memcpy(psi, o->psi, sizeof(double) * 2 * o->N); }
void oracle_get_den(const oracle_t* o, double* d) { memcpy(d, o->den, sizeof(double) * o->N); }
void oracle_set_den(oracle_t* o, const double* d) { memcpy(o->den, d, sizeof(double) * o->N); }
int oracle_get_field(const oracle_t* o, int which, double* out) {
  const double* src[] = {o->pot4, o->dencg, o->delj4, o->denalf, o->ualphas, o->uimp,
                         o->dxden, o->dyden, o->dzden, o->sto1, o->intx, o->inty, o->intz};
  if (which < 0 || which > 12) return EFFPOT_ERR_ARG;
  memcpy(out, src[which], sizeof(double) * o->N);
  return EFFPOT_OK;
}
void oracle_get_kernels(const oracle_t* o, double* wcgk, double* fvlj4, double* kalfs) {
  memcpy(wcgk, o->wcgk, sizeof(double) * o->M);
  memcpy(fvlj4, o->fvlj4, sizeof(double) * o->M);
  memcpy(kalfs, o->kalfs, sizeof(double) * o->M);
}
void oracle_grid(const oracle_t* o, double* out3) { out3[0] = o->hx; out3[1] = o->hy; out3[2] = o->hz; }

/* ------------------------------------------------------------------------ */
/* Relaxation loop (new: the reference has the pieces but no driver).       */
/* ------------------------------------------------------------------------ */
int oracle_solve(oracle_t* o, const effpot_solve_opts* opt, effpot_result* res) {
  if (!o->have_imp) return EFFPOT_ERR_STATE;
  if (!(opt->dt > 0.0)) return EFFPOT_ERR_DT;
  memset(res, 0, sizeof(*res));
  update_den(o);
  oracle_poten(o);
  int steps = 0, rc;
  for (int s = 0; s < opt->n_rk; s++) {
    if ((rc = oracle_step_rk(o, opt->dt)) != EFFPOT_OK) return rc;
    oracle_renormalize(o);
    steps++;
  }
  effpot_energies e;
  oracle_poten(o);
  oracle_energy(o, &e);
  double eprev = e.etot;
  int since = 0;
  double err = 0.0;
  while (steps < opt->max_steps) {
    if ((rc = oracle_step_pc(o, opt->dt, &err)) != EFFPOT_OK) return rc;
    oracle_renormalize(o);
    steps++; since++;
    if (since == opt->pener || steps == opt->max_steps) {
      oracle_poten(o);
      oracle_energy(o, &e);
      double rel = fabs(e.etot - eprev) / fabs(e.etot) / (double)since;
      res->rel_de = rel;
      eprev = e.etot;
      since = 0;
      if (rel < opt->tol) { res->converged = 1; break; }
    }
  }
  if (since) { oracle_poten(o); oracle_energy(o, &e); }
  res->e = e;
  res->steps = steps;
  res->err_he = err;
  return EFFPOT_OK;
}
