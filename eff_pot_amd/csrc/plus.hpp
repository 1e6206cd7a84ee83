// X+-He potentials of the uplus energy term (energy.f90:146-163,
// Select_Pot(selec_plus, r, r_cutoff_plus, umax_plus), V_impur.f90:87).
//
// The reference's Ba+ - He fits (V_Ba_plus_gs/2D0/2D1/2D2,
// V_impur.f90:1218-1396) share one form: a plateau, a parabola and an
// inverse-power series whose 13 coefficients are given to 33 digits and summed
// in REAL*16.  Near r0 = 2 A the terms reach 1e10-1e11 K and cancel to
// 1e3-1e4 K, so a plain FP64 sum would lose ~1e-5 K per point.  Here the
// coefficients are stored as double-double pairs (pg + pg_lo) and the series
// is summed with double-double Horner steps (Dekker's exact product via FMA
// and Knuth's two-sum), which keeps the result within a few ulps of the
// REAL*16 sum.  Every function here turns FP contraction off: hipcc contracts
// by default, and a product fused into a later sum (p = a*b; s = p + e) would
// silently break the exact-product and two-sum identities.
#pragma once
#include <cmath>
#include <cstdlib>
#include <cstring>

#include "../../include/effpot.h"

namespace ep {

struct DD {
  double h, l;
};
__host__ __device__ inline DD dd_fast_two_sum(double a, double b) {
#pragma clang fp contract(off)
  const double s = a + b;
  return {s, b - (s - a)};
}
__host__ __device__ inline DD dd_add(DD a, DD b) {
#pragma clang fp contract(off)
  const double s = a.h + b.h, v = s - a.h;
  double e = (a.h - (s - v)) + (b.h - v);
  e += a.l + b.l;
  return dd_fast_two_sum(s, e);
}
__host__ __device__ inline DD dd_mul(DD a, DD b) {
#pragma clang fp contract(off)
  const double p = a.h * b.h;
  double e = fma(a.h, b.h, -p);
  e += a.h * b.l + a.l * b.h;
  return dd_fast_two_sum(p, e);
}

// The Koutselos alkali-ion - He form (V_Koutselos_*_plus_gs,
// V_impur.f90:1704-1883), in atomic units: x = r / a_bohr,
//   V = [v0 (A e^{-a x/rho} - B e^{-b x/rho})
//        + (-CC4/x^4 - CC6/x^6 - CC8/x^8) h(x)] * Hartree * eV_K,
// h = 1 for x >= 1.28 R_m, else exp(-(1.28 R_m / x - 1)^2) (0 at x = 0).
// Parameters in effpot_plus.pg (EFFPOT_KOUT_*), operations in the reference's
// order, FP64 as the reference evaluates it.
__host__ __device__ inline double koutselos_pot(const effpot_plus& p, double r) {
#pragma clang fp contract(off)
  const double* c = p.pg;
  const double x = r / c[EFFPOT_KOUT_ABOHR];
  const double rs = x / c[EFFPOT_KOUT_RHO];
  const double vex = c[EFFPOT_KOUT_V0] * (c[EFFPOT_KOUT_A] * exp(-c[EFFPOT_KOUT_ALPHA] * rs) -
                                          c[EFFPOT_KOUT_B] * exp(-c[EFFPOT_KOUT_BETA] * rs));
  const double aux = c[EFFPOT_KOUT_RM] * 1.28;
  double h = 1.0;
  if (!(x >= aux)) {
    h = 0.0;
    if (x != 0.0) {
      const double t = aux / x - 1.0;
      h = exp(-(t * t));
    }
  }
  double sto = vex;
  if (x != 0.0) {
    const double x2 = x * x, x4 = x2 * x2, x6 = x4 * x2, x8 = x6 * x2;
    sto = sto + (-c[EFFPOT_KOUT_CC4] / x4 - c[EFFPOT_KOUT_CC6] / x6 - c[EFFPOT_KOUT_CC8] / x8) * h;
  }
  return sto * c[EFFPOT_KOUT_HARTREE] * c[EFFPOT_KOUT_EVK];
}

// Select_Pot for an effpot_plus (host and device, FP64 in, FP64 out).
__host__ __device__ inline double plus_pot(const effpot_plus& p, double r) {
#pragma clang fp contract(off)
  if (p.kind == EFFPOT_POT_KOUTSELOS) return r < p.r_cutoff ? p.umax : koutselos_pot(p, r);
  if (p.kind != EFFPOT_POT_SERIES) return 0.0;
  if (r < p.r_cutoff) return p.umax;
  if (r <= p.rcutoff) return p.fcutoff;
  if (r <= p.r0) return fabs(p.aa * (r * r) + p.bb * r + p.cc);
  DD q{1.0 / r, 0.0};
  q.l = fma(-r, q.h, 1.0) / r;  // 1/r to double-double
  const int n = p.nterms;
  DD s{p.pg[n - 1], p.pg_lo[n - 1]};
  for (int k = n - 2; k >= 0; k--) s = dd_add(dd_mul(s, q), DD{p.pg[k], p.pg_lo[k]});
  for (int k = 0; k <= p.k0; k++) s = dd_mul(s, q);  // * r^-(k0+1)
  return s.h + s.l;
}

// One slot's X+ potential and the impurity position it is centred on, as
// the energy kernel reads it (on = 0: uplus is 0 and nothing is evaluated).
struct PlusSlot {
  effpot_plus p;
  double rimp[3];
  int32_t on;
  int32_t force_filter;  // Lfilter_exciplex_force: F_z only beyond excl in z
  double excl;  // z_exciplex_exclusion: points nearer than this to rimp stay out of eHeX and uplus
};

namespace plus_fits {
// rcutoff, fcutoff, r0, aa, bb, cc (REAL*8 in the reference), k0 and the
// REAL*16 coefficients as written in V_impur.f90's Block Data
// Inicio_LJ_V_Ba_plus_* (gs :1353-1371, 2D0 :1218-1236, 2D1 :1265-1283,
// 2D2 :1308-1326).
struct Fit {
  const char* name;
  double rcutoff, fcutoff, r0, aa, bb, cc;
  int k0;
  const char* pg[13];
};
static const Fit kFits[] = {
    {"Ba_plus_gs", 1.0000000000000000e+00, 2.7557869999999999e+04, 2.0000000000000000e+00, -5.1918208117033464e+03, 0.0000000000000000e+00, 3.2749691986506852e+04, 3,
     {"-1233025.18431181221429176595319949",
      "84965045.0899243014427808833458580",
      "-2329830901.26373970562255233678155",
      "34487035482.8661381911727426332743",
      "-312315418217.722824241631231266832",
      "1837177830634.86715823984693486481",
      "-7222749564186.92692954840041234856",
      "19113448483679.6067074196915946358",
      "-33537885793985.8415962556220185604",
      "37188922413954.1348677597526953193",
      "-23134511849820.2719590286530591524",
      "5477820267454.80225927552742559374",
      "657017642818.351074515977126633909"}},
    {"Ba_plus_2D0", 2.0000000000000000e+00, 2.3097950000000001e+04, 2.0000000000000000e+00, 2.4148340053884458e+04, -1.4167500343195186e+05, 2.0985459566015401e+05, 3,
     {"-1087334.00470567619271105807407299",
      "84828727.5548886783972206269121162",
      "-2706837566.34944393727530187448557",
      "47482174965.3049172156448636792347",
      "-519244658310.373320149026960630547",
      "3776794752968.39098091679878465104",
      "-18931184609496.2743530917730707405",
      "66467512849641.5867602022345173376",
      "-163349259737503.439380947298152502",
      "275471652871895.368050105366543138",
      "-303799513108731.163572027673779711",
      "197342887711895.715873097734435815",
      "-57247315878299.6181404889603871296"}},
    {"Ba_plus_2D1", 2.0000000000000000e+00, 8.9667530000000006e+03, 2.0000000000000000e+00, 1.6993055762720283e+04, -9.4455519278364838e+04, 1.2990556820533771e+05, 3,
     {"-17806.3555611251407400764850972719",
      "1823098.97448459689302902398930546",
      "-85830730.9944556694819916079148077",
      "1725318458.71278993648593521565983",
      "-19436563702.5972378924691700268134",
      "135815448941.441168085385272965158",
      "-614470726947.401636604960682655298",
      "1813794100230.25415271043202962245",
      "-3408845825677.80602502264479414645",
      "3767306350864.97856677416953497814",
      "-1870543050026.55759258430089080476",
      "-263830499030.803962487458419711050",
      "490468253633.822352576865392543355"}},
    {"Ba_plus_2D2", 2.0000000000000000e+00, 1.1687330000000000e+04, 2.0000000000000000e+00, 1.9539148260089081e+04, -1.0968122164089163e+05, 1.5289317408495379e+05, 3,
     {"3438.08194066326485411722455335026",
      "0.00000000000000000000000000000000",
      "-22271935.4558369081774443426658153",
      "509815799.148361882435180461889140",
      "-5103275255.84719311793286237036308",
      "24347543963.7285486850837178167031",
      "-22098022739.0755793944143793174099",
      "-371505231363.903388585217865272476",
      "2188591221460.46490291662922516704",
      "-6008459393013.55312031551791210851",
      "9242835510174.84636285804976040549",
      "-7681295358440.52888750989831902749",
      "2696390604822.62910359945307774040"}},
};
// The Koutselos et al. (J. Chem. Phys. 93, 7125 (1990)) He - alkali-ion
// parameters v0, rho, R_m, C4, C6, C8 (atomic units) of V_Koutselos_*_plus_gs
// (Cs :1704-1749, Na :1750-1795, Rb :1796-1839, K :1840-1883), selected as
// '<Ion>_plus_Koutselos' (V_impur.f90:232-292).
struct Kout {
  const char* name;
  double v0, rho, rm, c4, c6, c8;
};
static const Kout kKout[] = {
    {"Cs_plus_Koutselos", 0.5092, 0.9241, 6.38, 0.0, 13.69, 236.4},
    {"Na_plus_Koutselos", 0.2063, 0.7640, 4.55, 0.0, 1.424, 13.17},
    {"Rb_plus_Koutselos", 0.3474, 0.8868, 5.85, 0.0, 8.8, 126.2},
    {"K_plus_Koutselos", 0.2690, 0.8589, 5.49, 0.0, 5.83, 72.05},
};
}  // namespace plus_fits

// A decimal string ("-123.456...", no exponent) to a double-double: the
// digits are accumulated as an integer (exact below 2^106, i.e. 31 digits;
// the fits' 33-digit values round there at 1e-32 relative) and divided by the
// power of ten of the fraction digits (exact in double-double up to 10^45).
inline DD dd_sub_mul(DD r, DD b, double q) {  // r - b*q
#pragma clang fp contract(off)
  const DD t = dd_mul(b, DD{q, 0.0});
  return dd_add(r, DD{-t.h, -t.l});
}
inline DD dd_div(DD a, DD b) {  // three quotient digits (long division)
#pragma clang fp contract(off)
  const double q1 = a.h / b.h;
  const DD r1 = dd_sub_mul(a, b, q1);
  const double q2 = r1.h / b.h;
  const DD r2 = dd_sub_mul(r1, b, q2);
  return dd_add(dd_fast_two_sum(q1, q2), DD{r2.h / b.h, 0.0});
}
inline DD dd_from_decimal(const char* s) {
#pragma clang fp contract(off)
  bool neg = false;
  if (*s == '-' || *s == '+') neg = *s++ == '-';
  DD acc{0.0, 0.0}, p10{1.0, 0.0};
  bool frac = false;
  for (; *s; s++) {
    if (*s == '.') { frac = true; continue; }
    acc = dd_add(dd_mul(acc, DD{10.0, 0.0}), DD{(double)(*s - '0'), 0.0});
    if (frac) p10 = dd_mul(p10, DD{10.0, 0.0});
  }
  DD v = dd_div(acc, p10);
  return neg ? DD{-v.h, -v.l} : v;
}

// effpot_plus for a Select_Pot name; returns false for an unknown name.
inline bool plus_preset(const char* selec, double r_cutoff, double umax, effpot_plus* out) {
  size_t len = std::strlen(selec);
  while (len > 0 && selec[len - 1] == ' ') len--;  // Fortran Trim
  std::memset(out, 0, sizeof(*out));
  out->r_cutoff = r_cutoff;
  out->umax = umax;
  if (len == 4 && std::strncmp(selec, "null", 4) == 0) {
    out->kind = EFFPOT_POT_NULL;
    return true;
  }
  for (const plus_fits::Kout& f : plus_fits::kKout) {
    if (std::strlen(f.name) != len || std::strncmp(selec, f.name, len) != 0) continue;
    out->kind = EFFPOT_POT_KOUTSELOS;
    double* c = out->pg;
    c[EFFPOT_KOUT_V0] = f.v0; c[EFFPOT_KOUT_RHO] = f.rho; c[EFFPOT_KOUT_RM] = f.rm;
    // CC_n = C_n + alpha_n / 2 with the He polarisabilities (V_impur.f90:1824)
    c[EFFPOT_KOUT_CC4] = f.c4 + 1.3831 / 2.0;
    c[EFFPOT_KOUT_CC6] = f.c6 + 2.4434 / 2.0;
    c[EFFPOT_KOUT_CC8] = f.c8 + 10.614 / 2.0;
    c[EFFPOT_KOUT_A] = 146.98; c[EFFPOT_KOUT_ALPHA] = 1.5024;
    c[EFFPOT_KOUT_B] = 70.198; c[EFFPOT_KOUT_BETA] = 1.4041;
    c[EFFPOT_KOUT_ABOHR] = 0.529177; c[EFFPOT_KOUT_HARTREE] = 27.2114; c[EFFPOT_KOUT_EVK] = 11604.448;
    return true;
  }
  for (const plus_fits::Fit& f : plus_fits::kFits) {
    if (std::strlen(f.name) != len || std::strncmp(selec, f.name, len) != 0) continue;
    out->kind = EFFPOT_POT_SERIES;
    out->nterms = 13;
    out->k0 = f.k0;
    out->rcutoff = f.rcutoff; out->fcutoff = f.fcutoff; out->r0 = f.r0;
    out->aa = f.aa; out->bb = f.bb; out->cc = f.cc;
    for (int k = 0; k < 13; k++) {
      const DD v = dd_from_decimal(f.pg[k]);
      out->pg[k] = v.h;
      out->pg_lo[k] = v.l;
    }
    return true;
  }
  return false;
}

}  // namespace ep
