// Pointwise kernels: one-time k-space kernels, impurity field, pot4 assembly,
// renormalisation, deterministic reductions, precision conversion.
#pragma once
#include "common.hpp"

namespace ep {

// ---- sine integral (FP64): series for x <= 4, else pi/2 + Im E1(ix) by the
// continued fraction of E1 evaluated with the modified Lentz method -------------
EP_DEV double si_dev(double x) {
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
    double br = 1.0, bi = ax, cr = 1e300, ci = 0.0;
    double dd = br * br + bi * bi;
    double dr = br / dd, di = -bi / dd, hr = dr, hi = di;
    for (int i = 1; i < 1000; i++) {
      double an = -(double)i * (double)i;
      br += 2.0;
      double tr = an * dr + br, ti = an * di + bi;
      dd = tr * tr + ti * ti;
      dr = tr / dd; di = -ti / dd;
      double cc = cr * cr + ci * ci;
      double ir = cr / cc, ii = -ci / cc;
      cr = br + an * ir; ci = bi + an * ii;
      double delr = cr * dr - ci * di, deli = cr * di + ci * dr;
      double nr = hr * delr - hi * deli, ni = hr * deli + hi * delr;
      hr = nr; hi = ni;
      if (fabs(delr - 1.0) + fabs(deli) < 1e-16) break;
    }
    double c = cos(ax), s = sin(ax);
    res = 1.5707963267948966 + (c * hi - s * hr);
  }
  return x < 0 ? -res : res;
}

// OT-screened LJ transform (fforma.f90:40-60 with sinx11 :73-102, sinx5
// :108-137, sinint :143-196, bforce :256-280 for core OT).
EP_DEV double fvlj_dev(double p, double eps, double sigma, double hc) {
  const double pi = 3.14159265358979323846;
  double s6 = pow(sigma, 6.0);
  if (fabs(p) < 1e-10) {
    double pi4 = 4.0 * pi, eps4 = 4.0 * eps;
    return pi4 * eps4 * (s6 * s6) / (9.0 * pow(hc, 9.0)) - pi4 * eps4 * s6 / (3.0 * hc * hc * hc);
  }
  double a = 2.0 * pi * p;
  double y = a * hc, y2 = y * y;
  double si = si_dev(y);
  double sint = a > 0.0 ? 0.5 * pi - si : 0.5 * pi + si;
  double d1 = 1.0, d2 = y2, d3 = y2 * y2, d4 = d3 * y2, d5 = d4 * y2;
  double bd = 40320.0 * d1 - 720.0 * d2 + 24.0 * d3 - 2.0 * d4 + 1.0 * d5;
  double cd = 362880.0 * d1 - 5040.0 * d2 + 120.0 * d3 - 6.0 * d4 + 1.0 * d5;
  double cy = cos(y), sy = sin(y);
  double t1 = (a * bd * cy + cd * sy / hc) / (3628800.0 * pow(hc, 9.0)) - pow(a, 10.0) * sint / 3628800.0;
  double bd5 = -2.0 + y2, cd5 = -6.0 + y2;
  double t2 = -(a * bd5 * cy + cd5 * sy / hc) / (24.0 * hc * hc * hc) + pow(a, 4.0) * sint / 24.0;
  return (8.0 * eps * s6 * (s6 * t1 - t2)) / p;
}
// initcg.f90:32-49
EP_DEV double wcg_dev(double p, double h) {
  if (!(p > 0.0)) return 1.0;
  double arg = 2.0 * 3.14159265358979323846 * p * h;
  double aux = sin(arg) - arg * cos(arg);
  if (fabs(aux) < 1e-10) return 1.0;
  return 3.0 * aux / (arg * arg * arg);
}
// frequency of index i on an axis (DFT4He3d.f90:269-290)
EP_DEV double pfreq_dev(int i, int n, double h) {
  double hp = 1.0 / (n * h), pmax = 1.0 / (2.0 * h);
  return i <= n / 2 ? hp * i : -pmax + hp * (i - n / 2);
}

struct KernParams {
  int nx, ny, nz, nxh;
  double hx, hy, hz, eps, sigma, hc, l;
  int nxp;  // padded row length of the half grid
};
// fvlj4 (fforma), wcgk (initcg), kalfs (DFT4He3d.f90:396-400) on the half grid.
__global__ void k_kspace(KernParams kp, double* wcgk, double* fvlj, double* kalfs) {
  long M = (long)kp.nxp * kp.ny * kp.nz;
  for (long k = blockIdx.x * (long)blockDim.x + threadIdx.x; k < M; k += (long)gridDim.x * blockDim.x) {
    int ix = (int)(k % kp.nxp);
    long r = k / kp.nxp;
    if (ix >= kp.nxh) {  // row padding
      fvlj[k] = 0.0; wcgk[k] = 0.0; kalfs[k] = 0.0;
      continue;
    }
    int iy = (int)(r % kp.ny), iz = (int)(r / kp.ny);
    double px = (1.0 / (kp.nx * kp.hx)) * ix;
    double py = pfreq_dev(iy, kp.ny, kp.hy), pz = pfreq_dev(iz, kp.nz, kp.hz);
    double pm = sqrt(px * px + py * py + pz * pz);
    fvlj[k] = fvlj_dev(pm, kp.eps, kp.sigma, kp.hc);
    wcgk[k] = wcg_dev(pm, kp.hc);
    double t = 3.14159265358979323846 * kp.l * pm;
    kalfs[k] = exp(-(t * t));
  }
}

// ---- impurity: table (potenimpini.f90:35-58) and field (updatepoten /
// update / potenimp cap, potenimp.f90:515-632, 184-186) ----------------------
struct ImpParams {
  int kind;
  double D, re, a, r_cutoff, umax, cap;
  double rx, ry, rz;
};
EP_DEV double select_pot_dev(const ImpParams& ip, double r) {
  if (ip.kind == 0) return 0.0;  // 'null' is 0 on both sides of r_cutoff (V_impur.f90:171-176)
  if (r < ip.r_cutoff) return ip.umax;
  double e = exp(-ip.a * (r - ip.re));
  return ip.D * (e * e - 2.0 * e);
}
__global__ void k_imp_table(ImpParams ip, int npot, double delinter, double* tab) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i <= npot + 1; i += gridDim.x * blockDim.x) {
    int j = i == 0 ? 1 : (i == npot + 1 ? npot : i);
    tab[i] = select_pot_dev(ip, (double)(j - 1) * delinter);
  }
}
template <typename T>
__global__ void k_imp_field(ImpParams ip, int nx, int ny, int nz, double x0, double y0, double z0, double hx,
                            double hy, double hz, int npot, double rmax, const double* f, T* uimp) {
  long N = (long)nx * ny * nz;
  const double h = rmax / (double)(npot - 1);
  const double c1oh = 1.0 / h, c1o2h2 = c1oh * c1oh * 0.5;
  for (long k = blockIdx.x * (long)blockDim.x + threadIdx.x; k < N; k += (long)gridDim.x * blockDim.x) {
    int ix = (int)(k % nx);
    long rr = k / nx;
    int iy = (int)(rr % ny), iz = (int)(rr / ny);
    double zt = (z0 + hz * iz - ip.rz) * (z0 + hz * iz - ip.rz);
    double yt = (y0 + hy * iy - ip.ry) * (y0 + hy * iy - ip.ry) + zt;
    double r2 = (x0 + hx * ix - ip.rx) * (x0 + hx * ix - ip.rx) + yt;
    double r = sqrt(r2);
    int i0 = (int)(1.5 + r / h);
    if (i0 > npot) i0 = npot;  // beyond the table (r > rmax): last node
    double r0 = (double)(i0 - 1) * h;
    double f0 = f[i0], fm = f[i0 - 1], fp = f[i0 + 1];
    double aa = r < r0 ? (fp - 2.0 * f0 + fm) * c1o2h2 : (fm - 2.0 * f0 + fp) * c1o2h2;
    double bb = (fp - f0) * c1oh - aa * (2.0 * r0 + h);
    double cc = f0 - r0 * (bb + r0 * aa);
    double u = r * (aa * r + bb) + cc;
    uimp[k] = (T)(u < ip.cap ? u : ip.cap);
  }
}

// ---- pot4 assembly (poten.f90:56-126; term_alfa.f90:216-236) ----------------
template <typename T> struct Pot4Args {
  long N, slot;
  const int* active;
  int otc;
  T a0, a1, den0s, h2o2m4, alphas, C, beta, den_m;
  const T *delj4, *dencg, *sto1, *ua, *div, *den;
  T* pot4; T* ualphas;
  const T* corr;  // defect correction added to pot4 (FP32 fields of a mixed context), or null
  const double* dscale;  // [slot] factor den is read with (the step's density to n4 atoms), or null
};
template <typename T> __global__ void k_pot4(Pot4Args<T> a) {
  const int s = blockIdx.y;
  if (!a.active[s]) return;
  const long so = (long)s * a.slot;
  const T sc = a.dscale ? (T)a.dscale[s] : T(1);
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < a.N; i += (long)gridDim.x * blockDim.x) {
    const long k = so + i;
    T cg = a.dencg[k], d = sc * a.den[k];
    T p = (a.delj4[k] + cg * cg * (a.a0 + a.a1 * cg)) + a.sto1[k];
    if (a.otc) {
      T ua = ((a.ua[k] / a.den0s + a.div[k]) * a.h2o2m4) * a.alphas;
      a.ualphas[k] = ua;
      p = p + ua;
    }
    T t = tanh(a.beta * (d - a.den_m));
    T pen = a.C * ((T(1) + t) + a.beta * d * (T(1) - t * t));
    a.pot4[k] = a.corr ? (p + pen) + a.corr[k] : p + pen;
  }
}

// array[slot][i] = 0 for the slots in the mask (a kernel, so it can sit in code
// that callers may capture into a graph)
template <typename T> __global__ void k_zero_slots(T* a, const int* active, long N) {
  const int s = blockIdx.y;
  if (!active[s]) return;
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < N; i += (long)gridDim.x * blockDim.x)
    a[(long)s * N + i] = T(0);
}

// Defect correction of FP32 fields from the FP64 ones at the same state:
// corr += (pot4_64 + uimp_64) - (pot4_32 + uimp_32), evaluated in FP64 (pot4_32
// already holds the previous correction), then pot4_32 = pot4_64 + uimp_64 -
// uimp_32 so the FP32 step sees the FP64 potential up to one rounding.
__global__ void k_defect(const double* p64, const double* u64, float* p32, const float* u32, float* corr,
                         const int* active, long N) {
  const int s = blockIdx.y;
  if (!active[s]) return;
  const long so = (long)s * N;
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < N; i += (long)gridDim.x * blockDim.x) {
    const long k = so + i;
    const double want = p64[k] + u64[k];
    corr[k] = (float)((double)corr[k] + (want - ((double)p32[k] + (double)u32[k])));
    p32[k] = (float)(want - (double)u32[k]);
  }
}

// ---- renormalisation: psi *= scale[s]; den = |psi|^2 -------------------------
template <typename T>
__global__ void k_renorm(cx<T>* psi, T* den, const double* scale, const int* active, long N, long slot) {
  const int s = blockIdx.y;
  if (!active[s]) return;
  const T f = (T)scale[s];
  const long so = (long)s * slot;
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < N; i += (long)gridDim.x * blockDim.x) {
    cx<T> p = psi[so + i];
    p = f * p;
    psi[so + i] = p;
    den[so + i] = p.x * p.x + p.y * p.y;
  }
}

// ---- deterministic reduction: partial[s][nblk][K] -> out[s][K] -------------
// mode 1: also scale[s] = sqrt(n4[s] / (out[s][1] * dxyz)).
// ---- buffer roles of the psi / H psi history --------------------------------
enum Role { R_CUR = 0, R_O1, R_O2, R_O3, R_SPARE, R_H1, R_H2, R_OUT };
enum RingOp { RING_RESET = 0, RING_BEFORE_RK0, RING_AFTER_RK0, RING_BEFORE_RKJ, RING_AFTER_RKJ, RING_AFTER_PRED,
              RING_AFTER_CORR };

// Role rotation of one slot's buffers (the reference's ioldp/ioldh rotation).
EP_HD void ring_apply(int* r, int op) {
  switch (op) {
    case RING_RESET:
      r[R_CUR] = 0; r[R_O1] = 1; r[R_O2] = 2; r[R_O3] = 3; r[R_SPARE] = 4; r[R_H1] = 0; r[R_H2] = 1; r[R_OUT] = 0;
      break;
    case RING_BEFORE_RK0: {
      r[R_OUT] = r[R_O3]; r[R_O3] = r[R_O2]; r[R_O2] = r[R_O1]; r[R_O1] = r[R_CUR];
      int t = r[R_H1]; r[R_H1] = r[R_H2]; r[R_H2] = t;
    } break;
    case RING_AFTER_RK0: r[R_CUR] = r[R_OUT]; break;
    case RING_BEFORE_RKJ: r[R_OUT] = r[R_SPARE]; break;
    case RING_AFTER_RKJ: r[R_SPARE] = r[R_CUR]; r[R_CUR] = r[R_OUT]; break;
    case RING_AFTER_PRED: {
      int t = r[R_CUR]; r[R_CUR] = r[R_O3]; r[R_O3] = t;
      t = r[R_H1]; r[R_H1] = r[R_H2]; r[R_H2] = t;
    } break;
    case RING_AFTER_CORR: {
      int fin = r[R_SPARE], mod = r[R_CUR], n = r[R_O3];
      r[R_O3] = r[R_O2]; r[R_O2] = r[R_O1]; r[R_O1] = n; r[R_SPARE] = mod; r[R_CUR] = fin;
    } break;
    default: break;
  }
}

template <int K>
__global__ __launch_bounds__(kThreads) void k_reduce(const double* partial, int nblk, double* out,
                                                      const int* active, int mode, const double* n4, double dxyz,
                                                      double* scale, int* ring, int op1, int op2) {
  __shared__ double red[4 * K];
  const int s = blockIdx.x;
  if (!active[s]) return;
  double v[K];
#pragma unroll
  for (int k = 0; k < K; k++) v[k] = 0.0;
  for (int b = threadIdx.x; b < nblk; b += kThreads) {
    const double* p = partial + ((long)s * nblk + b) * K;
#pragma unroll
    for (int k = 0; k < K; k++) v[k] += p[k];
  }
  block_sum<K>(v, red);
  if (threadIdx.x == 0) {
#pragma unroll
    for (int k = 0; k < K; k++) out[s * K + k] = v[k];
    if (mode == 1) scale[s] = sqrt(n4[s] / (v[1] * dxyz));
    if (mode == 2) scale[s] = n4[s] / (v[1] * dxyz);  // density factor to n4 atoms
    if (ring) {  // role rotations of the step that produced these sums
      ring_apply(ring + 8 * s, op1);
      ring_apply(ring + 8 * s, op2);
    }
  }
}

// den *= f[slot]: the density of a stage or predicted state scaled to n4
// atoms (f from k_reduce mode 2) before the step's poten.
template <typename P>
__global__ void k_scale_den(P* den, const double* f, const int* active, long N) {
  const int s = blockIdx.y;
  if (!active[s]) return;
  const double fs = f[s];
  P* d = den + (long)s * N;
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < N; i += (long)gridDim.x * blockDim.x)
    d[i] = (P)((double)d[i] * fs);
}

// Factor pairs of the packed column: kw[kz][ky] = ((K0 + KH)/2, (K0 - KH)/2)
// with K0, KH the k-space factor at kx = 0 and kx = P (= nx/2) of the padded
// table k (rows of nxp), in the precision of k.
template <typename T>
__global__ void k_pair_tables(const T* k, T* kw, int nxp, int P, int ny, int nz) {
  const long n = (long)ny * nz;
  for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < n; e += (long)gridDim.x * blockDim.x) {
    const long kz = e / ny, ky = e - kz * ny;
    const T k0 = k[ky * nxp + kz * (long)nxp * ny], kh = k[P + ky * nxp + kz * (long)nxp * ny];
    kw[2 * e] = T(0.5) * (k0 + kh);
    kw[2 * e + 1] = T(0.5) * (k0 - kh);
  }
}

// ---- precision conversion (FP32 phase -> FP64 polish) -----------------------
template <typename A, typename B>
__global__ void k_convert(const A* in, B* out, long n) {
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
    out[i] = (B)in[i];
}

// scan for non-finite or non-positive density (steprk.f90:321-340 guard)
template <typename T>
__global__ void k_check_den(const T* den, long N, long slot, const int* active, int* bad) {
  const int s = blockIdx.y;
  if (!active[s]) return;
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < N; i += (long)gridDim.x * blockDim.x) {
    T d = den[(long)s * slot + i];
    if (!(d > T(0)) || !(d < T(1e30))) { bad[s] = 1; }
  }
}

}  // namespace ep
