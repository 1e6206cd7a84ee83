// Anisotropic impurity in a P or D internal state (Lstate 'P' / 'D').
//
// potenimpini.f90:11-72 tabulates the radial potentials (P: V_pi and
// V_sigma - V_pi; D: V_delta, V_pi - V_delta, V_sigma - V_delta) on the
// npot-point grid; updatepoten (potenimp.f90:515-575) interpolates them at
// |r - rimp| with the parabola of update() and divides the anisotropic parts
// by r^2 (P) or r^4 (D); instates (instates.f90:1-800) builds the
// density-weighted potential matrix plus the spin-orbit matrix, diagonalises
// it and picks invar; potenimp (potenimp.f90:34-180) evaluates the He field of
// that internal state, capped at 1000 K.
//
// Here one kernel reduces the density-weighted moments of every slot (the
// grid half of instates), the host builds and diagonalises the 6x6 / 10x10
// Hermitian matrix (Jacobi rotations), and one kernel writes uimp.  The
// fields of updatepoten are recomputed from the tables where they are needed
// instead of being stored (three more N-sized arrays per slot otherwise).
#pragma once
#include <cmath>
#include <complex>
#include <cstring>

#include "common.hpp"
#include "plus.hpp"
#include "stencil.hpp"

namespace ep {

// update() with Lparab (potenimp.f90:588-632) on a table f[0..npot+1]; the
// index is clamped at npot (the reference reads past its table beyond rmax).
EP_DEV double update_dev(double r, int npot, double h, double c1oh, double c1o2h2, const double* f) {
  int i0 = (int)(1.5 + r / h);
  if (i0 > npot) i0 = npot;
  const double r0 = (double)(i0 - 1) * h;
  const double f0 = f[i0], fm = f[i0 - 1], fp = f[i0 + 1];
  const double aa = r < r0 ? (fp - 2.0 * f0 + fm) * c1o2h2 : (fm - 2.0 * f0 + fp) * c1o2h2;
  const double bb = (fp - f0) * c1oh - aa * (2.0 * r0 + h);
  const double cc = f0 - r0 * (bb + r0 * aa);
  return r * (aa * r + bb) + cc;
}

// Moments of instates: P (instates.f90:136-157) A, Vx, Vy, Vz, Vxy, Vxz, Vyz;
// D (:455-519) MP(i<=j), MS(i<=j) row-major over the 5x5 upper triangles, then
// the diagonal sum of Delta.
constexpr int kMomP = 7, kMomD = 31, kStateBlocks = 256;
constexpr double kZeroR = 1e-3;  // updatepoten's zeror

struct StateArgs {
  int state;  // EFFPOT_STATE_P | EFFPOT_STATE_D
  int nx, ny, nz;
  double x0, y0, z0, hx, hy, hz;
  int npot;
  double h, c1oh, c1o2h2;
  const double* tab;  // [3][npot+2] of this slot
  double umax_pi, umax_sigma;
  double rx, ry, rz;
  double inv[20];  // invar (complex, interleaved)
  double cap;
};

// updatepoten fields at one point: P (t0, t1) = (VPi, Delta/r^2);
// D (t0, t1, t2) = (Delta, Pi_Del/r^4, Sig_Del/r^4) (potenimp.f90:531-565).
EP_DEV void state_fields(const StateArgs& a, double r2, double& t0, double& t1, double& t2) {
  const double r = sqrt(r2);
  const int L = a.npot + 2;
  if (a.state == EFFPOT_STATE_P) {
    t0 = update_dev(r, a.npot, a.h, a.c1oh, a.c1o2h2, a.tab);
    t1 = r > kZeroR ? update_dev(r, a.npot, a.h, a.c1oh, a.c1o2h2, a.tab + L) / r2 : a.umax_sigma;
    t2 = 0.0;
  } else {
    const double r4 = r2 * r2;
    t0 = update_dev(r, a.npot, a.h, a.c1oh, a.c1o2h2, a.tab);
    t1 = r > kZeroR ? update_dev(r, a.npot, a.h, a.c1oh, a.c1o2h2, a.tab + L) / r4 : a.umax_pi;
    t2 = r > kZeroR ? update_dev(r, a.npot, a.h, a.c1oh, a.c1o2h2, a.tab + 2 * L) / r4 : a.umax_sigma;
  }
}

// The five d-orbital quadratic forms of instates / potenimp (:466-499):
// armonicd d1..d5 and the upper triangle of MP (the Sig_Del-weighted matrix).
struct DForms {
  double d[6];   // d1..d5, d0
  double mp[15]; // MP(i<=j) row-major, without the radial factor
};
EP_DEV void d_forms(double xx, double yy, double zz, DForms& f) {
  const double sqr3 = 1.7320508075688772, sqr3o2 = 0.8660254037844386;
  const double insqr3 = 0.5773502691896258, in3 = 1.0 / 3.0;
  const double d0 = xx * xx + yy * yy + zz * zz, d1 = sqr3 * xx * yy, d2 = sqr3 * yy * zz;
  const double d3 = (-xx * xx - yy * yy + 2.0 * zz * zz) / 2.0, d4 = sqr3 * xx * zz, d5 = sqr3o2 * (xx * xx - yy * yy);
  const double sd1 = d1 * d1, sd2 = d2 * d2, sd4 = d4 * d4, sd5 = d5 * d5;
  const double s35 = d3 + insqr3 * d5, f35 = d3 - insqr3 * d5;
  f.d[0] = d1; f.d[1] = d2; f.d[2] = d3; f.d[3] = d4; f.d[4] = d5; f.d[5] = d0;
  f.mp[0] = in3 * (sd2 + sd4 + 4.0 * sd5);               // (1,1)
  f.mp[1] = -4.0 * in3 * d1 * d2 + insqr3 * d4 * d0;     // (1,2)
  f.mp[2] = -2.0 * insqr3 * d2 * d4;                     // (1,3)
  f.mp[3] = -4.0 * in3 * d1 * d4 + insqr3 * d2 * d0;     // (1,4)
  f.mp[4] = -4.0 * in3 * d1 * d5;                        // (1,5)
  f.mp[5] = in3 * (sd1 + sd4) + s35 * s35;               // (2,2)
  f.mp[6] = insqr3 * d1 * d4 - d2 * s35;                 // (2,3)
  f.mp[7] = -4.0 * in3 * d2 * d4 + insqr3 * d1 * d0;     // (2,4)
  f.mp[8] = -d1 * d4 - insqr3 * d2 * s35;                // (2,5)
  f.mp[9] = sd4 + sd2;                                   // (3,3)
  f.mp[10] = insqr3 * d1 * d2 - d4 * f35;                // (3,4)
  f.mp[11] = insqr3 * (sd2 - sd4);                       // (3,5)
  f.mp[12] = in3 * (sd1 + sd2) + f35 * f35;              // (4,4)
  f.mp[13] = d1 * d2 + insqr3 * d4 * f35;                // (4,5)
  f.mp[14] = in3 * (sd2 + sd4 + 4.0 * sd1);              // (5,5)
}

// Table of one radial potential (mode 0: va) or of a difference (mode 1:
// vb - va), potenimpini.f90:35-58, with the end copies of :60-66.
__global__ void k_state_table(effpot_plus va, effpot_plus vb, int mode, int npot, double delinter, double* tab) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i <= npot + 1; i += gridDim.x * blockDim.x) {
    const int j = i == 0 ? 1 : (i == npot + 1 ? npot : i);
    const double r = (double)(j - 1) * delinter;
    const double a = plus_pot(va, r);
    tab[i] = mode == 0 ? a : plus_pot(vb, r) - a;
  }
}

// Density-weighted moments of one slot: block partials [kStateBlocks][K].
template <typename V, int K>
__global__ __launch_bounds__(256) void k_state_moments(const V* psi, StateArgs a, double* partial) {
  __shared__ double red[4 * K];
  double v[K];
#pragma unroll
  for (int k = 0; k < K; k++) v[k] = 0.0;
  const long N = (long)a.nx * a.ny * a.nz;
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < N; i += (long)gridDim.x * blockDim.x) {
    const int ix = (int)(i % a.nx);
    const long rr = i / a.nx;
    const int iy = (int)(rr % a.ny), iz = (int)(rr / a.ny);
    const double xx = (a.x0 + a.hx * ix) - a.rx, yy = (a.y0 + a.hy * iy) - a.ry, zz = (a.z0 + a.hz * iz) - a.rz;
    const double den = (double)norm2(psi[i]);
    double t0, t1, t2;
    state_fields(a, (xx * xx + yy * yy) + zz * zz, t0, t1, t2);
    if constexpr (K == kMomP) {  // instates.f90:145-153
      v[0] += den;
      v[1] += den * (t0 + t1 * xx * xx);
      v[2] += den * (t0 + t1 * yy * yy);
      v[3] += den * (t0 + t1 * zz * zz);
      v[4] += den * t1 * xx * yy;
      v[5] += den * t1 * xx * zz;
      v[6] += den * t1 * yy * zz;
    } else {  // instates.f90:466-517: MP weighted by Sig_Del, MS by Pi_Del
      DForms f;
      d_forms(xx, yy, zz, f);
      const double dv1 = t1 * den, dv2 = t2 * den;
#pragma unroll
      for (int k = 0; k < 15; k++) v[k] += f.mp[k] * dv2;
      int k = 15;
#pragma unroll
      for (int p = 0; p < 5; p++)
#pragma unroll
        for (int q = p; q < 5; q++) v[k++] += (f.d[p] * f.d[q]) * dv1;
      v[30] += t0 * den;
    }
  }
  block_sum<K>(v, red);
  if (threadIdx.x == 0) {
#pragma unroll
    for (int k = 0; k < K; k++) partial[(long)blockIdx.x * K + k] = v[k];
  }
}

// potenimp (:34-45 P, :54-158 D) and the cap (:184-186) at grid point i.
EP_DEV double state_uimp_at(const StateArgs& a, long i) {
  const int ix = (int)(i % a.nx);
  const long rr = i / a.nx;
  const int iy = (int)(rr % a.ny), iz = (int)(rr / a.ny);
  const double xx = (a.x0 + a.hx * ix) - a.rx, yy = (a.y0 + a.hy * iy) - a.ry, zz = (a.z0 + a.hz * iz) - a.rz;
  double t0, t1, t2;
  state_fields(a, (xx * xx + yy * yy) + zz * zz, t0, t1, t2);
  double u;
  if (a.state == EFFPOT_STATE_P) {
    // odd (spin up) and even (spin down) components of iv . r
    const double ore = a.inv[0] * xx + a.inv[4] * yy + a.inv[8] * zz;
    const double oim = a.inv[1] * xx + a.inv[5] * yy + a.inv[9] * zz;
    const double ere = a.inv[2] * xx + a.inv[6] * yy + a.inv[10] * zz;
    const double eim = a.inv[3] * xx + a.inv[7] * yy + a.inv[11] * zz;
    u = t0 + t1 * ((ore * ore + oim * oim) + (ere * ere + eim * eim));
  } else {
    DForms f;
    d_forms(xx, yy, zz, f);
    // M = Delta*I + Pi_Del*(d d^T) + Sig_Del*MP; u = sum_ij M_ij Re(conj(c_i) c_j)
    // over both spin components, c_i = iv(2i-1) (up) and iv(2i) (down)
    u = 0.0;
    int k = 0;
#pragma unroll
    for (int p = 0; p < 5; p++)
#pragma unroll
      for (int q = p; q < 5; q++, k++) {
        double m = f.mp[k] * t2 + (f.d[p] * f.d[q]) * t1;
        if (p == q) m += t0;
        const double* cp = a.inv + 4 * p;
        const double* cq = a.inv + 4 * q;
        const double g = (cp[0] * cq[0] + cp[1] * cq[1]) + (cp[2] * cq[2] + cp[3] * cq[3]);
        u += (p == q ? 1.0 : 2.0) * m * g;
      }
  }
  return u < a.cap ? u : a.cap;
}
template <typename T>
__global__ __launch_bounds__(256) void k_state_uimp(StateArgs a, T* uimp) {
  const long N = (long)a.nx * a.ny * a.nz;
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < N; i += (long)gridDim.x * blockDim.x)
    uimp[i] = (T)state_uimp_at(a, i);
}

// ---- host half of instates -------------------------------------------------
using zd = std::complex<double>;

// V + (Als/2) SO of the moments (instates.f90:159-195 P, :520-565 D; the
// same matrix potinvar applies, potenimp.f90:318-488), n x n row-major, real
// and imaginary parts apart so that the device can build it too.
__host__ __device__ inline int state_matrix_ri(int state, const double* m, double als, double* Hr, double* Hi) {
  const int n = state == EFFPOT_STATE_P ? 6 : 10;
  for (int i = 0; i < n * n; i++) { Hr[i] = 0.0; Hi[i] = 0.0; }
  const double h = 0.5 * als;  // Also2
  // upper-triangle spin-orbit entries (1-based i < j), value h * (re + i im)
  auto so = [&](int i, int j, double re, double im) {
    Hr[(i - 1) * n + (j - 1)] += h * re; Hi[(i - 1) * n + (j - 1)] += h * im;
    Hr[(j - 1) * n + (i - 1)] += h * re; Hi[(j - 1) * n + (i - 1)] -= h * im;
  };
  if (state == EFFPOT_STATE_P) {
    const double Vx = m[1], Vy = m[2], Vz = m[3], Vxy = m[4], Vxz = m[5], Vyz = m[6];
    const double d[3] = {Vx, Vy, Vz};
    for (int a = 0; a < 3; a++) { Hr[(2 * a) * n + 2 * a] = d[a]; Hr[(2 * a + 1) * n + 2 * a + 1] = d[a]; }
    auto sym = [&](int i, int j, double v) { Hr[(i - 1) * n + (j - 1)] = v; Hr[(j - 1) * n + (i - 1)] = v; };
    sym(1, 3, Vxy); sym(2, 4, Vxy); sym(1, 5, Vxz); sym(2, 6, Vxz); sym(3, 5, Vyz); sym(4, 6, Vyz);
    // SO(1,3)=-i SO(2,4)=i SO(2,5)=-1 SO(1,6)=1 SO(4,5)=-i SO(3,6)=-i
    so(1, 3, 0.0, -1.0); so(2, 4, 0.0, 1.0); so(2, 5, -1.0, 0.0); so(1, 6, 1.0, 0.0);
    so(4, 5, 0.0, -1.0); so(3, 6, 0.0, -1.0);
  } else {
    double mp[5][5], ms[5][5];
    int k = 0;
    for (int p = 0; p < 5; p++)
      for (int q = p; q < 5; q++, k++) { mp[p][q] = m[k]; mp[q][p] = m[k]; }
    for (int p = 0; p < 5; p++)
      for (int q = p; q < 5; q++, k++) { ms[p][q] = m[k]; ms[q][p] = m[k]; }
    for (int p = 0; p < 5; p++)
      for (int q = 0; q < 5; q++) {
        const double v = (p == q ? m[30] : 0.0) + ms[p][q] + mp[p][q];
        Hr[(2 * p) * n + 2 * q] = v;          // MatD(2i-1, 2j-1)
        Hr[(2 * p + 1) * n + 2 * q + 1] = v;  // MatD(2i, 2j)
      }
    const double s3 = 1.7320508075688772;
    so(1, 4, 1.0, 0.0); so(1, 8, 0.0, -1.0); so(1, 9, 0.0, 2.0); so(2, 3, -1.0, 0.0); so(2, 7, 0.0, -1.0);
    so(2, 10, 0.0, -2.0); so(3, 6, 0.0, -s3); so(3, 7, 0.0, 1.0); so(3, 10, 0.0, -1.0); so(4, 5, 0.0, -s3);
    so(4, 8, 0.0, -1.0); so(4, 9, 0.0, -1.0); so(5, 8, -s3, 0.0); so(6, 7, s3, 0.0); so(7, 10, -1.0, 0.0);
    so(8, 9, 1.0, 0.0);
  }
  return n;
}
inline int state_matrix(int state, const double* m, double als, zd* H) {
  double Hr[100], Hi[100];
  const int n = state_matrix_ri(state, m, als, Hr, Hi);
  for (int i = 0; i < n * n; i++) H[i] = zd(Hr[i], Hi[i]);
  return n;
}

// (Als/2) <v|SO|v>: energy.f90:170-192 (P and D).
inline double state_eso(int state, const double* inv, double als) {
  const int n = state == EFFPOT_STATE_P ? 6 : 10;
  zd H[100];
  double zero[kMomD] = {};
  state_matrix(state, zero, als, H);  // with zero moments H = (Als/2) SO
  zd acc = 0.0;
  for (int i = 0; i < n; i++) {
    zd hv = 0.0;
    for (int j = 0; j < n; j++) hv += H[i * n + j] * zd(inv[2 * j], inv[2 * j + 1]);
    acc += std::conj(zd(inv[2 * i], inv[2 * i + 1])) * hv;
  }
  return acc.real();
}

// Eigenvalues (ascending) and eigenvectors (columns of V) of a Hermitian
// n x n matrix by cyclic complex Jacobi rotations: each rotation first
// removes the phase of a_pq, then is the real symmetric Jacobi rotation of
// the 2x2 block (Numerical Recipes' tan formula).
inline void herm_eigen(int n, zd* A, double* w, zd* V) {
  for (int i = 0; i < n * n; i++) V[i] = (i / n == i % n) ? 1.0 : 0.0;
  for (int sweep = 0; sweep < 64; sweep++) {
    double off = 0.0, diag = 0.0;
    for (int p = 0; p < n; p++) {
      diag += std::norm(A[p * n + p]);
      for (int q = p + 1; q < n; q++) off += std::norm(A[p * n + q]);
    }
    if (off == 0.0 || off <= 1e-33 * (diag + 2.0 * off)) break;
    for (int p = 0; p < n; p++)
      for (int q = p + 1; q < n; q++) {
        const zd apq = A[p * n + q];
        const double g = std::abs(apq);
        if (g == 0.0) continue;
        const zd e = apq / g;  // e^{i phi}
        const double app = A[p * n + p].real(), aqq = A[q * n + q].real();
        const double th = (aqq - app) / (2.0 * g);
        const double t = (th >= 0.0 ? 1.0 : -1.0) / (std::fabs(th) + std::sqrt(th * th + 1.0));
        const double c = 1.0 / std::sqrt(t * t + 1.0), s = t * c;
        const zd ec = std::conj(e);
        // U = [[c, s], [-s e^{-i phi}, c e^{-i phi}]] on (p, q); A <- U^H A U, V <- V U
        for (int k = 0; k < n; k++) {  // columns
          const zd akp = A[k * n + p], akq = A[k * n + q];
          A[k * n + p] = c * akp - s * ec * akq;
          A[k * n + q] = s * akp + c * ec * akq;
          const zd vkp = V[k * n + p], vkq = V[k * n + q];
          V[k * n + p] = c * vkp - s * ec * vkq;
          V[k * n + q] = s * vkp + c * ec * vkq;
        }
        for (int k = 0; k < n; k++) {  // rows
          const zd apk = A[p * n + k], aqk = A[q * n + k];
          A[p * n + k] = c * apk - s * e * aqk;
          A[q * n + k] = s * apk + c * e * aqk;
        }
        A[p * n + q] = A[q * n + p] = 0.0;
        A[p * n + p] = A[p * n + p].real();
        A[q * n + q] = A[q * n + q].real();
      }
  }
  int idx[10];
  for (int i = 0; i < n; i++) { idx[i] = i; w[i] = A[i * n + i].real(); }
  for (int i = 1; i < n; i++)  // insertion sort, ascending
    for (int j = i; j > 0 && w[idx[j]] < w[idx[j - 1]]; j--) std::swap(idx[j], idx[j - 1]);
  double ws[10];
  zd Vs[100];
  for (int k = 0; k < n; k++) {
    ws[k] = w[idx[k]];
    for (int i = 0; i < n; i++) Vs[i * n + k] = V[i * n + idx[k]];
  }
  std::memcpy(w, ws, sizeof(double) * n);
  for (int i = 0; i < n * n; i++) V[i] = Vs[i];
}

// instates' default selection (:360-364 P, :722-726 D): the equal-weight
// sum of the two vectors of Kramers pair `instate` (0 = lowest).
inline void state_select(int n, const zd* V, int instate, double* inv) {
  const double c = 1.0 / std::sqrt(2.0);
  std::memset(inv, 0, sizeof(double) * 20);
  for (int i = 0; i < n; i++) {
    const zd v = c * V[i * n + 2 * instate] + c * V[i * n + 2 * instate + 1];
    inv[2 * i] = v.real();
    inv[2 * i + 1] = v.imag();
  }
}

// Jz = Lz + Sz in the instates basis (even 0-based components spin up).  P:
// orbitals (x, y, z) with <x|Lz|y> = -i (instates.f90:239-248); D: orbitals
// 1..5 with Lz coupling 1 and 5 (m = +-2, weight 2) and 2 and 4 (m = +-1),
// orbital 3 m = 0 (instates.f90:663-680).  Row-major n x n.
inline void state_jz(int state, zd* J) {
  const int n = state == EFFPOT_STATE_P ? 6 : 10;
  for (int i = 0; i < n * n; i++) J[i] = 0.0;
  for (int i = 0; i < n; i++) J[i * n + i] = (i % 2 == 0) ? 0.5 : -0.5;
  auto lz = [&](int a, int b, double w) {  // orbitals a, b (0-based): <a|Lz|b> = -i w
    for (int s = 0; s < 2; s++) {
      J[(2 * a + s) * n + 2 * b + s] = zd(0.0, -w);
      J[(2 * b + s) * n + 2 * a + s] = zd(0.0, w);
    }
  };
  if (state == EFFPOT_STATE_P) {
    lz(0, 1, 1.0);
  } else {
    lz(4, 0, 2.0);
    lz(1, 3, -1.0);
  }
}

// The Jz manifold of Ldiag_jz: first eigenvector (0-based) and count, and
// the reference's default Ljz (as 2 jz) for `instate` (instates.f90:229-234,
// 259-274 for P; 614-661 for D).
inline void jz_manifold(int state, int instate, int* k0, int* nj, int* jz2_default) {
  if (state == EFFPOT_STATE_P) {
    if (instate == 0) { *k0 = 0; *nj = 2; *jz2_default = -1; }
    else { *k0 = 2; *nj = 4; *jz2_default = instate == 1 ? -3 : -1; }
  } else {
    if (instate <= 1) { *k0 = 0; *nj = 4; *jz2_default = instate == 0 ? -3 : -1; }
    else { *k0 = 4; *nj = 6; *jz2_default = instate == 2 ? -5 : (instate == 3 ? -3 : -1); }
  }
}

// Ldiag_jz: diagonalise Jz within the eigenvectors k0..k0+nj-1 of V + SO
// (columns of V) and return the eigenvector of the (jz2 + nj - 1)/2-th Jz
// eigenvalue (ascending), expressed in the instates basis.  EFFPOT_ERR_ARG
// for an Ljz outside the manifold (the reference stops there).
inline int state_select_jz(int state, int n, const zd* V, int instate, int jz2, double* inv) {
  int k0, nj, jdef;
  jz_manifold(state, instate, &k0, &nj, &jdef);
  if (jz2 == 0) jz2 = jdef;
  if ((jz2 & 1) == 0 || jz2 < -(nj - 1) || jz2 > nj - 1) return EFFPOT_ERR_ARG;
  const int il = (jz2 + nj - 1) / 2;
  zd J[100], M[36], W[36];
  double w[10];
  state_jz(state, J);
  for (int k = 0; k < nj; k++)
    for (int l = 0; l < nj; l++) {
      zd acc = 0.0;
      for (int a = 0; a < n; a++) {
        zd jv = 0.0;
        for (int b = 0; b < n; b++) jv += J[a * n + b] * V[b * n + k0 + l];
        acc += std::conj(V[a * n + k0 + k]) * jv;
      }
      M[k * nj + l] = acc;
    }
  herm_eigen(nj, M, w, W);
  std::memset(inv, 0, sizeof(double) * 20);
  for (int i = 0; i < n; i++) {
    zd v = 0.0;
    for (int k = 0; k < nj; k++) v += W[k * nj + il] * V[i * n + k0 + k];
    inv[2 * i] = v.real();
    inv[2 * i + 1] = v.imag();
  }
  return EFFPOT_OK;
}

// Lexcite_state_fix (instates.f90:301-327): a fixed spin-up P orbital.
inline void state_select_fix(int instate, double* inv) {
  std::memset(inv, 0, sizeof(double) * 20);
  const int m = instate >= 2 ? 3 : instate + 1;
  for (int o = 0; o < m; o++) inv[4 * o] = 1.0 / std::sqrt((double)m);
}

// Laverage_P_value (:328-335) and Lmixture (:336-358) for P.
inline void state_select_average(int n, const zd* V, double* inv) {
  std::memset(inv, 0, sizeof(double) * 20);
  const double c = 1.0 / std::sqrt((double)n);
  for (int i = 0; i < n; i++) {
    zd v = 0.0;
    for (int k = 0; k < n; k++) v += c * V[i * n + k];
    inv[2 * i] = v.real();
    inv[2 * i + 1] = v.imag();
  }
}
inline void state_select_mixture(int n, const zd* V, double a_pi, double a_sig, double* inv) {
  const double c = 1.0 / std::sqrt(2.0);
  const double amp[3] = {std::sqrt(std::max(1.0 - a_pi - a_sig, 0.0)), std::sqrt(a_pi), std::sqrt(a_sig)};
  std::memset(inv, 0, sizeof(double) * 20);
  for (int i = 0; i < n; i++) {
    zd v = 0.0;
    for (int p = 0; p < 3; p++) v += amp[p] * (c * V[i * n + 2 * p] + c * V[i * n + 2 * p + 1]);
    inv[2 * i] = v.real();
    inv[2 * i + 1] = v.imag();
  }
}

}  // namespace ep
