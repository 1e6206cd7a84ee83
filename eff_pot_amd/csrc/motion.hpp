// Classical motion of an isotropic impurity through the helium integrators.
//
// steprk.f90:117-150 and steppc.f90:68-116,160-203 advance the impurity
// position and velocity with the same RK-Gill starter and Hamming
// predictor-modifier-corrector as psi, the acceleration being forceimp
// (potenimp.f90:193-258: F = -sum grad(den) uimp dV) over the impurity mass.
// Here the per-slot state (r, v, a, their histories, the index rotations
// ioldr/ioldv/iolda and the pc/q accumulators) lives in device memory and one
// thread per slot runs each update, so a step never waits for the host; the
// field uimp is re-interpolated at the new position by a grid kernel that reads
// the position from that state.
#pragma once
#include "common.hpp"
#include "misc.hpp"
#include "state.hpp"

namespace ep {

struct Motion {
  int32_t on, iso;  // iso: isotropic field from the slot's table (else a P/D state's)
  double m;  // impurity mass in K^-1 A^-2 (mimpur * mp_u, DFT4He3d.f90:203)
  double r[3], v[3], a[3];
  double rold[3][3], vold[3][3], aold[2][3];  // rimpold, vimpold, aimpold: [storage][xyz]
  int32_t ir[3], iv[3], ia[2], pad2_;         // ioldr, ioldv, iolda (0-based storage slots)
  double pcr[3], pcv[3], qr[3], qv[3];
  ImpParams ip;  // the field (its rx, ry, rz are not used: r is)
};

// One RK-Gill stage j (0-based) of r and v (steprk.f90:117-150); after the
// last stage the history indices are the identity again (:178-190).
__global__ void k_motion_rk(Motion* M, const int* active, int B, int j, const double* dtv, double arun, double brun,
                            double crun) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= B || !active[s] || !M[s].on) return;
  Motion& q = M[s];
  const double dt = dtv[s];
  for (int c = 0; c < 3; c++) {  // positions
    const double a2 = arun * (q.v[c] - brun * q.qr[c]);
    q.qr[c] = q.qr[c] + 3.0 * a2 - crun * q.v[c];
    if (j == 0) { q.rold[2][c] = q.rold[1][c]; q.rold[1][c] = q.rold[0][c]; q.rold[0][c] = q.r[c]; }
    q.r[c] = q.r[c] + dt * a2;
  }
  for (int c = 0; c < 3; c++) {  // velocities
    const double a2 = arun * (q.a[c] - brun * q.qv[c]);
    q.qv[c] = q.qv[c] + 3.0 * a2 - crun * q.a[c];
    if (j == 0) {
      q.aold[1][c] = q.aold[0][c]; q.aold[0][c] = q.a[c];
      q.vold[2][c] = q.vold[1][c]; q.vold[1][c] = q.vold[0][c]; q.vold[0][c] = q.v[c];
    }
    q.v[c] = q.v[c] + dt * a2;
  }
  if (j == 3) {
    for (int k = 0; k < 3; k++) { q.ir[k] = k; q.iv[k] = k; }
    q.ia[0] = 0; q.ia[1] = 1;
  }
}

// Predictor + modifier of r and v, then aimpold(iolda(2)) = a and the iolda
// swap (steppc.f90:68-103).  The modified velocity goes to vimpold(ioldv(3)),
// v itself keeps the old value until the corrector.
__global__ void k_motion_pred(Motion* M, const int* active, int B, const double* dtv) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= B || !active[s] || !M[s].on) return;
  Motion& q = M[s];
  const double dt = dtv[s], c4o3 = 4.0 / 3.0, c112 = 112.0 / 121.0;
  for (int c = 0; c < 3; c++) {
    const double p = q.rold[q.ir[2]][c] + c4o3 * dt * (2.0 * q.v[c] - q.vold[q.iv[0]][c] + 2.0 * q.vold[q.iv[1]][c]);
    q.rold[q.ir[2]][c] = q.r[c];
    q.r[c] = p - c112 * q.pcr[c];
    q.pcr[c] = p;
  }
  for (int c = 0; c < 3; c++) {
    const double p = q.vold[q.iv[2]][c] + c4o3 * dt * (2.0 * q.a[c] - q.aold[q.ia[0]][c] + 2.0 * q.aold[q.ia[1]][c]);
    q.vold[q.iv[2]][c] = p - c112 * q.pcv[c];
    q.pcv[c] = p;
  }
  for (int c = 0; c < 3; c++) q.aold[q.ia[1]][c] = q.a[c];
  const int t = q.ia[1]; q.ia[1] = q.ia[0]; q.ia[0] = t;
}

// a = F / m from the reduced sums sum(d_x den * uimp), ... (forceimp).
__global__ void k_motion_force(Motion* M, const int* active, int B, const double* sums, double dxyz) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= B || !active[s] || !M[s].on) return;
  for (int c = 0; c < 3; c++) M[s].a[c] = (-sums[3 * s + c] * dxyz) / M[s].m;
}

// Corrector of r and v and the ioldr / ioldv rotations (steppc.f90:160-203).
__global__ void k_motion_corr(Motion* M, const int* active, int B, const double* dtv) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= B || !active[s] || !M[s].on) return;
  Motion& q = M[s];
  const double dt = dtv[s], c9 = 9.0 / 121.0;
  for (int c = 0; c < 3; c++) {
    const double x = 0.125 * (9.0 * q.rold[q.ir[2]][c] - q.rold[q.ir[1]][c] +
                              3.0 * dt * (q.vold[q.iv[2]][c] + 2.0 * q.v[c] - q.vold[q.iv[0]][c]));
    q.pcr[c] = q.pcr[c] - x;
    q.r[c] = x + c9 * q.pcr[c];
  }
  { const int t = q.ir[2]; q.ir[2] = q.ir[1]; q.ir[1] = q.ir[0]; q.ir[0] = t; }
  for (int c = 0; c < 3; c++) {
    const double x = 0.125 * (9.0 * q.v[c] - q.vold[q.iv[1]][c] +
                              3.0 * dt * (q.a[c] + 2.0 * q.aold[q.ia[0]][c] - q.aold[q.ia[1]][c]));
    q.pcv[c] = q.pcv[c] - x;
    q.vold[q.iv[2]][c] = q.v[c];
    q.v[c] = x + c9 * q.pcv[c];
  }
  { const int t = q.iv[2]; q.iv[2] = q.iv[1]; q.iv[1] = q.iv[0]; q.iv[0] = t; }
}

// uimp of every moving slot at its current position (updatepoten + the cap),
// from the slot's table f[slot][npot+2].
__global__ void k_imp_field_moving(const Motion* M, const int* active, int nx, int ny, int nz, double x0, double y0,
                                   double z0, double hx, double hy, double hz, int npot, double rmax, const double* f,
                                   double* uimp, long N) {
  const int s = blockIdx.y;
  if (!active[s] || !M[s].on || !M[s].iso) return;
  const Motion& q = M[s];
  const double* tab = f + (long)s * (npot + 2);
  const double h = rmax / (double)(npot - 1), c1oh = 1.0 / h, c1o2h2 = c1oh * c1oh * 0.5;
  double* u = uimp + (long)s * N;
  for (long k = blockIdx.x * (long)blockDim.x + threadIdx.x; k < N; k += (long)gridDim.x * blockDim.x) {
    const int ix = (int)(k % nx);
    const long rr = k / nx;
    const int iy = (int)(rr % ny), iz = (int)(rr / ny);
    const double zt = (z0 + hz * iz - q.r[2]) * (z0 + hz * iz - q.r[2]);
    const double yt = (y0 + hy * iy - q.r[1]) * (y0 + hy * iy - q.r[1]) + zt;
    const double r2 = (x0 + hx * ix - q.r[0]) * (x0 + hx * ix - q.r[0]) + yt;
    const double v = update_dev(sqrt(r2), npot, h, c1oh, c1o2h2, tab);
    u[k] = v < q.ip.cap ? v : q.ip.cap;
  }
}

// forceimp partial sums [slot][block][3]: sum d_c den * uimp over the block's
// points (grid: nblk x slots, grid-stride).
// With Lfilter_exciplex_force (potenimp.f90:216-245) F_z sums only the planes
// on the far side of the impurity beyond z_exciplex_exclusion: 1-based iz >=
// iz0 = int((rz - z(1) + excl)/hz + 1.5) when rz < 0, else iz <= nz0 =
// int((rz - z(1) - excl)/hz + 1.5); z1 = z(1), plane = nx ny.
__global__ __launch_bounds__(256) void k_force_partials(const Motion* M, const PlusSlot* P, const int* active,
                                                        const double* dx, const double* dy, const double* dz,
                                                        const double* uimp, long N, long plane, int nz, double z1,
                                                        double hz, double* partial) {
  const int s = blockIdx.y;
  if (!active[s] || !M[s].on) return;
  __shared__ double red[4 * 3];
  double v[3] = {0.0, 0.0, 0.0};
  const long so = (long)s * N;
  int zlo = 0, zhi = nz - 1;  // 0-based planes of F_z
  if (P[s].force_filter) {
    const double rz = M[s].r[2], gap = P[s].excl;
    if (rz < 0.0) {
      zlo = max(1, (int)((rz - z1 + gap) / hz + 1.5)) - 1;
    } else {
      zhi = min(nz, (int)((rz - z1 - gap) / hz + 1.5)) - 1;
    }
  }
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < N; i += (long)gridDim.x * blockDim.x) {
    const double u = uimp[so + i];
    const int iz = (int)(i / plane);
    v[0] += dx[so + i] * u;
    v[1] += dy[so + i] * u;
    if (iz >= zlo && iz <= zhi) v[2] += dz[so + i] * u;
  }
  block_sum<3>(v, red);
  if (threadIdx.x == 0) {
    double* p = partial + ((long)s * gridDim.x + blockIdx.x) * 3;
    p[0] = v[0]; p[1] = v[1]; p[2] = v[2];
  }
}

// ---- internal-state dynamics of a P/D impurity (potinvar) -----------------
// steprk.f90:152-167 and steppc.f90:84-102,205-226 advance invar with the
// same integrators under Hinvar = -i (H(rho) invar - Ev0 invar), H(rho) the
// density-weighted V + (Als/2) SO matrix of instates (potinvar,
// potenimp.f90:261-492); potenimp then rebuilds uimp from the new invar.
struct InvarDyn {
  int32_t on, state, n, pad_;
  double als, ev0;
  double iv[20], hiv[20];                         // invar, Hinvar (complex, interleaved)
  double ivold[3][20], hivold[2][20], pciv[20], qiv[20];
  int32_t iiv[3], ihiv[2], pad2_;                 // ioldiv, ioldhiv (0-based storage slots)
  StateArgs sa;                                   // the slot's tables and position
};

template <int K>
EP_DEV void dyn_moments_body(const StateArgs& a, const double* den, double* partial) {
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
    const double d = den[i];
    double t0, t1, t2;
    state_fields(a, (xx * xx + yy * yy) + zz * zz, t0, t1, t2);
    if constexpr (K == kMomP) {
      v[0] += d;
      v[1] += d * (t0 + t1 * xx * xx);
      v[2] += d * (t0 + t1 * yy * yy);
      v[3] += d * (t0 + t1 * zz * zz);
      v[4] += d * t1 * xx * yy;
      v[5] += d * t1 * xx * zz;
      v[6] += d * t1 * yy * zz;
    } else {
      DForms f;
      d_forms(xx, yy, zz, f);
      const double dv1 = t1 * d, dv2 = t2 * d;
#pragma unroll
      for (int k = 0; k < 15; k++) v[k] += f.mp[k] * dv2;
      int k = 15;
#pragma unroll
      for (int p = 0; p < 5; p++)
#pragma unroll
        for (int q = p; q < 5; q++) v[k++] += (f.d[p] * f.d[q]) * dv1;
      v[30] += t0 * d;
    }
  }
  block_sum<K>(v, red);
  if (threadIdx.x == 0) {
#pragma unroll
    for (int k = 0; k < K; k++) partial[(long)blockIdx.x * kMomD + k] = v[k];
  }
}
// Density moments of every slot whose invar evolves: partial[slot][block][31].
// The slot's state set-up at the impurity's current position (it may move).
EP_DEV StateArgs dyn_args(const InvarDyn& d, const Motion& m) {
  StateArgs a = d.sa;
  if (m.on) { a.rx = m.r[0]; a.ry = m.r[1]; a.rz = m.r[2]; }
  return a;
}
__global__ __launch_bounds__(256) void k_dyn_moments(const InvarDyn* D, const Motion* M, const int* active,
                                                     const double* den, long N, double* partial) {
  const int s = blockIdx.y;
  if (!active[s] || !D[s].on) return;
  double* p = partial + (long)s * kStateBlocks * kMomD;
  const StateArgs a = dyn_args(D[s], M[s]);
  if (D[s].state == EFFPOT_STATE_P) dyn_moments_body<kMomP>(a, den + (long)s * N, p);
  else dyn_moments_body<kMomD>(a, den + (long)s * N, p);
}
// potinvar: Hinvar = -i (H invar - Ev0 invar) from the reduced moments.
__global__ void k_dyn_hiv(InvarDyn* D, const int* active, int B, const double* partial, double dxyz) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= B || !active[s] || !D[s].on) return;
  InvarDyn& q = D[s];
  const double* p = partial + (long)s * kStateBlocks * kMomD;
  double m[kMomD];
  const int K = q.state == EFFPOT_STATE_P ? kMomP : kMomD;
  for (int k = 0; k < kMomD; k++) m[k] = 0.0;
  for (int k = 0; k < K; k++) {
    double acc = 0.0;
    for (int b = 0; b < kStateBlocks; b++) acc += p[(long)b * kMomD + k];
    m[k] = acc * dxyz;
  }
  double Hr[100], Hi[100];
  const int n = state_matrix_ri(q.state, m, q.als, Hr, Hi);
  for (int i = 0; i < n; i++) {
    double re = -q.ev0 * q.iv[2 * i], im = -q.ev0 * q.iv[2 * i + 1];
    for (int j = 0; j < n; j++) {
      re += Hr[i * n + j] * q.iv[2 * j] - Hi[i * n + j] * q.iv[2 * j + 1];
      im += Hr[i * n + j] * q.iv[2 * j + 1] + Hi[i * n + j] * q.iv[2 * j];
    }
    q.hiv[2 * i] = im;       // -i (re + i im) = im - i re
    q.hiv[2 * i + 1] = -re;
  }
}
// potenimp with the slot's current invar (grid: blocks x slots).
__global__ __launch_bounds__(256) void k_dyn_uimp(const InvarDyn* D, const Motion* M, const int* active,
                                                  double* uimp, long N) {
  const int s = blockIdx.y;
  if (!active[s] || !D[s].on) return;
  StateArgs a = dyn_args(D[s], M[s]);
  for (int k = 0; k < 20; k++) a.inv[k] = D[s].iv[k];
  double* u = uimp + (long)s * N;
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < N; i += (long)gridDim.x * blockDim.x)
    u[i] = state_uimp_at(a, i);
}
// One RK-Gill stage of invar (steprk.f90:152-167); the history indices are the
// identity after the last stage.
__global__ void k_dyn_rk(InvarDyn* D, const int* active, int B, int j, const double* dtv, double arun, double brun,
                         double crun) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= B || !active[s] || !D[s].on) return;
  InvarDyn& q = D[s];
  const double dt = dtv[s];
  for (int k = 0; k < 20; k++) {
    const double a2 = arun * (q.hiv[k] - brun * q.qiv[k]);
    q.qiv[k] = q.qiv[k] + 3.0 * a2 - crun * q.hiv[k];
    if (j == 0) {
      q.hivold[1][k] = q.hivold[0][k]; q.hivold[0][k] = q.hiv[k];
      q.ivold[2][k] = q.ivold[1][k]; q.ivold[1][k] = q.ivold[0][k]; q.ivold[0][k] = q.iv[k];
    }
    q.iv[k] = q.iv[k] + dt * a2;
  }
  if (j == 3) {
    for (int k = 0; k < 3; k++) q.iiv[k] = k;
    q.ihiv[0] = 0; q.ihiv[1] = 1;
  }
}
// Predictor + modifier of invar, then Hinvarold(ioldhiv(2)) = Hinvar and the
// ioldhiv swap (steppc.f90:84-102).
__global__ void k_dyn_pred(InvarDyn* D, const int* active, int B, const double* dtv) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= B || !active[s] || !D[s].on) return;
  InvarDyn& q = D[s];
  const double dt = dtv[s], c4o3 = 4.0 / 3.0, c112 = 112.0 / 121.0;
  for (int k = 0; k < 20; k++) {
    const double p = q.ivold[q.iiv[2]][k] +
                     c4o3 * dt * (2.0 * q.hiv[k] - q.hivold[q.ihiv[0]][k] + 2.0 * q.hivold[q.ihiv[1]][k]);
    q.ivold[q.iiv[2]][k] = q.iv[k];
    q.iv[k] = p - c112 * q.pciv[k];
    q.pciv[k] = p;
  }
  for (int k = 0; k < 20; k++) q.hivold[q.ihiv[1]][k] = q.hiv[k];
  const int t = q.ihiv[1]; q.ihiv[1] = q.ihiv[0]; q.ihiv[0] = t;
}
// Corrector of invar and the ioldiv rotation (steppc.f90:205-226); a P state
// keeps components 7-10 at zero.
__global__ void k_dyn_corr(InvarDyn* D, const int* active, int B, const double* dtv) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= B || !active[s] || !D[s].on) return;
  InvarDyn& q = D[s];
  const double dt = dtv[s], c9 = 9.0 / 121.0;
  const int nn = 2 * q.n;
  for (int k = 0; k < 20; k++) {
    const double x = 0.125 * (9.0 * q.ivold[q.iiv[2]][k] - q.ivold[q.iiv[1]][k] +
                              3.0 * dt * (q.hiv[k] + 2.0 * q.hivold[q.ihiv[0]][k] - q.hivold[q.ihiv[1]][k]));
    if (k < nn) {
      q.pciv[k] = q.pciv[k] - x;
      q.iv[k] = x + c9 * q.pciv[k];
    } else {
      q.iv[k] = 0.0;
    }
  }
  const int t = q.iiv[2]; q.iiv[2] = q.iiv[1]; q.iiv[1] = q.iiv[0]; q.iiv[0] = t;
}

}  // namespace ep
