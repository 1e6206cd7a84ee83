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
  int32_t on, pad_;
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
  if (!active[s] || !M[s].on) return;
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
__global__ __launch_bounds__(256) void k_force_partials(const Motion* M, const int* active, const double* dx,
                                                        const double* dy, const double* dz, const double* uimp, long N,
                                                        double* partial) {
  const int s = blockIdx.y;
  if (!active[s] || !M[s].on) return;
  __shared__ double red[4 * 3];
  double v[3] = {0.0, 0.0, 0.0};
  const long so = (long)s * N;
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < N; i += (long)gridDim.x * blockDim.x) {
    const double u = uimp[so + i];
    v[0] += dx[so + i] * u;
    v[1] += dy[so + i] * u;
    v[2] += dz[so + i] * u;
  }
  block_sum<3>(v, red);
  if (threadIdx.x == 0) {
    double* p = partial + ((long)s * gridDim.x + blockIdx.x) * 3;
    p[0] = v[0]; p[1] = v[1]; p[2] = v[2];
  }
}

}  // namespace ep
