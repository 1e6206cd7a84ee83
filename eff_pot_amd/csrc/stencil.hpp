// 13-point finite-difference stencils fused with their consumers.
//
// derivnd (init_deriv_parallel.f90:1-46 -> Deriv3D_p :368-642 / Deriv3D_pc
// :923-1197) computes one derivative along one axis per call and the callers
// combine them in separate array passes.  Here one kernel computes the needed
// derivatives along x, y and z of a field and hands them, with the field
// value, to a fused epilogue (H psi + predictor/corrector, RK stage, energy
// densities, poten side arrays).
//
// Work decomposition: a 256-thread group owns a TX x TY column tile and
// marches over a chunk of z planes.  Each thread keeps the 13 z-neighbours of
// its column in registers (one new load per plane); the centre plane is put
// in LDS with a 6-point halo in x and y for the x/y derivatives.  Boundary
// mode icon 13 reads zeros outside the grid (the redef.f:133-170 boundary
// rows), icon 8 wraps periodically.
#pragma once
#include "common.hpp"

namespace ep {

struct Grid {
  int nx, ny, nz, periodic;
  long N;
  EP_DEV long idx(int x, int y, int z) const { return (long)x + (long)nx * ((long)y + (long)ny * z); }
};

// Scaled 13-point coefficients for each axis (c_m / h^id), first and second order.
template <typename T> struct StenCoef { T d1[3][13]; T d2[3][13]; };

template <typename V> EP_DEV V gload(const V* f, int x, int y, int z, const Grid& g) {
  if (g.periodic) {
    x = (x + g.nx) % g.nx; y = (y + g.ny) % g.ny; z = (z + g.nz) % g.nz;
  } else if (x < 0 || x >= g.nx || y < 0 || y >= g.ny || z < 0 || z >= g.nz) {
    return vzero((V*)nullptr);
  }
  return f[g.idx(x, y, z)];
}

template <int TX> struct Tile {
  static constexpr int TY = kThreads / TX, LX = TX + 12, LY = TY + 12;
};

// The z-march.  epi(x, y, z, inside, v, d1[3], d2[3]) is called for every
// plane by every thread (inside=false for threads past the grid edge, which
// must still take part in barriers).  Threads past the edge still load their
// (zero or wrapped) values: they fill the LDS centre rows that inside
// threads read as halo.
template <typename T, typename V, int TX, bool D1, bool D2, class Epi>
EP_DEV void zmarch(const V* __restrict__ f, const Grid& g, const StenCoef<T>& sc, int z0, int z1, V* lds,
                   Epi&& epi) {
  constexpr int TY = Tile<TX>::TY, LX = Tile<TX>::LX, LY = Tile<TX>::LY;
  const int tx = threadIdx.x % TX, ty = threadIdx.x / TX;
  const int ntx = (g.nx + TX - 1) / TX;
  const int bx = blockIdx.x % ntx, by = blockIdx.x / ntx;
  const int x = bx * TX + tx, y = by * TY + ty;
  const bool inside = x < g.nx && y < g.ny;
  V q[13];
  const V zero = vzero((V*)nullptr);
#pragma unroll
  for (int k = 0; k < 12; k++) q[k] = gload(f, x, y, z0 - 6 + k, g);
  for (int z = z0; z < z1; z++) {
    q[12] = gload(f, x, y, z + 6, g);
    __syncthreads();
    lds[(ty + 6) * LX + tx + 6] = q[6];
    for (int e = threadIdx.x; e < LY * LX; e += kThreads) {
      const int ly = e / LX, lx = e - ly * LX;
      const bool inx = lx >= 6 && lx < TX + 6, iny = ly >= 6 && ly < TY + 6;
      if (inx == iny) continue;  // centre (own thread) or unused corner
      lds[e] = gload(f, bx * TX + lx - 6, by * TY + ly - 6, z, g);
    }
    __syncthreads();
    V d1[3], d2[3];
    if constexpr (D1) {
      V ax = zero, ay = zero, az = zero;
#pragma unroll
      for (int m = 0; m < 13; m++) {
        ax = vfma(sc.d1[0][m], lds[(ty + 6) * LX + tx + m], ax);
        ay = vfma(sc.d1[1][m], lds[(ty + m) * LX + tx + 6], ay);
        az = vfma(sc.d1[2][m], q[m], az);
      }
      d1[0] = ax; d1[1] = ay; d1[2] = az;
    }
    if constexpr (D2) {
      V ax = zero, ay = zero, az = zero;
#pragma unroll
      for (int m = 0; m < 13; m++) {
        ax = vfma(sc.d2[0][m], lds[(ty + 6) * LX + tx + m], ax);
        ay = vfma(sc.d2[1][m], lds[(ty + m) * LX + tx + 6], ay);
        az = vfma(sc.d2[2][m], q[m], az);
      }
      d2[0] = ax; d2[1] = ay; d2[2] = az;
    }
    epi(x, y, z, inside, q[6], d1, d2);
#pragma unroll
    for (int k = 0; k < 12; k++) q[k] = q[k + 1];
  }
}

// Number of blocks of the z-march for a grid: (x tiles * y tiles, z chunks, slots).
template <int TX> inline dim3 zmarch_grid(int nx, int ny, int nz, int zc, int nslots) {
  constexpr int TY = Tile<TX>::TY;
  return dim3(((nx + TX - 1) / TX) * ((ny + TY - 1) / TY), (nz + zc - 1) / zc, nslots);
}

// ---------------------------------------------------------------------------
// H psi = (lap(psi) h2o2m4 - pot4 psi) - uimp psi   (steppc.f90:53, steprk.f90:304
// with timec = 1 and the impurity term taken real for imaginary time)
template <typename T, typename V> EP_DEV V hpsi_of(V psi, const V (&d2)[3], T h2o2m4, T p4, T u) {
  V lap = (d2[0] + d2[1]) + d2[2];
  V a = h2o2m4 * lap - p4 * psi;
  return a - u * psi;
}
template <typename T> EP_DEV T norm2(cx<T> v) { return v.x * v.x + v.y * v.y; }
template <typename T> EP_DEV T norm2(T v) { return v * v; }
template <typename T> EP_DEV T dotre(cx<T> a, cx<T> b) { return a.x * b.x + a.y * b.y; }
template <typename T> EP_DEV T dotre(T a, T b) { return a * b; }

// V = cx<T> (complex psi) or T (real psi: imaginary time from a real state
// keeps the imaginary part exactly zero, so the real path is the same
// arithmetic on half the bytes).
template <typename T, typename V> struct StepArgs {
  Grid g;
  StenCoef<T> sc;
  int zc;
  T h2o2m4, dt;
  const int* active;
  long slot;  // elements per slot (N)
  const V* cur;   // psi read through the stencil
  V* out;         // psi written (never == cur)
  V* o1; V* o2; V* o3;  // psi history psiold(ioldp(1..3))
  V* h1; V* h2;         // hpsi history hpsiold(ioldh(1..2))
  V* pc; V* q;          // slot stride of pc/q is `pcq_slot` elements of V
  long pcq_slot;
  const T* pot4; const T* uimp;
  T* den;
  double* partial;  // [slot][block][2]: sum c9|pc| (or 0), sum |psi_new|^2
  int nblk;
  // RK-Gill stage coefficients
  T arun, brun, crun;
  int first_stage;
};

// Hamming predictor + modifier (steppc.f90:53-64).  out = o3's buffer.
template <typename T, typename V, int TX>
EP_DEV void k_pc_predict_body(const StepArgs<T, V>& a) {
  const int s = blockIdx.z;
  if (!a.active[s]) return;
  __shared__ V lds[Tile<TX>::LX * Tile<TX>::LY];
  const long so = (long)s * a.slot;
  const V* cur = a.cur + so;
  V* o3 = a.o3 + so; V* h1 = a.h1 + so; V* h2 = a.h2 + so; V* pc = a.pc + (long)s * a.pcq_slot;
  const T* pot4 = a.pot4 + so; const T* uimp = a.uimp + so; T* den = a.den + so;
  const T c4o3dt = T(4.0 / 3.0) * a.dt, c112 = T(112.0 / 121.0);
  const int z0 = blockIdx.y * a.zc, z1 = min(z0 + a.zc, a.g.nz);
  zmarch<T, V, TX, false, true>(cur, a.g, a.sc, z0, z1, lds,
    [&](int x, int y, int z, bool in, V psi, V(&)[3], V(&d2)[3]) {
      if (!in) return;
      const long i = a.g.idx(x, y, z);
      V hp = hpsi_of(psi, d2, a.h2o2m4, pot4[i], uimp[i]);
      V pred = o3[i] + c4o3dt * ((T(2) * hp - h1[i]) + T(2) * h2[i]);
      V pcv = pc[i];
      V pn = pred - c112 * pcv;
      o3[i] = pn;  // modified psi goes to the retired psi_{n-3} buffer
      pc[i] = pred;
      h2[i] = hp;
      den[i] = norm2(pn);
    });
}

// Hamming corrector + final value (steppc.f90:135-151), errHe and norm partials.
// Here o3 holds psi_n (psiold(ioldp(3)) after the modifier), o2 psi_{n-2},
// h1 = H psi_n, h2 = H psi_{n-1} (after the ioldh swap).
template <typename T, typename V, int TX>
EP_DEV void k_pc_correct_body(const StepArgs<T, V>& a) {
  const int s = blockIdx.z;
  if (!a.active[s]) return;
  __shared__ V lds[Tile<TX>::LX * Tile<TX>::LY];
  __shared__ double red[8];
  const long so = (long)s * a.slot;
  const V* cur = a.cur + so;
  V* out = a.out + so;
  const V* o3 = a.o3 + so; const V* o2 = a.o2 + so;
  const V* h1 = a.h1 + so; const V* h2 = a.h2 + so; V* pc = a.pc + (long)s * a.pcq_slot;
  const T* pot4 = a.pot4 + so; const T* uimp = a.uimp + so; T* den = a.den + so;
  const T c9 = T(9.0 / 121.0), dt3 = T(3) * a.dt;
  double v[2] = {0.0, 0.0};
  const int z0 = blockIdx.y * a.zc, z1 = min(z0 + a.zc, a.g.nz);
  zmarch<T, V, TX, false, true>(cur, a.g, a.sc, z0, z1, lds,
    [&](int x, int y, int z, bool in, V psi, V(&)[3], V(&d2)[3]) {
      if (!in) return;
      const long i = a.g.idx(x, y, z);
      V hp = hpsi_of(psi, d2, a.h2o2m4, pot4[i], uimp[i]);
      V corr = T(0.125) * ((T(9) * o3[i] - o2[i]) + dt3 * ((hp + T(2) * h1[i]) - h2[i]));
      V pcv = pc[i] - corr;
      V pn = corr + c9 * pcv;
      pc[i] = pcv;
      out[i] = pn;
      T d = norm2(pn);
      den[i] = d;
      v[0] += (double)(c9 * sqrt(norm2(pcv)));
      v[1] += (double)d;
    });
  block_sum<2>(v, red);
  if (threadIdx.x == 0) {
    const long b = (long)blockIdx.y * gridDim.x + blockIdx.x;
    double* p = a.partial + ((long)s * a.nblk + b) * 2;
    p[0] = v[0]; p[1] = v[1];
  }
}

// One Runge-Kutta-Gill stage (steprk.f90:301-316).  At the first stage the host
// has rotated the history so that o1 := cur (psi at the start) and h1 is the
// retired buffer that receives H psi.
template <typename T, typename V, int TX>
EP_DEV void k_rk_stage_body(const StepArgs<T, V>& a) {
  const int s = blockIdx.z;
  if (!a.active[s]) return;
  __shared__ V lds[Tile<TX>::LX * Tile<TX>::LY];
  __shared__ double red[8];
  const long so = (long)s * a.slot;
  const V* cur = a.cur + so;
  V* out = a.out + so; V* h1 = a.h1 + so; V* q = a.q + (long)s * a.pcq_slot;
  const T* pot4 = a.pot4 + so; const T* uimp = a.uimp + so; T* den = a.den + so;
  double v[2] = {0.0, 0.0};
  const int z0 = blockIdx.y * a.zc, z1 = min(z0 + a.zc, a.g.nz);
  zmarch<T, V, TX, false, true>(cur, a.g, a.sc, z0, z1, lds,
    [&](int x, int y, int z, bool in, V psi, V(&)[3], V(&d2)[3]) {
      if (!in) return;
      const long i = a.g.idx(x, y, z);
      V hp = hpsi_of(psi, d2, a.h2o2m4, pot4[i], uimp[i]);
      V qv = q[i];
      V a2 = a.arun * (hp - a.brun * qv);
      q[i] = (qv + T(3) * a2) - a.crun * hp;
      if (a.first_stage) h1[i] = hp;
      V pn = psi + a.dt * a2;
      out[i] = pn;
      T d = norm2(pn);
      den[i] = d;
      v[1] += (double)d;
    });
  block_sum<2>(v, red);
  if (threadIdx.x == 0) {
    const long b = (long)blockIdx.y * gridDim.x + blockIdx.x;
    double* p = a.partial + ((long)s * a.nblk + b) * 2;
    p[0] = 0.0; p[1] = v[1];
  }
}

// ---------------------------------------------------------------------------
// poten side arrays: derden (derden.f90:16-18) fused with the FFT sources of
// the correlation convolution (poten.f90:62-64) and, for OTC, of the alpha_s
// integrals (term_alfa.f90:157-165).
template <typename T> struct PotenArgs {
  Grid g;
  StenCoef<T> sc;
  int zc;
  const int* active;
  long slot;
  int otc;
  T cp4, cpp4, den0s;
  const T* den; const T* dencg; const T* denalf;
  T* dx; T* dy; T* dz;
  T* s0; T* s1; T* s2; T* s3;  // sources: correlation, falfs*d{x,y,z}den
};
template <typename T, int TX>
__global__ __launch_bounds__(kThreads) void k_grad(PotenArgs<T> a) {
  const int s = blockIdx.z;
  if (!a.active[s]) return;
  __shared__ T lds[Tile<TX>::LX * Tile<TX>::LY];
  const long so = (long)s * a.slot;
  const T* den = a.den + so; const T* dencg = a.dencg + so; const T* denalf = a.denalf + so;
  T *dx = a.dx + so, *dy = a.dy + so, *dz = a.dz + so;
  T *s0 = a.s0 + so, *s1 = a.s1 + so, *s2 = a.s2 + so, *s3 = a.s3 + so;
  const int z0 = blockIdx.y * a.zc, z1 = min(z0 + a.zc, a.g.nz);
  zmarch<T, T, TX, true, false>(den, a.g, a.sc, z0, z1, lds,
    [&](int x, int y, int z, bool in, T d, T(&d1)[3], T(&)[3]) {
      if (!in) return;
      const long i = a.g.idx(x, y, z);
      dx[i] = d1[0]; dy[i] = d1[1]; dz[i] = d1[2];
      T cg = dencg[i];
      s0[i] = d * cg * (a.cp4 + a.cpp4 * cg);
      if (a.otc) {
        T fa = denalf[i];  // holds falfs
        s1[i] = fa * d1[0]; s2[i] = fa * d1[1]; s3[i] = fa * d1[2];
      }
    });
}

// alpha_s field pieces (term_alfa.f90:194-198, 216-225): the FFT source
// sum_i d_i den * int_i, and div = sum_i d/dx_i (falfs int_i), each axis
// differentiating its own component.
template <typename T> struct DivArgs {
  Grid g;
  StenCoef<T> sc;
  int zc;
  const int* active;
  long slot;
  T den0s;
  const T* denalf; const T* ix; const T* iy; const T* iz;
  const T* dx; const T* dy; const T* dz;
  T* src; T* div;
};
template <typename T, int TX>
__global__ __launch_bounds__(kThreads) void k_div(DivArgs<T> a) {
  constexpr int TY = Tile<TX>::TY, LX = Tile<TX>::LX, LY = Tile<TX>::LY;
  const int s = blockIdx.z;
  if (!a.active[s]) return;
  __shared__ T lx_[TY * LX];  // falfs*intx with x halo
  __shared__ T ly_[LY * TX];  // falfs*inty with y halo
  const long so = (long)s * a.slot;
  const Grid& g = a.g;
  const T* denalf = a.denalf + so;
  const T *ix = a.ix + so, *iy = a.iy + so, *iz = a.iz + so;
  const T *dx = a.dx + so, *dy = a.dy + so, *dz = a.dz + so;
  T *src = a.src + so, *div = a.div + so;
  auto gfa = [&](const T* comp, int x, int y, int z) -> T {
    if (g.periodic) {
      x = (x + g.nx) % g.nx; y = (y + g.ny) % g.ny; z = (z + g.nz) % g.nz;
    } else if (x < 0 || x >= g.nx || y < 0 || y >= g.ny || z < 0 || z >= g.nz) {
      return T(0);
    }
    long i = g.idx(x, y, z);
    return denalf[i] * comp[i];  // denalf holds falfs
  };
  const int tx = threadIdx.x % TX, ty = threadIdx.x / TX;
  const int ntx = (g.nx + TX - 1) / TX;
  const int bx = blockIdx.x % ntx, by = blockIdx.x / ntx;
  const int x = bx * TX + tx, y = by * TY + ty;
  const bool inside = x < g.nx && y < g.ny;
  const int z0 = blockIdx.y * a.zc, z1 = min(z0 + a.zc, g.nz);
  T q[13];
#pragma unroll
  for (int k = 0; k < 12; k++) q[k] = gfa(iz, x, y, z0 - 6 + k);
  for (int z = z0; z < z1; z++) {
    q[12] = gfa(iz, x, y, z + 6);
    __syncthreads();
    for (int e = threadIdx.x; e < TY * LX; e += kThreads) {
      const int r = e / LX, c = e - r * LX;
      lx_[e] = gfa(ix, bx * TX + c - 6, by * TY + r, z);
    }
    for (int e = threadIdx.x; e < LY * TX; e += kThreads) {
      const int r = e / TX, c = e - r * TX;
      ly_[e] = gfa(iy, bx * TX + c, by * TY + r - 6, z);
    }
    __syncthreads();
    if (inside) {
      T ax = T(0), ay = T(0), az = T(0);
#pragma unroll
      for (int m = 0; m < 13; m++) {
        ax += a.sc.d1[0][m] * lx_[ty * LX + tx + m];
        ay += a.sc.d1[1][m] * ly_[(ty + m) * TX + tx];
        az += a.sc.d1[2][m] * q[m];
      }
      const long i = g.idx(x, y, z);
      div[i] = (ax + ay) + az;
      src[i] = (dx[i] * ix[i] + dy[i] * iy[i]) + dz[i] * iz[i];
    }
#pragma unroll
    for (int k = 0; k < 12; k++) q[k] = q[k + 1];
  }
}

// ---------------------------------------------------------------------------
// energy() densities (energy.f90:47-165), the chemical potential and the
// helium centre of mass (r_cm.f90:38-54): 11 block partial sums.
constexpr int kEnergySums = 11;
template <typename T> struct EnergyArgs {
  Grid g;
  StenCoef<T> sc;
  int zc;
  const int* active;
  long slot;
  int otc;
  T h2o2m4, aux1, aux2, beta, den_m, den0s;
  double x0, y0, z0, hx, hy, hz;
  const cx<T>* psi_all;  // [slot][5][N]; the live buffer is ring[8*slot]
  const int* ring;
  const T *den, *dencg, *delj4, *denalf, *dx, *dy, *dz, *ix, *iy, *iz, *uimp, *pot4;
  double* partial;
  int nblk;
};
template <typename T, typename V, int TX>
__global__ __launch_bounds__(kThreads) void k_energy(EnergyArgs<T> a) {
  const int s = blockIdx.z;
  if (!a.active[s]) return;
  __shared__ V lds[Tile<TX>::LX * Tile<TX>::LY];
  __shared__ double red[4 * kEnergySums];
  const long so = (long)s * a.slot;
  const V* psi = reinterpret_cast<const V*>(a.psi_all + ((long)s * 5 + a.ring[8 * s]) * a.slot);
  double v[kEnergySums];
#pragma unroll
  for (int k = 0; k < kEnergySums; k++) v[k] = 0.0;
  const int z0 = blockIdx.y * a.zc, z1 = min(z0 + a.zc, a.g.nz);
  zmarch<T, V, TX, true, true>(psi, a.g, a.sc, z0, z1, lds,
    [&](int x, int y, int z, bool in, V p, V(&d1)[3], V(&d2)[3]) {
      if (!in) return;
      const long i = so + a.g.idx(x, y, z);
      const T d = a.den[i], cg = a.dencg[i];
      T ek = (norm2(d1[0]) + norm2(d1[1])) + norm2(d1[2]);
      v[0] += (double)ek;
      v[1] += (double)(a.delj4[i] * d);
      v[2] += (double)(d * cg * cg * (a.aux1 + a.aux2 * cg));
      if (a.otc) {
        T fa = a.denalf[i];  // holds falfs
        v[3] += (double)(fa * ((a.dx[i] * a.ix[i] + a.dy[i] * a.iy[i]) + a.dz[i] * a.iz[i]));
      }
      v[4] += (double)(d * (T(1) + tanh(a.beta * (d - a.den_m))));
      v[5] += (double)(d * a.uimp[i]);
      v[6] += (double)d;
      v[7] += (double)d * (a.x0 + a.hx * x);
      v[8] += (double)d * (a.y0 + a.hy * y);
      v[9] += (double)d * (a.z0 + a.hz * z);
      V hp = hpsi_of(p, d2, a.h2o2m4, a.pot4[i], a.uimp[i]);
      v[10] -= (double)dotre(p, hp);
    });
  block_sum<kEnergySums>(v, red);
  if (threadIdx.x == 0) {
    const long b = (long)blockIdx.y * gridDim.x + blockIdx.x;
    double* pp = a.partial + ((long)s * a.nblk + b) * kEnergySums;
#pragma unroll
    for (int k = 0; k < kEnergySums; k++) pp[k] = v[k];
  }
}

}  // namespace ep
