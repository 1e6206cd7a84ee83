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
//
// The march is software-pipelined: while plane z is differentiated and its
// epilogue runs, the loads of plane z+1 (the new z-neighbour, the halo and the
// pointwise operands the epilogue will need) are already in flight, and the
// LDS plane is double-buffered so one barrier per plane suffices.
#pragma once
#include "common.hpp"
#include "plus.hpp"

namespace ep {

struct Grid {
  int nx, ny, nz, periodic;
  long N;
  EP_DEV long idx(int x, int y, int z) const { return (long)x + (long)nx * ((long)y + (long)ny * z); }
  // the same offset as 32 bits (a field of one slot holds fewer than 2^32 values)
  EP_DEV unsigned at(int x, int y, int z) const {
    return (unsigned)x + (unsigned)nx * ((unsigned)y + (unsigned)ny * (unsigned)z);
  }
};

// Scaled 13-point coefficients for each axis (c_m / h^id), first and second
// order, indexed by m = 0..12 (offset m - 6).  The first-derivative stencil is
// odd and the second even in the offset, exactly, so only offsets 0..6 are
// kept (h1, h2): half the kernel-argument words, which the step kernels hold
// in scalar registers, and the same values as the full tables.
template <typename T> struct StenCoef {
  T h1[3][7], h2[3][7];  // [axis][offset 0..6]
  EP_DEV T d1(int a, int m) const { return m >= 6 ? h1[a][m - 6] : -h1[a][6 - m]; }
  EP_DEV T d2(int a, int m) const { return h2[a][m >= 6 ? m - 6 : 6 - m]; }
};

// Offset of column (x, y) within a plane, wrapped for icon 8; -1 outside the
// grid for icon 13.
EP_DEV int colofs(int x, int y, const Grid& g) {
  if (g.periodic) {
    x = (x + g.nx) % g.nx; y = (y + g.ny) % g.ny;
  } else if (x < 0 || x >= g.nx || y < 0 || y >= g.ny) {
    return -1;
  }
  return x + g.nx * y;
}
// Offset of plane z from a field's base, wrapped for icon 8; -1 when the
// plane lies outside the grid for icon 13 (its values are zero).  It is
// uniform over the workgroup: the z-marches form it once per plane and share
// it between the loads of that plane (formed inside every load, the wrap and
// the 64-bit product sat under each load's lane branch and were repeated per
// load: some 30 scalar instructions each).
EP_DEV int zofs(int z, const Grid& g) {
  if (z < 0 || z >= g.nz) {
    if (!g.periodic) return -1;
    z = (z + g.nz) % g.nz;
  }
  return g.nx * g.ny * z;
}
// f at column c (colofs) of the plane at offset zo (zofs); zero outside the grid.
template <typename V> EP_DEV V ldz(const V* f, int c, int zo) {
  return (c < 0 || zo < 0) ? vzero((V*)nullptr) : ld32(f, (unsigned)(c + zo));
}

template <int TX> struct Tile {
  static constexpr int TY = kThreads / TX, LX = TX + 12, LY = TY + 12, PL = LX * LY;
  static constexpr int H = 12 * (TX + TY);              // halo points of a plane (no corners)
  static constexpr int NH = (H + kThreads - 1) / kThreads;  // halo loads per thread
};

// Workgroup -> tile order.  Workgroup i runs on XCD i % 8 and each XCD has its
// own L2; giving every XCD a contiguous range of tiles keeps the neighbours
// whose planes a tile reads as halo in the same L2.
EP_DEV int xcd_order(int i, int G) { return (G & 7) ? i : (i & 7) * (G >> 3) + (i >> 3); }

// Column tile and z chunk of this workgroup (blockIdx.z is the slot).
template <int TX> struct TilePos {
  int bx, by, z0, z1, tx, ty, x, y;
  bool inside;
  EP_DEV TilePos(const Grid& g, int zc) {
    constexpr int TY = Tile<TX>::TY;
    const int L = xcd_order(blockIdx.x + gridDim.x * blockIdx.y, gridDim.x * gridDim.y);
    const int tile = L % gridDim.x, chunk = L / gridDim.x;
    const int ntx = (g.nx + TX - 1) / TX;
    bx = tile % ntx; by = tile / ntx;
    z0 = chunk * zc; z1 = min(z0 + zc, g.nz);
    tx = threadIdx.x % TX; ty = threadIdx.x / TX;
    x = bx * TX + tx; y = by * TY + ty;
    inside = x < g.nx && y < g.ny;
  }
};

// Halo point h of a tile plane -> LDS offset (lo) and global column (go).
// Points 0 .. 12 TY-1 are the x strips (6 columns left and right of the
// tile), the rest the y strips (6 rows below and above).
template <int TX> EP_DEV void halo_point(int h, int bx, int by, const Grid& g, int& lo, int& go) {
  constexpr int TY = Tile<TX>::TY, LX = Tile<TX>::LX;
  if (h >= Tile<TX>::H) { lo = -1; go = -1; return; }
  int lx, ly;
  if (h < 12 * TY) {
    const int r = h / 12, c = h - 12 * (h / 12);
    ly = 6 + r; lx = c < 6 ? c : TX + c;
  } else {
    const int k = h - 12 * TY, r = k / TX, c = k - TX * (k / TX);
    ly = r < 6 ? r : TY + r; lx = 6 + c;
  }
  lo = ly * LX + lx;
  go = colofs(bx * TX + lx - 6, by * TY + ly - 6, g);
}

struct NoPre {};

// The z-march.  pre(i) loads the pointwise operands of grid point i (called
// for inside threads only, one plane ahead); epi(x, y, z, inside, v, d1, d2,
// pre_value) is called for every plane by every thread (inside=false for
// threads past the grid edge).  Threads past the edge still load their (zero
// or wrapped) column: it is LDS halo of inside threads.  lds holds 2 planes.
template <typename T, typename V, int TX, bool D1, bool D2, class Pre, class Epi>
EP_DEV void zmarch(const V* __restrict__ f, const Grid& g, const StenCoef<T>& sc, const TilePos<TX>& tp, V* lds,
                   Pre&& pre, Epi&& epi) {
  constexpr int LX = Tile<TX>::LX, PL = Tile<TX>::PL, NH = Tile<TX>::NH;
  const int tx = tp.tx, ty = tp.ty, z0 = tp.z0, z1 = tp.z1;
  if (z0 >= z1) return;  // uniform over the group
  const int col = colofs(tp.x, tp.y, g);
  const unsigned plane = (unsigned)(g.nx * g.ny);
  int lo[NH], go[NH];
#pragma unroll
  for (int k = 0; k < NH; k++) halo_point<TX>(threadIdx.x + k * kThreads, tp.bx, tp.by, g, lo[k], go[k]);
  V q[13], hv[NH];
  const V zero = vzero((V*)nullptr);
#pragma unroll
  for (int k = 0; k < 13; k++) q[k] = ldz(f, col, zofs(z0 - 6 + k, g));
  const int zo0 = zofs(z0, g);
#pragma unroll
  for (int k = 0; k < NH; k++) hv[k] = ldz(f, go[k], zo0);
  using PV = decltype(pre(0u));
  PV p0{};
  if (tp.inside) p0 = pre((unsigned)col + plane * (unsigned)z0);
  lds[(ty + 6) * LX + tx + 6] = q[6];
#pragma unroll
  for (int k = 0; k < NH; k++)
    if (lo[k] >= 0) lds[lo[k]] = hv[k];
  __syncthreads();
  int buf = 0;
  for (int z = z0; z < z1; z++) {
    const bool more = z + 1 < z1;
    V qn = zero;
    PV p1{};
    if (more) {
      qn = ldz(f, col, zofs(z + 7, g));
      const int zo = zofs(z + 1, g);
#pragma unroll
      for (int k = 0; k < NH; k++) hv[k] = ldz(f, go[k], zo);
      if (tp.inside) p1 = pre((unsigned)col + plane * (unsigned)(z + 1));
    }
    const V* pl = lds + buf * PL;
    V d1[3], d2[3];
    if constexpr (D1) {
      V ax = zero, ay = zero, az = zero;
#pragma unroll
      for (int m = 0; m < 13; m++) {
        ax = vfma(sc.d1(0, m), pl[(ty + 6) * LX + tx + m], ax);
        ay = vfma(sc.d1(1, m), pl[(ty + m) * LX + tx + 6], ay);
        az = vfma(sc.d1(2, m), q[m], az);
      }
      d1[0] = ax; d1[1] = ay; d1[2] = az;
    }
    if constexpr (D2) {
      V ax = zero, ay = zero, az = zero;
#pragma unroll
      for (int m = 0; m < 13; m++) {
        ax = vfma(sc.d2(0, m), pl[(ty + 6) * LX + tx + m], ax);
        ay = vfma(sc.d2(1, m), pl[(ty + m) * LX + tx + 6], ay);
        az = vfma(sc.d2(2, m), q[m], az);
      }
      d2[0] = ax; d2[1] = ay; d2[2] = az;
    }
    epi(tp.x, tp.y, z, tp.inside, q[6], d1, d2, p0);
    if (more) {
      V* nl = lds + (buf ^ 1) * PL;
      nl[(ty + 6) * LX + tx + 6] = q[7];
#pragma unroll
      for (int k = 0; k < NH; k++)
        if (lo[k] >= 0) nl[lo[k]] = hv[k];
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < 12; k++) q[k] = q[k + 1];
    q[12] = qn;
    p0 = p1;
    buf ^= 1;
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
// Real time (timec = i, steppc.f90:53): the step integrates i * H psi.
template <typename T> EP_DEV cx<T> times_i(cx<T> v) { return cx<T>{-v.y, v.x}; }
template <typename T> EP_DEV T times_i(T v) { return v; }  // never used: real time needs complex psi
template <typename T> EP_DEV T norm2(cx<T> v) { return v.x * v.x + v.y * v.y; }
template <typename T> EP_DEV T norm2(T v) { return v * v; }
template <typename T> EP_DEV T dotre(cx<T> a, cx<T> b) { return a.x * b.x + a.y * b.y; }
template <typename T> EP_DEV T dotre(T a, T b) { return a * b; }

// V = cx<T> (complex psi) or T (real psi: imaginary time from a real state
// keeps the imaginary part exactly zero, so the real path is the same
// arithmetic on half the bytes, up to the compiler's FMA contraction).
// P: precision of the fields the step reads/writes (pot4, uimp, den): T, or
// float when a FP64 state is driven by FP32 mean fields.
template <typename T, typename V, typename P> struct StepArgs {
  Grid g;
  StenCoef<T> sc;
  int zc;
  T h2o2m4;
  const double* dtv;  // time step of each slot
  const int* active;
  long slot;  // elements per slot (N)
  const V* cur;   // psi read through the stencil
  V* out;         // psi written (never == cur)
  V* o1; V* o2; V* o3;  // psi history psiold(ioldp(1..3))
  V* h1; V* h2;         // hpsi history hpsiold(ioldh(1..2))
  V* pc; V* q;          // slot stride of pc/q is `pcq_slot` elements of V
  long pcq_slot;
  const P* pot4; const P* uimp;
  P* den;
  double* partial;  // [slot][block][2]: sum c9|pc| (or 0), sum |psi_new|^2
  int nblk;
  // RK-Gill stage coefficients
  T arun, brun, crun;
  int first_stage;
  int write_den;  // corrector: also store the density (no renormalisation follows)
  int rt;         // real time: H psi -> i H psi (complex psi only)
};

// Hamming predictor + modifier (steppc.f90:53-64).  out = o3's buffer.  The
// block partial sums of the modified density feed the scaling of that density
// to n4 atoms before the step's poten.
template <typename T, typename V, typename P, int TX>
EP_DEV void k_pc_predict_body(const StepArgs<T, V, P>& a) {
  const int s = blockIdx.z;
  if (!a.active[s]) return;
  __shared__ V lds[2 * Tile<TX>::PL];
  __shared__ double red[8];
  double v[2] = {0.0, 0.0};
  const long so = (long)s * a.slot;
  const V* cur = a.cur + so;
  V* o3 = a.o3 + so; const V* h1 = a.h1 + so; V* h2 = a.h2 + so; V* pc = a.pc + (long)s * a.pcq_slot;
  const P* pot4 = a.pot4 + so; const P* uimp = a.uimp + so; P* den = a.den + so;
  const T c4o3dt = T(4.0 / 3.0) * (T)a.dtv[s], c112 = T(112.0 / 121.0);
  struct Op { V o3, h1, h2, pc; T p4, u; };
  zmarch<T, V, TX, false, true>(cur, a.g, a.sc, TilePos<TX>(a.g, a.zc), lds,
    [&](unsigned i) {
      return Op{ld32(o3, i), ld32(h1, i), ld32(h2, i), ld32(pc, i), (T)ld32(pot4, i), (T)ld32(uimp, i)};
    },
    [&](int x, int y, int z, bool in, V psi, V(&)[3], V(&d2)[3], const Op& o) {
      if (!in) return;
      const unsigned i = a.g.at(x, y, z);
      V hp = hpsi_of(psi, d2, a.h2o2m4, o.p4, o.u);
      if (a.rt) hp = times_i(hp);
      V pred = o.o3 + c4o3dt * ((T(2) * hp - o.h1) + T(2) * o.h2);
      V pn = pred - c112 * o.pc;
      st32(o3, i, pn);  // modified psi goes to the retired psi_{n-3} buffer
      st32(pc, i, pred);
      st32(h2, i, hp);
      const T d = norm2(pn);
      st32(den, i, (P)d);
      v[1] += (double)d;
    });
  block_sum<2>(v, red);
  if (threadIdx.x == 0) {
    const long b = (long)blockIdx.y * gridDim.x + blockIdx.x;
    double* p = a.partial + ((long)s * a.nblk + b) * 2;
    p[0] = 0.0; p[1] = v[1];
  }
}

// Hamming corrector + final value (steppc.f90:135-151), errHe and norm partials.
// Here o3 holds psi_n (psiold(ioldp(3)) after the modifier), o2 psi_{n-2},
// h1 = H psi_n, h2 = H psi_{n-1} (after the ioldh swap).  The density is
// written only when no renormalisation (which rewrites it) follows.
template <typename T, typename V, typename P, int TX>
EP_DEV void k_pc_correct_body(const StepArgs<T, V, P>& a) {
  const int s = blockIdx.z;
  if (!a.active[s]) return;
  __shared__ V lds[2 * Tile<TX>::PL];
  __shared__ double red[8];
  const long so = (long)s * a.slot;
  const V* cur = a.cur + so;
  V* out = a.out + so;
  const V* o3 = a.o3 + so; const V* o2 = a.o2 + so;
  const V* h1 = a.h1 + so; const V* h2 = a.h2 + so; V* pc = a.pc + (long)s * a.pcq_slot;
  const P* pot4 = a.pot4 + so; const P* uimp = a.uimp + so; P* den = a.den + so;
  const T c9 = T(9.0 / 121.0), dt3 = T(3) * (T)a.dtv[s];
  const bool wden = a.write_den;
  double v[2] = {0.0, 0.0};
  struct Op { V o3, o2, h1, h2, pc; T p4, u; };
  const TilePos<TX> tp(a.g, a.zc);
  zmarch<T, V, TX, false, true>(cur, a.g, a.sc, tp, lds,
    [&](unsigned i) {
      return Op{ld32(o3, i), ld32(o2, i), ld32(h1, i), ld32(h2, i), ld32(pc, i), (T)ld32(pot4, i), (T)ld32(uimp, i)};
    },
    [&](int x, int y, int z, bool in, V psi, V(&)[3], V(&d2)[3], const Op& o) {
      if (!in) return;
      const unsigned i = a.g.at(x, y, z);
      V hp = hpsi_of(psi, d2, a.h2o2m4, o.p4, o.u);
      if (a.rt) hp = times_i(hp);
      V corr = T(0.125) * ((T(9) * o.o3 - o.o2) + dt3 * ((hp + T(2) * o.h1) - o.h2));
      V pcv = o.pc - corr;
      V pn = corr + c9 * pcv;
      st32(pc, i, pcv);
      st32(out, i, pn);
      T d = norm2(pn);
      if (wden) st32(den, i, (P)d);
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
template <typename T, typename V, typename P, int TX>
EP_DEV void k_rk_stage_body(const StepArgs<T, V, P>& a) {
  const int s = blockIdx.z;
  if (!a.active[s]) return;
  __shared__ V lds[2 * Tile<TX>::PL];
  __shared__ double red[8];
  const long so = (long)s * a.slot;
  const V* cur = a.cur + so;
  V* out = a.out + so; V* h1 = a.h1 + so; V* q = a.q + (long)s * a.pcq_slot;
  const T dt = (T)a.dtv[s];
  const P* pot4 = a.pot4 + so; const P* uimp = a.uimp + so; P* den = a.den + so;
  double v[2] = {0.0, 0.0};
  struct Op { V q; T p4, u; };
  zmarch<T, V, TX, false, true>(cur, a.g, a.sc, TilePos<TX>(a.g, a.zc), lds,
    [&](unsigned i) { return Op{ld32(q, i), (T)ld32(pot4, i), (T)ld32(uimp, i)}; },
    [&](int x, int y, int z, bool in, V psi, V(&)[3], V(&d2)[3], const Op& o) {
      if (!in) return;
      const unsigned i = a.g.at(x, y, z);
      V hp = hpsi_of(psi, d2, a.h2o2m4, o.p4, o.u);
      if (a.rt) hp = times_i(hp);
      V a2 = a.arun * (hp - a.brun * o.q);
      st32(q, i, (o.q + T(3) * a2) - a.crun * hp);
      if (a.first_stage) st32(h1, i, hp);
      V pn = psi + dt * a2;
      st32(out, i, pn);
      T d = norm2(pn);
      st32(den, i, (P)d);
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
  T* s0; T* s1; T* s2; T* s3;  // sources: correlation, falfs*d{x,y,z}den (s1..s3 null: not stored)
  const double* dscale;        // [slot] factor den is read with (the step's density to n4 atoms), or null
};
template <typename T, int TX>
__global__ __launch_bounds__(kThreads) void k_grad(PotenArgs<T> a) {
  const int s = blockIdx.z;
  if (!a.active[s]) return;
  __shared__ T lds[2 * Tile<TX>::PL];
  const long so = (long)s * a.slot;
  const T* den = a.den + so; const T* dencg = a.dencg + so; const T* denalf = a.denalf + so;
  T *dx = a.dx ? a.dx + so : nullptr, *dy = a.dy ? a.dy + so : nullptr, *dz = a.dz ? a.dz + so : nullptr;
  T *s0 = a.s0 + so, *s1 = a.s1 + so, *s2 = a.s2 + so, *s3 = a.s3 + so;
  const bool otc = a.otc;
  const T sc = a.dscale ? (T)a.dscale[s] : T(1);  // the stencil is linear: scale its results
  struct Op { T cg, fa; };
  zmarch<T, T, TX, true, false>(den, a.g, a.sc, TilePos<TX>(a.g, a.zc), lds,
    [&](unsigned i) { return Op{ld32(dencg, i), otc ? ld32(denalf, i) : T(0)}; },
    [&](int x, int y, int z, bool in, T d, T(&d1)[3], T(&)[3], const Op& o) {
      if (!in) return;
      const unsigned i = a.g.at(x, y, z);
      const T gx = sc * d1[0], gy = sc * d1[1], gz = sc * d1[2];
      if (dx) { st32(dx, i, gx); st32(dy, i, gy); st32(dz, i, gz); }  // null: k_div forms them (GRAD)
      st32(s0, i, (sc * d) * o.cg * (a.cp4 + a.cpp4 * o.cg));
      if (otc && s1) {  // o.fa holds falfs; with s1 null the forward transform forms these products
        st32(s1, i, o.fa * gx); st32(s2, i, o.fa * gy); st32(s3, i, o.fa * gz);
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
  const T* den; const double* dscale;  // GRAD: the density and its [slot] factor (or null)
  T* src; T* div;
};
// GRAD: the density gradient of the source is formed here from den (k_grad's
// 13-point sums in k_grad's order; equal to FP32 rounding, as the compiler
// contracts the two kernels' sums into FMAs its own way) instead of read from
// dx, dy, dz, which k_grad then does not store: k_grad is bound by its stores
// (7 fields written; 4 without dx, dy, dz).
template <typename T, int TX, bool GRAD = false>
__global__ __launch_bounds__(kThreads) void k_div(DivArgs<T> a) {
  constexpr int TY = Tile<TX>::TY, LX = Tile<TX>::LX, LY = Tile<TX>::LY;
  constexpr int NXS = TY * LX, NYS = LY * TX;  // x strip (x halo), y strip (y halo) of a plane
  constexpr int KX = (NXS + kThreads - 1) / kThreads, KY = (NYS + kThreads - 1) / kThreads;
  const int s = blockIdx.z;
  if (!a.active[s]) return;
  __shared__ T lx_[2 * NXS];  // falfs*intx
  __shared__ T ly_[2 * NYS];  // falfs*inty
  __shared__ T gx_[GRAD ? 2 * NXS : 1];  // den, x strip
  __shared__ T gy_[GRAD ? 2 * NYS : 1];  // den, y strip
  const long so = (long)s * a.slot;
  const Grid& g = a.g;
  const T* fa = a.denalf + so;  // denalf holds falfs
  const T *ix = a.ix + so, *iy = a.iy + so, *iz = a.iz + so;
  const T *dx = GRAD ? nullptr : a.dx + so, *dy = GRAD ? nullptr : a.dy + so, *dz = GRAD ? nullptr : a.dz + so;
  const T* den = GRAD ? a.den + so : nullptr;
  const T dsc = (GRAD && a.dscale) ? (T)a.dscale[s] : T(1);
  T *src = a.src + so, *div = a.div + so;
  const TilePos<TX> tp(g, a.zc);
  if (tp.z0 >= tp.z1) return;
  const int tx = tp.tx, ty = tp.ty;
  const unsigned plane = (unsigned)(g.nx * g.ny);
  const int col = colofs(tp.x, tp.y, g);
  int cx_[KX], cy_[KY];
#pragma unroll
  for (int k = 0; k < KX; k++) {
    const int e = threadIdx.x + k * kThreads, r = e / LX, c = e - LX * (e / LX);
    cx_[k] = e < NXS ? colofs(tp.bx * TX + c - 6, tp.by * TY + r, g) : -1;
  }
#pragma unroll
  for (int k = 0; k < KY; k++) {
    const int e = threadIdx.x + k * kThreads, r = e / TX, c = e - TX * (e / TX);
    cy_[k] = e < NYS ? colofs(tp.bx * TX + c, tp.by * TY + r - 6, g) : -1;
  }
  auto prod = [&](const T* comp, int c, int zo) { return ldz(fa, c, zo) * ldz(comp, c, zo); };
  struct Op { T dx, dy, dz, ix, iy, iz; };
  auto pre = [&](unsigned i) {
    if constexpr (GRAD) return Op{T(0), T(0), T(0), ld32(ix, i), ld32(iy, i), ld32(iz, i)};
    else return Op{ld32(dx, i), ld32(dy, i), ld32(dz, i), ld32(ix, i), ld32(iy, i), ld32(iz, i)};
  };
  T q[13], hx[KX], hy[KY];
  T qd[GRAD ? 13 : 1], hdx[GRAD ? KX : 1], hdy[GRAD ? KY : 1];  // den: z column, strips
#pragma unroll
  for (int k = 0; k < 13; k++) q[k] = prod(iz, col, zofs(tp.z0 - 6 + k, g));
  const int zo0 = zofs(tp.z0, g);
#pragma unroll
  for (int k = 0; k < KX; k++) hx[k] = prod(ix, cx_[k], zo0);
#pragma unroll
  for (int k = 0; k < KY; k++) hy[k] = prod(iy, cy_[k], zo0);
  if constexpr (GRAD) {
#pragma unroll
    for (int k = 0; k < 13; k++) qd[k] = ldz(den, col, zofs(tp.z0 - 6 + k, g));
#pragma unroll
    for (int k = 0; k < KX; k++) hdx[k] = ldz(den, cx_[k], zo0);
#pragma unroll
    for (int k = 0; k < KY; k++) hdy[k] = ldz(den, cy_[k], zo0);
  }
  Op p0{};
  if (tp.inside) p0 = pre((unsigned)col + plane * (unsigned)tp.z0);
#pragma unroll
  for (int k = 0; k < KX; k++)
    if (threadIdx.x + k * kThreads < NXS) lx_[threadIdx.x + k * kThreads] = hx[k];
#pragma unroll
  for (int k = 0; k < KY; k++)
    if (threadIdx.x + k * kThreads < NYS) ly_[threadIdx.x + k * kThreads] = hy[k];
  if constexpr (GRAD) {
#pragma unroll
    for (int k = 0; k < KX; k++)
      if (threadIdx.x + k * kThreads < NXS) gx_[threadIdx.x + k * kThreads] = hdx[k];
#pragma unroll
    for (int k = 0; k < KY; k++)
      if (threadIdx.x + k * kThreads < NYS) gy_[threadIdx.x + k * kThreads] = hdy[k];
  }
  __syncthreads();
  int buf = 0;
  for (int z = tp.z0; z < tp.z1; z++) {
    const bool more = z + 1 < tp.z1;
    T qn = T(0), qdn = T(0);
    Op p1{};
    if (more) {
      qn = prod(iz, col, zofs(z + 7, g));
      const int zo = zofs(z + 1, g);
#pragma unroll
      for (int k = 0; k < KX; k++) hx[k] = prod(ix, cx_[k], zo);
#pragma unroll
      for (int k = 0; k < KY; k++) hy[k] = prod(iy, cy_[k], zo);
      if constexpr (GRAD) {
        qdn = ldz(den, col, zofs(z + 7, g));
#pragma unroll
        for (int k = 0; k < KX; k++) hdx[k] = ldz(den, cx_[k], zo);
#pragma unroll
        for (int k = 0; k < KY; k++) hdy[k] = ldz(den, cy_[k], zo);
      }
      if (tp.inside) p1 = pre((unsigned)col + plane * (unsigned)(z + 1));
    }
    if (tp.inside) {
      const T* px = lx_ + buf * NXS;
      const T* py = ly_ + buf * NYS;
      T ax = T(0), ay = T(0), az = T(0);
#pragma unroll
      for (int m = 0; m < 13; m++) {
        ax += a.sc.d1(0, m) * px[ty * LX + tx + m];
        ay += a.sc.d1(1, m) * py[(ty + m) * TX + tx];
        az += a.sc.d1(2, m) * q[m];
      }
      T gxv = p0.dx, gyv = p0.dy, gzv = p0.dz;
      if constexpr (GRAD) {  // k_grad's sums (zmarch, D1) on the same values
        const T* gxs = gx_ + buf * NXS;
        const T* gys = gy_ + buf * NYS;
        T bx = T(0), by = T(0), bz = T(0);
#pragma unroll
        for (int m = 0; m < 13; m++) {
          bx = vfma(a.sc.d1(0, m), gxs[ty * LX + tx + m], bx);
          by = vfma(a.sc.d1(1, m), gys[(ty + m) * TX + tx], by);
          bz = vfma(a.sc.d1(2, m), qd[m], bz);
        }
        gxv = dsc * bx; gyv = dsc * by; gzv = dsc * bz;
      }
      const unsigned i = g.at(tp.x, tp.y, z);
      st32(div, i, (ax + ay) + az);
      st32(src, i, (gxv * p0.ix + gyv * p0.iy) + gzv * p0.iz);
    }
    if (more) {
      T* nx_ = lx_ + (buf ^ 1) * NXS;
      T* ny_ = ly_ + (buf ^ 1) * NYS;
#pragma unroll
      for (int k = 0; k < KX; k++)
        if (threadIdx.x + k * kThreads < NXS) nx_[threadIdx.x + k * kThreads] = hx[k];
#pragma unroll
      for (int k = 0; k < KY; k++)
        if (threadIdx.x + k * kThreads < NYS) ny_[threadIdx.x + k * kThreads] = hy[k];
      if constexpr (GRAD) {
        T* ngx = gx_ + (buf ^ 1) * NXS;
        T* ngy = gy_ + (buf ^ 1) * NYS;
#pragma unroll
        for (int k = 0; k < KX; k++)
          if (threadIdx.x + k * kThreads < NXS) ngx[threadIdx.x + k * kThreads] = hdx[k];
#pragma unroll
        for (int k = 0; k < KY; k++)
          if (threadIdx.x + k * kThreads < NYS) ngy[threadIdx.x + k * kThreads] = hdy[k];
      }
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < 12; k++) q[k] = q[k + 1];
    q[12] = qn;
    if constexpr (GRAD) {
#pragma unroll
      for (int k = 0; k < 12; k++) qd[k] = qd[k + 1];
      qd[12] = qdn;
    }
    p0 = p1;
    buf ^= 1;
  }
}

// ---------------------------------------------------------------------------
// energy() densities (energy.f90:47-165), the chemical potential, the
// helium centre of mass (r_cm.f90:38-54), the X+-He term and |H psi|^2 (the
// residual of the eigenvalue equation): 13 block partial sums.
constexpr int kEnergySums = 13;
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
  const PlusSlot* plus;  // [slot] X+ potential of uplus
  double* partial;
  int nblk;
};
template <typename T, typename V, int TX>
__global__ __launch_bounds__(kThreads) void k_energy(EnergyArgs<T> a) {
  const int s = blockIdx.z;
  if (!a.active[s]) return;
  __shared__ V lds[2 * Tile<TX>::PL];
  __shared__ double red[4 * kEnergySums];
  const long so = (long)s * a.slot;
  const V* psi = reinterpret_cast<const V*>(a.psi_all + ((long)s * 5 + a.ring[8 * s]) * a.slot);
  const PlusSlot& pl = a.plus[s];
  const bool plus_on = pl.on != 0, excl_on = pl.excl > 0.0;
  double v[kEnergySums];
#pragma unroll
  for (int k = 0; k < kEnergySums; k++) v[k] = 0.0;
  zmarch<T, V, TX, true, true>(psi, a.g, a.sc, TilePos<TX>(a.g, a.zc), lds, [](unsigned) { return NoPre{}; },
    [&](int x, int y, int z, bool in, V p, V(&d1)[3], V(&d2)[3], const NoPre&) {
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
      // |rimp - r| for uplus and the exclusion sphere (energy.f90:149-157)
      double r = 0.0;
      if (plus_on || excl_on) {
        const double rx = pl.rimp[0] - (a.x0 + a.hx * x), ry = pl.rimp[1] - (a.y0 + a.hy * y),
                     rz = pl.rimp[2] - (a.z0 + a.hz * z);
        r = sqrt((rx * rx + ry * ry) + rz * rz);
      }
      const bool keep = !excl_on || r >= pl.excl;
      if (keep) v[5] += (double)(d * a.uimp[i]);
      v[6] += (double)d;
      v[7] += (double)d * (a.x0 + a.hx * x);
      v[8] += (double)d * (a.y0 + a.hy * y);
      v[9] += (double)d * (a.z0 + a.hz * z);
      V hp = hpsi_of(p, d2, a.h2o2m4, a.pot4[i], a.uimp[i]);
      v[10] -= (double)dotre(p, hp);
      v[12] += (double)norm2(hp);
      if (plus_on && keep) v[11] += (double)d * plus_pot(pl.p, r);  // uplus: Select_Pot(selec_plus, r)
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
