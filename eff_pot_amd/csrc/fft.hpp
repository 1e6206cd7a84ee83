// Batched 3-D real<->complex FFTs for gfx950, built from one device routine:
// a length-N complex transform done as N = N1*N2 with a register FFT of
// length N1 per thread, a twiddle, one LDS transpose and a register FFT of
// length N2.  Three kernel families use it:
//   x_r2c / x_c2r : rows along x (contiguous), two real rows packed into one
//                   complex row; output is the half spectrum (nx/2+1,ny,nz);
//   strided       : complex lines along y or z of the half spectrum, C lines
//                   (consecutive kx) per work-group so every global access is
//                   a contiguous run of C complex values;
//   z_conv        : forward z transform, multiply by up to 3 real k-space
//                   kernels (times 1/(nx ny nz)), inverse z transform, all
//                   without leaving the work-group (the spectrum stays in
//                   registers between the products).
// Replaces fft.f90:6-144 (FFTW r2c/c2r plans) and the k-space products of
// poten.f90:34-71,95-97 and term_alfa.f90:172-205.
#pragma once
#include "common.hpp"

namespace ep {

// ---- 48th roots of unity (all register-FFT twiddles are powers of them) ----
struct Tw48 { double c[48], s[48]; };
constexpr double ce_sin_(double x) {
  double x2 = x * x, t = x, s = x;
  for (int k = 1; k < 14; k++) { t *= -x2 / ((2.0 * k) * (2.0 * k + 1.0)); s += t; }
  return s;
}
constexpr double ce_cos_(double x) {
  double x2 = x * x, t = 1.0, s = 1.0;
  for (int k = 1; k < 14; k++) { t *= -x2 / ((2.0 * k - 1.0) * (2.0 * k)); s += t; }
  return s;
}
constexpr Tw48 make_tw48() {
  Tw48 w{};
  const double pi24 = 3.14159265358979323846 / 24.0;
  for (int j = 0; j < 48; j++) {
    int q = j / 6, r = j % 6;
    double Cr = ce_cos_(r * pi24), Sr = ce_sin_(r * pi24);
    double Cm = ce_cos_((6 - r) * pi24), Sm = ce_sin_((6 - r) * pi24);
    double c = 0, s = 0;  // cos, sin of 2 pi j / 48
    switch (q) {
      case 0: c = Cr; s = Sr; break;
      case 1: c = Sm; s = Cm; break;
      case 2: c = -Sr; s = Cr; break;
      case 3: c = -Cm; s = Sm; break;
      case 4: c = -Cr; s = -Sr; break;
      case 5: c = -Sm; s = -Cm; break;
      case 6: c = Sr; s = -Cr; break;
      default: c = Cm; s = -Sm; break;
    }
    w.c[j] = c;
    w.s[j] = -s;  // exp(-i theta)
  }
  return w;
}
struct Tw48f { float c[48], s[48]; };
constexpr Tw48f make_tw48f() {
  Tw48 d = make_tw48();
  Tw48f f{};
  for (int j = 0; j < 48; j++) { f.c[j] = (float)d.c[j]; f.s[j] = (float)d.s[j]; }
  return f;
}
__constant__ Tw48 c_tw48 = make_tw48();
__constant__ Tw48f c_tw48f = make_tw48f();

template <typename T> EP_DEV cx<T> tw48(int j);
template <> EP_DEV cx<double> tw48<double>(int j) { return {c_tw48.c[j], c_tw48.s[j]}; }
template <> EP_DEV cx<float> tw48<float>(int j) { return {c_tw48f.c[j], c_tw48f.s[j]}; }

// ---- small DFT butterflies (natural order in and out) -----------------------
template <typename T> EP_DEV void dft2(cx<T>* v) {
  cx<T> a = v[0], b = v[1];
  v[0] = a + b; v[1] = a - b;
}
template <typename T> EP_DEV void dft3(cx<T>* v, int sign) {
  const T h = T(0.86602540378443864676);
  cx<T> t = v[1] + v[2], u = v[1] - v[2];
  cx<T> m = v[0] - T(0.5) * t;
  cx<T> r = mul_si(h * u, sign);
  v[0] = v[0] + t; v[1] = m + r; v[2] = m - r;
}
template <typename T> EP_DEV void dft4(cx<T>* v, int sign) {
  cx<T> s02 = v[0] + v[2], d02 = v[0] - v[2], s13 = v[1] + v[3], d13 = mul_si(v[1] - v[3], sign);
  v[0] = s02 + s13; v[2] = s02 - s13; v[1] = d02 + d13; v[3] = d02 - d13;
}
template <int R, typename T> EP_DEV void dft_r(cx<T>* v, int sign) {
  if constexpr (R == 2) dft2(v);
  else if constexpr (R == 3) dft3(v, sign);
  else dft4(v, sign);
}

// One Stockham pass on a register array (all indices compile-time after unrolling).
template <typename T, int L, int P, int R>
EP_DEV void reg_pass(const cx<T>* src, cx<T>* dst, int sign) {
  constexpr int M = L / R;
  static_assert(48 % (P * R) == 0, "twiddle must be a 48th root");
#pragma unroll
  for (int i = 0; i < M; i++) {
    const int k = i % P;
    cx<T> v[R];
#pragma unroll
    for (int j = 0; j < R; j++) {
      v[j] = src[i + j * M];
      if (j > 0 && k > 0) {
        cx<T> w = tw48<T>((j * k * (48 / (P * R))) % 48);
        if (sign > 0) w.y = -w.y;
        v[j] = cmul(v[j], w);
      }
    }
    dft_r<R>(v, sign);
    const int base = (i - k) * R + k;
#pragma unroll
    for (int j = 0; j < R; j++) dst[base + j * P] = v[j];
  }
}

template <typename T, int L> struct RegFFT;
template <typename T> struct RegFFT<T, 3> { static EP_DEV void run(cx<T>* a, int s) { dft3(a, s); } };
template <typename T> struct RegFFT<T, 4> { static EP_DEV void run(cx<T>* a, int s) { dft4(a, s); } };
template <typename T> struct RegFFT<T, 6> {
  static EP_DEV void run(cx<T>* a, int s) { cx<T> b[6]; reg_pass<T, 6, 1, 2>(a, b, s); reg_pass<T, 6, 2, 3>(b, a, s); }
};
template <typename T> struct RegFFT<T, 8> {
  static EP_DEV void run(cx<T>* a, int s) { cx<T> b[8]; reg_pass<T, 8, 1, 4>(a, b, s); reg_pass<T, 8, 4, 2>(b, a, s); }
};
template <typename T> struct RegFFT<T, 12> {
  static EP_DEV void run(cx<T>* a, int s) { cx<T> b[12]; reg_pass<T, 12, 1, 4>(a, b, s); reg_pass<T, 12, 4, 3>(b, a, s); }
};
template <typename T> struct RegFFT<T, 16> {
  static EP_DEV void run(cx<T>* a, int s) { cx<T> b[16]; reg_pass<T, 16, 1, 4>(a, b, s); reg_pass<T, 16, 4, 4>(b, a, s); }
};

// N = N1*N2 (N1 >= N2, N1 a power of two dividing 256): C = 256/N1 lines per group.
template <int N> struct Shape;
template <> struct Shape<16> { static constexpr int N1 = 4, N2 = 4; };
template <> struct Shape<24> { static constexpr int N1 = 8, N2 = 3; };
template <> struct Shape<32> { static constexpr int N1 = 8, N2 = 4; };
template <> struct Shape<48> { static constexpr int N1 = 8, N2 = 6; };
template <> struct Shape<64> { static constexpr int N1 = 8, N2 = 8; };
template <> struct Shape<96> { static constexpr int N1 = 16, N2 = 6; };
template <> struct Shape<128> { static constexpr int N1 = 16, N2 = 8; };
template <> struct Shape<192> { static constexpr int N1 = 16, N2 = 12; };
template <> struct Shape<256> { static constexpr int N1 = 16, N2 = 16; };

inline bool fft_size_supported(int n) {
  switch (n) { case 16: case 24: case 32: case 48: case 64: case 96: case 128: case 192: case 256: return true; }
  return false;
}

// Accessor describing where line c, element i lives in global memory.
// Lines c = cx + ncx*co; address = base + cx*1 + co*s_outer + i*s_elem.  A run
// is a padded row of the half spectrum: ncx (a multiple of 16) lines of which
// the first nvalid exist, so groups of 16 consecutive lines start on 128-byte
// boundaries; the padding lines are neither read nor written.
struct LineMap {
  int ncx;          // lines per run (padded row length)
  long s_outer;     // stride between runs
  long s_elem;      // stride between elements of a line
  int nlines;       // lines per field and slot (padding included)
  int nvalid;       // lines of a run that exist (nx/2+1)
  int first = 0;    // first line of a run that is processed (1: the packed column 0 is skipped)
  // 32-bit element offset from the field's base (a field of one slot holds
  // fewer than 2^32 elements): with a uniform base the loads and stores take
  // the SGPR-base + 32-bit VGPR-offset form, which frees the per-lane 64-bit
  // addresses (4-7% faster z convolutions at 128^3 than 64-bit offsets)
  EP_DEV unsigned addr(int c, int i) const {
    int co = c / ncx, ci = c - co * ncx;
    return (unsigned)ci + (unsigned)co * (unsigned)s_outer + (unsigned)i * (unsigned)s_elem;
  }
  EP_DEV bool valid(int c) const {
    const int ci = c - (c / ncx) * ncx;
    return c < nlines && ci >= first && ci < nvalid;
  }
  // A thread of the strided and z-convolution passes reads and writes one
  // line only: its base offset (addr(c, 0)) and validity are formed once, so
  // the runtime division by ncx is not repeated at every access.  A line that
  // is skipped (a padding line or the unused packed column) keeps its own
  // address, which lies in the allocated row next to its neighbours' (lines
  // past the last are given the last line's), so its loads may be issued
  // and their values discarded without touching other cache lines.
  struct Line {
    unsigned base, step;
    bool ok;
    EP_DEV unsigned at(int i) const { return base + (unsigned)i * step; }
  };
  EP_DEV Line line(int c) const {
    return Line{addr(min(c, nlines - 1), 0), (unsigned)s_elem, valid(c)};
  }
};

// ---- the two-level transform, step by step -------------------------------------
// LDS layout between the steps: index (k1*N2 + n2)*C + c (c fastest: conflict free).

// The group has TH threads (kThreads unless a kernel says otherwise) and
// works on C = TH / N1 lines.
// Twiddles: every kernel multiplies by twN[a b] whenever the unrolled index a
// is nonzero, whatever the lane's index b.  twN[0] = (1, -0) and a product by
// it returns its operand (up to the sign of a zero), so the lanes with b = 0
// lose nothing; under a lane-dependent branch each twiddle load waited out
// its whole latency before the next was issued.
// Step 1+2 for the lines of this group: a[n1] = x[N2 n1 + n2]; FFT_N1; twiddle; to LDS.
template <typename T, int N, int TH = kThreads, class Load>
EP_DEV void fft_step1(cx<T>* lds, const cx<T>* twN, int sign, Load load) {
  constexpr int N1 = Shape<N>::N1, N2 = Shape<N>::N2, C = TH / N1;
  const int t = threadIdx.x;
  if (t < C * N2) {
    const int c = t % C, n2 = t / C;
    cx<T> a[N1];
#pragma unroll
    for (int n1 = 0; n1 < N1; n1++) a[n1] = load(c, N2 * n1 + n2);
    RegFFT<T, N1>::run(a, sign);
#pragma unroll
    for (int k1 = 0; k1 < N1; k1++) {
      cx<T> v = a[k1];
      if (k1 > 0) {
        cx<T> w = ld32(twN, n2 * k1);
        if (sign > 0) w.y = -w.y;
        v = cmul(v, w);
      }
      lds[(k1 * N2 + n2) * C + c] = v;
    }
  }
}
// Step 3: thread (c,k1) gathers b[n2] from LDS, FFT_N2; X[k1 + N1 k2] = b[k2].
template <typename T, int N, int TH = kThreads>
EP_DEV void fft_step3(const cx<T>* lds, cx<T> (&b)[Shape<N>::N2], int sign) {
  constexpr int N1 = Shape<N>::N1, N2 = Shape<N>::N2, C = TH / N1;
  const int t = threadIdx.x;
  const int c = t % C, k1 = t / C;
#pragma unroll
  for (int n2 = 0; n2 < N2; n2++) b[n2] = lds[(k1 * N2 + n2) * C + c];
  RegFFT<T, N2>::run(b, sign);
}
// Inverse of the layout above starting from registers: thread (c,k1) holds
// X[k1 + N1 j] (j < N2).  Decomposition N = N2*N1: FFT_N2 over j, twiddle
// w^{k1 k1'}, LDS, FFT_N1 -> x[k1' + N2 k2'] delivered to store(c, i, v).
template <typename T, int N, int TH = kThreads, class Store>
EP_DEV void fft_from_regs(cx<T> (&a)[Shape<N>::N2], cx<T>* lds, const cx<T>* twN, int sign, Store store) {
  constexpr int N1 = Shape<N>::N1, N2 = Shape<N>::N2, C = TH / N1;
  const int t = threadIdx.x;
  {
    const int c = t % C, k1 = t / C;
    RegFFT<T, N2>::run(a, sign);
#pragma unroll
    for (int kk = 0; kk < N2; kk++) {
      cx<T> v = a[kk];
      if (kk > 0) {
        cx<T> w = ld32(twN, k1 * kk);
        if (sign > 0) w.y = -w.y;
        v = cmul(v, w);
      }
      lds[(kk * N1 + k1) * C + c] = v;
    }
  }
  __syncthreads();
  if (t < C * N2) {
    const int c = t % C, kk = t / C;
    cx<T> b[N1];
#pragma unroll
    for (int m = 0; m < N1; m++) b[m] = lds[(kk * N1 + m) * C + c];
    RegFFT<T, N1>::run(b, sign);
#pragma unroll
    for (int m = 0; m < N1; m++) store(c, kk + N2 * m, b[m]);
  }
}

// ---- strided complex lines (y or z axis of the half spectrum) ---------------
// grid: (ceil(nlines/C), nfields, nslots); in place on f[blockIdx.y] + slot*sstride.
template <typename T> struct StridedArgs {
  cx<T>* f[4];
  long sstride;
  LineMap lm;
  const cx<T>* twN;
  int sign;
  const int* active;
};
template <typename T, int N>
__global__ __launch_bounds__(kThreads) void k_fft_strided(StridedArgs<T> a) {
  constexpr int N1 = Shape<N>::N1, N2 = Shape<N>::N2, C = kThreads / N1;
  if (!a.active[blockIdx.z]) return;
  __shared__ cx<T> lds[N * C];
  const LineMap lm = a.lm;
  cx<T>* d = a.f[blockIdx.y] + (long)blockIdx.z * a.sstride;
  const int c0 = blockIdx.x * C;
  // every access of this thread (step 1 loads, step 3 stores) is on line c0 + t % C
  const LineMap::Line ln = lm.line(c0 + (int)(threadIdx.x % C));
  // unconditional loads (a skipped line reads its own, allocated, positions)
  // and a select: no lane branch around each load
  auto load = [&](int, int i) -> cx<T> {
    const cx<T> v = ld32(d, ln.at(i));
    return ln.ok ? v : cx<T>{T(0), T(0)};
  };
  fft_step1<T, N>(lds, a.twN, a.sign, load);
  __syncthreads();
  cx<T> b[N2];
  fft_step3<T, N>(lds, b, a.sign);
  const int k1 = threadIdx.x / C;
  if (ln.ok) {
#pragma unroll
    for (int k2 = 0; k2 < N2; k2++) d[ln.at(k1 + N1 * k2)] = b[k2];
  }
}

// ---- z transform with the k-space products in between ----------------------
// Group g (blockIdx.y): input in[g]; nout[g] outputs out[g][j] = IFFT_z(K * FFT_z(in))
// with K = kern[g][j] (M reals shared by all slots) times `scale`.
template <typename T> struct ConvArgs {
  const cx<T>* in[4];
  cx<T>* out[4][3];
  const T* kern[4][3];
  int nout[4];
  long sstride;
  LineMap lm;
  LineMap km;  // the same lines in the k-space factor tables (padded rows of nx/2+1)
  int nw;      // packed layout: the first nw groups (blockIdx.x) convolve the packed column W
  int P, ny;   // packed row length (nx/2), rows per plane
  long wofs;   // W(ky, kz) at the slot's spectrum + wofs + kz ny + ky
  const T* kw[4][3];  // factor pairs ((K0+KH)/2, (K0-KH)/2) of kern[g][j] at [kz][ky]
  T scale;
  const cx<T>* twN;
  const int* active;
};
// ---- z convolution of the packed column (packed layout) --------------------
// In the packed layout the rows hold kx = 1 .. H-1 (H = nx/2; their column 0
// is unused) and a block of its own holds W = S(0) + i S(H) for every (ky, z),
// ky fastest: the y-transformed columns kx = 0 and kx = H, both spectra of real
// columns, hence Hermitian in (ky, kz).  After the z transform the packed value
// at k = (ky, kz) is W(k) = S0(k) + i SH(k) with S0(k) = (W(k) + conj W(-k))/2
// and i SH(k) = (W(k) - conj W(-k))/2, so the products K0 S0 + i KH SH with the
// (real, even) k-space factors of columns 0 and H are
//   O(k) = (K0 + KH)/2 W(k) + (K0 - KH)/2 conj W(-k),   -k = (-ky, -kz),
// the packed form of the two convolved columns.  Group gx transforms C of the
// ny lines (kx = 0, ky) with their partners (kx = 0, -ky) in the same group:
// positions 0 and 1 hold ky = 0 and ny/2 (their own partners), then ky = m and
// ny - m at positions 2m and 2m + 1.  The partner's spectrum comes through LDS
// (rows of N + 1: no bank conflicts).
template <typename T, int N, int TH, bool ONE>
EP_DEV void zconv_packed_column(const ConvArgs<T>& a, cx<T>* lds, int gx) {
  constexpr int N1 = Shape<N>::N1, N2 = Shape<N>::N2, C = TH / N1;
  const int g = blockIdx.y;
  const long so = (long)blockIdx.z * a.sstride;
  const cx<T>* in = a.in[g] + so;
  const int c = threadIdx.x % C, kk1 = threadIdx.x / C;
  const int ny = a.ny, hy = ny / 2;
  const int p = gx * C + c;
  const bool valid = p < ny;
  const int ky = p == 0 ? 0 : p == 1 ? hy : ((p & 1) ? ny - (p >> 1) : (p >> 1));
  const int pc = p <= 1 ? c : (c ^ 1);  // the partner's lane in this group
  const int kyl = valid ? ky : 0;  // the row the loads use (in range for every lane)
  const long srow = ny;  // W is stored kz-major, ky fastest
  const long sb = a.wofs + kyl;
  auto load = [&](int, int i) -> cx<T> {
    const cx<T> v = in[sb + (long)i * srow];
    return valid ? v : cx<T>{T(0), T(0)};
  };
  fft_step1<T, N, TH>(lds, a.twN, -1, load);
  __syncthreads();
  cx<T> spec[N2];
  fft_step3<T, N, TH>(lds, spec, -1);  // spec[k2] = W(ky, kz = kk1 + N1 k2)
  __syncthreads();
#pragma unroll
  for (int k2 = 0; k2 < N2; k2++) lds[c * (N + 1) + kk1 + N1 * k2] = spec[k2];
  __syncthreads();
  cx<T> part[N2];  // conj W(-ky, -kz)
#pragma unroll
  for (int k2 = 0; k2 < N2; k2++) {
    const int kz = kk1 + N1 * k2, km = kz == 0 ? 0 : N - kz;
    part[k2] = conjg(lds[pc * (N + 1) + km]);
  }
  const int nout = ONE ? 1 : a.nout[g];
  for (int j = 0; j < nout; j++) {
    const T* K = a.kw[g][j];
    cx<T> v[N2];
#pragma unroll
    for (int k2 = 0; k2 < N2; k2++) {
      const long kk = 2 * ((long)(kk1 + N1 * k2) * ny + kyl);
      const T kp = K[kk], kq = K[kk + 1];
      const T ap = (valid ? kp : T(0)) * a.scale, am = (valid ? kq : T(0)) * a.scale;
      v[k2] = cx<T>{ap * spec[k2].x + am * part[k2].x, ap * spec[k2].y + am * part[k2].y};
    }
    __syncthreads();  // LDS reuse
    cx<T>* out = a.out[g][j] + so;
    auto store = [&](int, int i, cx<T> val) {
      if (valid) out[sb + (long)i * srow] = val;
    };
    fft_from_regs<T, N, TH>(v, lds, a.twN, +1, store);
  }
}

// TH = 256 or 512 threads: C = 16 or 32 lines per group (128 B or 256 B runs
// of FP32 complex per z plane); the autotuner picks one per context.  ONE:
// every input has one output, so nothing is kept between outputs (a separate
// instance because the register allocation follows the longest path).  In the
// packed layout the first nw groups (blockIdx.x) convolve the packed column,
// so that small piece of work runs alongside the other lines.
template <typename T, int N, int TH = kThreads, bool ONE = false>
__global__ __launch_bounds__(TH) void k_fft_zconv(ConvArgs<T> a) {
  constexpr int N1 = Shape<N>::N1, N2 = Shape<N>::N2, C = TH / N1;
  if (!a.active[blockIdx.z]) return;
  __shared__ cx<T> lds[(N + 1) * C];
  if ((int)blockIdx.x < a.nw) {
    zconv_packed_column<T, N, TH, ONE>(a, lds, blockIdx.x);
    return;
  }
  const int g = blockIdx.y;
  const LineMap lm = a.lm;
  const long so = (long)blockIdx.z * a.sstride;
  const cx<T>* in = a.in[g] + so;
  const int c0 = (blockIdx.x - a.nw) * C;
  // every access of this thread (loads, k-space factors, stores) is on line c0 + t % C
  const int kk1 = threadIdx.x / C;
  const LineMap::Line ln = lm.line(c0 + (int)(threadIdx.x % C));
  const LineMap::Line lk = a.km.line(c0 + (int)(threadIdx.x % C));
  const bool valid = ln.ok;
  // unconditional loads and selects, as in k_fft_strided
  auto load = [&](int, int i) -> cx<T> {
    const cx<T> v = ld32(in, ln.at(i));
    return valid ? v : cx<T>{T(0), T(0)};
  };
  const int nout = ONE ? 1 : a.nout[g];
  // k-space factors of this thread's outputs: the first output's are loaded
  // before the forward transform, each next output's during the previous
  // inverse transform, so their latency overlaps arithmetic
  T kn[N2];
  auto load_k = [&](int j) {
    const T* K = a.kern[g][j];
#pragma unroll
    for (int k2 = 0; k2 < N2; k2++) {
      const T k = ld32(K, lk.at(kk1 + N1 * k2));
      kn[k2] = valid ? k : T(0);
    }
  };
  load_k(0);
  fft_step1<T, N, TH>(lds, a.twN, -1, load);
  __syncthreads();
  cx<T> spec[N2];
  fft_step3<T, N, TH>(lds, spec, -1);
  for (int j = 0; j < nout; j++) {
    cx<T> v[N2];
#pragma unroll
    for (int k2 = 0; k2 < N2; k2++) v[k2] = (kn[k2] * a.scale) * spec[k2];
    if (j + 1 < nout) load_k(j + 1);
    __syncthreads();  // LDS reuse
    cx<T>* out = a.out[g][j] + so;
    auto store = [&](int, int i, cx<T> val) {
      if (valid) st32(out, ln.at(i), val);
    };
    fft_from_regs<T, N, TH>(v, lds, a.twN, +1, store);
  }
}

// ---- x axis: two real rows <-> one complex row ------------------------------
// Row r (0 <= r < ny*nz) of field f, slot s: real at r[f] + s*rslot + r*nx,
// half spectrum at c[f] + s*cslot + r*nxh.
template <typename T> struct XArgs {
  T* r[4];
  cx<T>* c[4];
  long rslot, cslot;
  int nrows;  // ny*nz (even)
  int pitch;  // complex elements per half-spectrum row (nx/2+1 padded)
  T scale;    // c2r only: out = aff0[f] + (scale*aff1[f]) * x
  const double* rscale;  // r2c only: [slot] factor field 0 is read with, or null
  const T* rmul;         // r2c only: fields 1.. are read times rmul (a field laid out like r), or null
  T aff0[4], aff1[4];
  const cx<T>* twN;
  const int* active;
};
template <typename T, int N>
__global__ __launch_bounds__(kThreads) void k_fft_x_r2c(XArgs<T> rm) {
  const cx<T>* twN = rm.twN;
  constexpr int N1 = Shape<N>::N1, N2 = Shape<N>::N2, C = kThreads / N1, NH = N / 2 + 1;
  constexpr int P = C + 1;  // padded line stride in the staging layout [i][c]
  if (!rm.active[blockIdx.z]) return;
  __shared__ cx<T> lds[N * P];
  const int f = blockIdx.y;
  const T* src = rm.r[f] + (long)blockIdx.z * rm.rslot;
  cx<T>* dst = rm.c[f] + (long)blockIdx.z * rm.cslot;
  const int pair0 = blockIdx.x * C;  // pairs of rows
  const int npairs = rm.nrows / 2;
  const T sc = (rm.rscale && f == 0) ? (T)rm.rscale[blockIdx.z] : T(1);
  const T* mul = (rm.rmul && f > 0) ? rm.rmul + (long)blockIdx.z * rm.rslot : nullptr;
  // stage: z_c[i] = a_c[i] + i b_c[i]
  for (int e = threadIdx.x; e < C * N; e += kThreads) {
    const int c = e / N, i = e - c * N;
    const int pr = pair0 + c;
    cx<T> v{T(0), T(0)};
    if (pr < npairs) {
      const T* ra = src + (long)(2 * pr) * N;
      v = cx<T>{sc * ra[i], sc * ra[N + i]};
      if (mul) {
        const T* ma = mul + (long)(2 * pr) * N;
        v = cx<T>{v.x * ma[i], v.y * ma[N + i]};
      }
    }
    lds[i * P + c] = v;
  }
  __syncthreads();
  cx<T> a[N1];
  const int t = threadIdx.x;
  const bool s1 = t < C * N2;
  if (s1) {
    const int c = t % C, n2 = t / C;
#pragma unroll
    for (int n1 = 0; n1 < N1; n1++) a[n1] = lds[(N2 * n1 + n2) * P + c];
  }
  __syncthreads();
  if (s1) {
    const int c = t % C, n2 = t / C;
    RegFFT<T, N1>::run(a, -1);
#pragma unroll
    for (int k1 = 0; k1 < N1; k1++) {
      cx<T> v = a[k1];
      if (k1 > 0) v = cmul(v, twN[n2 * k1]);
      lds[(k1 * N2 + n2) * C + c] = v;
    }
  }
  __syncthreads();
  cx<T> b[N2];
  fft_step3<T, N>(lds, b, -1);
  __syncthreads();
  {
    const int c = t % C, k1 = t / C;
#pragma unroll
    for (int k2 = 0; k2 < N2; k2++) lds[(k1 + N1 * k2) * P + c] = b[k2];
  }
  __syncthreads();
  for (int e = threadIdx.x; e < C * NH; e += kThreads) {
    const int c = e / NH, k = e - c * NH;
    const int pr = pair0 + c;
    if (pr >= npairs) continue;
    cx<T> z = lds[k * P + c], zm = lds[((N - k) % N) * P + c];
    cx<T> s = cx<T>{z.x + zm.x, z.y - zm.y};  // z + conj(zm)
    cx<T> d = cx<T>{z.x - zm.x, z.y + zm.y};  // z - conj(zm)
    cx<T>* oa = dst + (long)(2 * pr) * rm.pitch;
    oa[k] = cx<T>{T(0.5) * s.x, T(0.5) * s.y};
    oa[rm.pitch + k] = cx<T>{T(0.5) * d.y, T(-0.5) * d.x};
  }
}

// Half-spectrum rows -> real rows (out = aff0 + scale*aff1*x per field); the imaginary parts of
// the self-conjugate bins 0 and N/2 are ignored, as a c2r transform does.
template <typename T, int N>
__global__ __launch_bounds__(kThreads) void k_fft_x_c2r(XArgs<T> rm) {
  const cx<T>* twN = rm.twN;
  const T off = rm.aff0[blockIdx.y];
  const T scale = rm.scale * rm.aff1[blockIdx.y];
  constexpr int N1 = Shape<N>::N1, N2 = Shape<N>::N2, C = kThreads / N1, NH = N / 2 + 1;
  constexpr int P = C + 1;
  if (!rm.active[blockIdx.z]) return;
  __shared__ cx<T> lds[N * P];
  const int f = blockIdx.y;
  const cx<T>* src = rm.c[f] + (long)blockIdx.z * rm.cslot;
  T* dst = rm.r[f] + (long)blockIdx.z * rm.rslot;
  const int pair0 = blockIdx.x * C;
  const int npairs = rm.nrows / 2;
  for (int e = threadIdx.x; e < C * NH; e += kThreads) {
    const int c = e / NH, k = e - c * NH;
    const int pr = pair0 + c;
    cx<T> A{T(0), T(0)}, B{T(0), T(0)};
    if (pr < npairs) {
      const cx<T>* ra = src + (long)(2 * pr) * rm.pitch;
      A = ra[k]; B = ra[rm.pitch + k];
    }
    if (k == 0 || k == N / 2) { A.y = T(0); B.y = T(0); }
    lds[k * P + c] = cx<T>{A.x - B.y, A.y + B.x};
    if (k > 0 && k < N / 2) lds[(N - k) * P + c] = cx<T>{A.x + B.y, B.x - A.y};
  }
  __syncthreads();
  const int t = threadIdx.x;
  cx<T> a[N1];
  const bool s1 = t < C * N2;
  if (s1) {
    const int c = t % C, n2 = t / C;
#pragma unroll
    for (int n1 = 0; n1 < N1; n1++) a[n1] = lds[(N2 * n1 + n2) * P + c];
  }
  __syncthreads();
  if (s1) {
    const int c = t % C, n2 = t / C;
    RegFFT<T, N1>::run(a, +1);
#pragma unroll
    for (int k1 = 0; k1 < N1; k1++) {
      cx<T> v = a[k1];
      if (k1 > 0) v = cmul(v, conjg(twN[n2 * k1]));
      lds[(k1 * N2 + n2) * C + c] = v;
    }
  }
  __syncthreads();
  cx<T> b[N2];
  fft_step3<T, N>(lds, b, +1);
  __syncthreads();
  {
    const int c = t % C, k1 = t / C;
#pragma unroll
    for (int k2 = 0; k2 < N2; k2++) lds[(k1 + N1 * k2) * P + c] = b[k2];
  }
  __syncthreads();
  for (int e = threadIdx.x; e < C * N; e += kThreads) {
    const int c = e / N, i = e - c * N;
    const int pr = pair0 + c;
    if (pr >= npairs) continue;
    cx<T> z = lds[i * P + c];
    T* ra = dst + (long)(2 * pr) * N;
    ra[i] = off + scale * z.x;
    ra[N + i] = off + scale * z.y;
  }
}

// ---- forward x and y transforms of one z plane in one work-group ------------
// The half spectrum of a whole (nx = ny = N) plane stays in LDS between the x
// and the y transform, so the forward y pass costs no HBM round trip: real
// rows -> x r2c (pairs of rows as in k_fft_x_r2c) -> plane -> y FFT of the NH
// columns -> half spectrum in global memory.  (The same fusion of the inverse
// y and x c2r passes measured slower than the two separate passes.)
// The work-group is G sub-groups of kThreads threads, each with its own
// scratch, working on different row-pair or column batches at the same time
// (more waves in flight per CU than one 256-thread group with an 80 KB plane).
// grid: (nz, nfields, nslots).  Used where the plane and the G scratches fit
// in LDS (PlaneFits).
template <typename T, int N> struct PlaneFits {
  static constexpr int C = kThreads / Shape<N>::N1;
  static constexpr int G = sizeof(T) == 4 ? 4 : 1;
  static constexpr long bytes = ((long)N * (N / 2 + 1) + (long)G * N * (C + 1)) * (long)sizeof(cx<T>);
  static constexpr bool value = bytes <= 160 * 1024;
};
template <typename T> struct PlaneArgs {
  T* r[4];
  cx<T>* c[4];
  long rslot, cslot;
  int pitch;  // complex elements per half-spectrum row (nx/2+1 padded, or nx/2 packed)
  int packed; // packed layout: rows of nx/2 (column 0 unused) and W = S(kx=0) + i S(kx=nx/2)
  long wofs;  // packed: W of plane z, row ky at the slot's spectrum + wofs + z ny + ky
  const double* rscale;  // forward only: [slot] factor field 0 is read with, or null
  const cx<T>* twx;
  const cx<T>* twy;
  const int* active;
  T scale;            // inverse only: out = aff0[f] + (scale*aff1[f]) * x
  T aff0[4], aff1[4];
};

// fft_step1 / fft_step3 for sub-group thread t (0 <= t < kThreads).
template <typename T, int N, class Load>
EP_DEV void sub_step1(cx<T>* lds, const cx<T>* twN, int sign, Load load, int t, bool on) {
  constexpr int N1 = Shape<N>::N1, N2 = Shape<N>::N2, C = kThreads / N1;
  if (on && t < C * N2) {
    const int c = t % C, n2 = t / C;
    cx<T> a[N1];
#pragma unroll
    for (int n1 = 0; n1 < N1; n1++) a[n1] = load(c, N2 * n1 + n2);
    RegFFT<T, N1>::run(a, sign);
#pragma unroll
    for (int k1 = 0; k1 < N1; k1++) {
      cx<T> v = a[k1];
      if (k1 > 0) v = cmul(v, sign < 0 ? twN[n2 * k1] : conjg(twN[n2 * k1]));
      lds[(k1 * N2 + n2) * C + c] = v;
    }
  }
}
template <typename T, int N>
EP_DEV void sub_step3(const cx<T>* lds, cx<T> (&b)[Shape<N>::N2], int sign, int t) {
  constexpr int N1 = Shape<N>::N1, N2 = Shape<N>::N2, C = kThreads / N1;
  const int c = t % C, k1 = t / C;
#pragma unroll
  for (int n2 = 0; n2 < N2; n2++) b[n2] = lds[(k1 * N2 + n2) * C + c];
  RegFFT<T, N2>::run(b, sign);
}
// x transform of C row pairs staged as lds[i*P + c]; leaves X[k] in lds[k*P + c].
template <typename T, int N>
EP_DEV void plane_x_fft(cx<T>* lds, const cx<T>* twN, int sign, int t, bool on) {
  constexpr int N1 = Shape<N>::N1, N2 = Shape<N>::N2, C = kThreads / N1, P = C + 1;
  cx<T> a[N1];
  const bool s1 = on && t < C * N2;
  if (s1) {
    const int c = t % C, n2 = t / C;
#pragma unroll
    for (int n1 = 0; n1 < N1; n1++) a[n1] = lds[(N2 * n1 + n2) * P + c];
  }
  __syncthreads();
  if (s1) {
    const int c = t % C, n2 = t / C;
    RegFFT<T, N1>::run(a, sign);
#pragma unroll
    for (int k1 = 0; k1 < N1; k1++) {
      cx<T> v = a[k1];
      if (k1 > 0) v = cmul(v, sign < 0 ? twN[n2 * k1] : conjg(twN[n2 * k1]));
      lds[(k1 * N2 + n2) * C + c] = v;
    }
  }
  __syncthreads();
  cx<T> b[N2];
  if (on) sub_step3<T, N>(lds, b, sign, t);
  __syncthreads();
  if (on) {
    const int c = t % C, k1 = t / C;
#pragma unroll
    for (int k2 = 0; k2 < N2; k2++) lds[(k1 + N1 * k2) * P + c] = b[k2];
  }
  __syncthreads();
}

template <typename T, int N>
__global__ __launch_bounds__((kThreads * PlaneFits<T, N>::G)) void k_fft_plane_r2c(PlaneArgs<T> a) {
  constexpr bool fits = PlaneFits<T, N>::value;
  constexpr int G = PlaneFits<T, N>::G;
  constexpr int N1 = Shape<N>::N1, N2 = Shape<N>::N2, C = kThreads / N1, NH = N / 2 + 1, P = C + 1;
  if (!a.active[blockIdx.z]) return;
  __shared__ cx<T> plane[fits ? N * NH : 1];
  __shared__ cx<T> scratch[fits ? G * N * P : 1];
  if constexpr (fits) {
    const int f = blockIdx.y, iz = blockIdx.x, grp = threadIdx.x / kThreads, t = threadIdx.x % kThreads;
    cx<T>* lds = scratch + grp * N * P;
    const T* src = a.r[f] + (long)blockIdx.z * a.rslot + (long)iz * N * N;
    cx<T>* dst = a.c[f] + (long)blockIdx.z * a.cslot + (long)iz * N * a.pitch;
    const T sc = (a.rscale && f == 0) ? (T)a.rscale[blockIdx.z] : T(1);
    constexpr int NB = (N / 2 + C - 1) / C;  // row-pair batches
    for (int j = 0; j < (NB + G - 1) / G; j++) {  // x: pairs of rows
      const int p0 = (j * G + grp) * C;
      const bool on = p0 < N / 2;
      if (on)
        for (int e = t; e < C * N; e += kThreads) {
          const int c = e / N, i = e - c * N, pr = p0 + c;
          cx<T> v{T(0), T(0)};
          if (pr < N / 2) {
            const T* ra = src + (long)(2 * pr) * N;
            v = cx<T>{sc * ra[i], sc * ra[N + i]};
          }
          lds[i * P + c] = v;
        }
      __syncthreads();
      plane_x_fft<T, N>(lds, a.twx, -1, t, on);
      if (on)
        for (int e = t; e < C * NH; e += kThreads) {
          const int c = e / NH, k = e - c * NH, pr = p0 + c;
          if (pr >= N / 2) continue;
          const cx<T> z = lds[k * P + c], zm = lds[((N - k) % N) * P + c];
          const cx<T> sm = cx<T>{z.x + zm.x, z.y - zm.y};  // z + conj(zm)
          const cx<T> df = cx<T>{z.x - zm.x, z.y + zm.y};  // z - conj(zm)
          plane[(2 * pr) * NH + k] = cx<T>{T(0.5) * sm.x, T(0.5) * sm.y};
          plane[(2 * pr + 1) * NH + k] = cx<T>{T(0.5) * df.y, T(-0.5) * df.x};
        }
      __syncthreads();
    }
    constexpr int NK = (NH + C - 1) / C;  // column batches
    for (int j = 0; j < (NK + G - 1) / G; j++) {  // y: columns of the half spectrum
      const int k0 = (j * G + grp) * C;
      const bool on = k0 < NH;
      auto load = [&](int c, int i) -> cx<T> {
        const int k = k0 + c;
        return k < NH ? plane[i * NH + k] : cx<T>{T(0), T(0)};
      };
      sub_step1<T, N>(lds, a.twy, -1, load, t, on);
      __syncthreads();
      if (on) {
        cx<T> b[N2];
        sub_step3<T, N>(lds, b, -1, t);
        const int c = t % C, k1 = t / C, k = k0 + c;
        if (k < NH) {
#pragma unroll
          for (int k2 = 0; k2 < N2; k2++) dst[k + (long)(k1 + N1 * k2) * a.pitch] = b[k2];
        }
      }
      __syncthreads();
    }
  }
}

// ---- the same forward x-y plane transform in place in one plane buffer -----
// LDS holds only the plane, as N rows of Q = N/2 + 1 complex values, so two
// work-groups fit per CU (FP32 at N = 128: 66.5 KB) and one group's HBM loads
// overlap the other's transforms:
//  1. row pairs packed as z_p = row(2p) + i row(2p+1) into plane[i Q + p]
//     (pair-major: consecutive pairs on consecutive banks);
//  2. x FFT of every pair in place: step 1 writes back exactly the positions
//     the thread read (elements = n2 mod N2), step 3 after a barrier;
//  3. y FFT of N/2 columns: column k (1 <= k < N/2) of the half spectrum is
//     formed on load from Z_p[k] and Z_p[N-k] (X[2p] = (Z_p[k] + conj Z_p[N-k])/2,
//     X[2p+1] = (Z_p[k] - conj Z_p[N-k])/(2i)); the real columns k = 0 and N/2
//     travel packed as one complex column W = X[.][0] + i X[.][N/2] and are
//     separated after the transform (Y0 = (W + conj W_-)/2, Y_N/2 =
//     (W - conj W_-)/(2i)).  Each column works in place in the 2 (N/2) slots
//     of the two plane rows it reads.
// The work-group (G x kThreads threads) transforms kThreads G / N1 lines at a time.
template <typename T, int N, int G> struct PlaneIP {
  static constexpr int H = N / 2, Q = H + 1;
  static constexpr long bytes = (long)N * Q * (long)sizeof(cx<T>);
  static constexpr bool value = bytes <= 80 * 1024;  // two work-groups per CU
};
template <typename T, int N, int G, int W = 1>  // W: waves per SIMD the registers must allow
__global__ __launch_bounds__(kThreads * G, W) void k_fft_plane_r2c_ip(PlaneArgs<T> a) {
  constexpr bool fits = PlaneIP<T, N, G>::value;
  constexpr int N1 = Shape<N>::N1, N2 = Shape<N>::N2, H = N / 2, Q = H + 1;
  constexpr int CW = kThreads * G / N1;  // lines (pairs or columns) per round
  constexpr int R = (H + CW - 1) / CW;   // rounds
  if (!a.active[blockIdx.z]) return;
  __shared__ cx<T> plane[fits ? N * Q : 1];
  if constexpr (fits) {
    const int f = blockIdx.y, iz = blockIdx.x;
    const T* src = a.r[f] + (long)blockIdx.z * a.rslot + (long)iz * N * N;
    cx<T>* dst = a.c[f] + (long)blockIdx.z * a.cslot + (long)iz * N * a.pitch;
    cx<T>* wdst = a.c[f] + (long)blockIdx.z * a.cslot + a.wofs + (long)iz * N;  // packed: W of this plane
    const T sc = (a.rscale && f == 0) ? (T)a.rscale[blockIdx.z] : T(1);
    // 1. pairs of real rows, packed: a lane takes element i of a row pair and
    //    loads it from both rows (each load coalesced along the row), every
    //    load of the thread issued before its first LDS store; consecutive
    //    lanes store at the pitch Q (2Q = 2 mod 64 words: no bank conflicts)
    {
      constexpr int NT = kThreads * G, NE = N * H, NL = (NE + NT - 1) / NT;
      T ra[NL], rb[NL];
#pragma unroll
      for (int k = 0; k < NL; k++) {
        const int e = threadIdx.x + k * NT;
        if (NE % NT == 0 || e < NE) {
          const int p = e / N, i = e - p * N;
          ra[k] = sc * src[(long)(2 * p) * N + i];
          rb[k] = sc * src[(long)(2 * p + 1) * N + i];
        }
      }
#pragma unroll
      for (int k = 0; k < NL; k++) {
        const int e = threadIdx.x + k * NT;
        if (NE % NT == 0 || e < NE) {
          const int p = e / N, i = e - p * N;
          plane[i * Q + p] = cx<T>{ra[k], rb[k]};
        }
      }
    }
    __syncthreads();
    // the whole work-group is one group of CW lines (pairs or columns): the
    // 32 lanes of a wave half then touch 32 consecutive lines with the same
    // element index, at the pitch Q (2Q = 2 mod 64 words), so no LDS access
    // of the transforms has a bank conflict
    const int tt = threadIdx.x, c = tt % CW;
    // the N1-point first steps of all H lines run at once (NT / N2 lines per
    // round keep every thread busy); the N2-point second steps then run in R
    // rounds of CW lines, which touch disjoint lines
    constexpr int CW1 = kThreads * G / N2;
    static_assert(!fits || CW1 >= H, "the first steps of all lines in one round");
    // 2. x transforms, in place in pair column p
    {
      const int p = tt % CW1, n2 = tt / CW1;
      if (p < H && n2 < N2) {
        cx<T> v[N1];
#pragma unroll
        for (int n1 = 0; n1 < N1; n1++) v[n1] = plane[(N2 * n1 + n2) * Q + p];
        RegFFT<T, N1>::run(v, -1);
#pragma unroll
        for (int k1 = 0; k1 < N1; k1++) {
          cx<T> w = v[k1];
          if (k1 > 0) w = cmul(w, a.twx[n2 * k1]);
          plane[(k1 * N2 + n2) * Q + p] = w;  // the positions this thread read
        }
      }
    }
    __syncthreads();
    for (int j = 0; j < R; j++) {
      const int p = j * CW + c;
      const bool on = p < H;
      const int k1 = tt / CW;
      cx<T> b[N2];
      if (on) {
#pragma unroll
        for (int m = 0; m < N2; m++) b[m] = plane[(k1 * N2 + m) * Q + p];
        RegFFT<T, N2>::run(b, -1);
      }
      __syncthreads();
      if (on) {
#pragma unroll
        for (int k2 = 0; k2 < N2; k2++) plane[(k1 + N1 * k2) * Q + p] = b[k2];
      }
    }
    __syncthreads();
    // 3. y transforms of the half-spectrum columns; column kk owns plane rows
    //    kk and N - kk (kk = 0: rows 0 and H): element m < H at row ra, else rb
    {
      const int kk = tt % CW1, n2 = tt / CW1;
      const bool s1 = kk < H && n2 < N2;
      const int ra = kk, rb = kk == 0 ? H : N - kk;
      cx<T> v[N1];
      if (s1) {
#pragma unroll
        for (int n1 = 0; n1 < N1; n1++) {
          const int y = N2 * n1 + n2, pp = y >> 1;
          const cx<T> za = plane[ra * Q + pp], zb = plane[rb * Q + pp];
          cx<T> x;
          if (kk == 0) {  // W[y] = X[y][0] + i X[y][H]
            x = (y & 1) ? cx<T>{za.y, zb.y} : cx<T>{za.x, zb.x};
          } else if (y & 1) {  // (Z[k] - conj Z[N-k]) / (2i)
            x = cx<T>{T(0.5) * (za.y + zb.y), T(-0.5) * (za.x - zb.x)};
          } else {  // (Z[k] + conj Z[N-k]) / 2
            x = cx<T>{T(0.5) * (za.x + zb.x), T(0.5) * (za.y - zb.y)};
          }
          v[n1] = x;
        }
        RegFFT<T, N1>::run(v, -1);
#pragma unroll
        for (int k1 = 0; k1 < N1; k1++)
          if (k1 > 0) v[k1] = cmul(v[k1], a.twy[n2 * k1]);
      }
      __syncthreads();  // every column's Z values are read before any is overwritten
      if (s1) {
#pragma unroll
        for (int k1 = 0; k1 < N1; k1++) {
          const int m = k1 * N2 + n2;
          plane[(m < H ? ra : rb) * Q + (m < H ? m : m - H)] = v[k1];
        }
      }
    }
    __syncthreads();
    for (int j = 0; j < R; j++) {
      const int kk = j * CW + c;
      const bool on = kk < H;
      const int ra = kk, rb = kk == 0 ? H : N - kk;
      const int k1 = tt / CW;
      cx<T> b[N2];
      if (on) {
#pragma unroll
        for (int m2 = 0; m2 < N2; m2++) {
          const int m = k1 * N2 + m2;
          b[m2] = plane[(m < H ? ra : rb) * Q + (m < H ? m : m - H)];
        }
        RegFFT<T, N2>::run(b, -1);
        if (kk > 0) {
#pragma unroll
          for (int k2 = 0; k2 < N2; k2++) dst[kk + (long)(k1 + N1 * k2) * a.pitch] = b[k2];
        } else if (a.packed) {  // packed layout: W kept as it is, in its own block
#pragma unroll
          for (int k2 = 0; k2 < N2; k2++) wdst[k1 + N1 * k2] = b[k2];
        }
      }
      if (j == 0 && !a.packed) {  // column 0, the packed W, lives in round 0
        __syncthreads();
        if (on && kk == 0) {  // W[ky] into its own slots
#pragma unroll
          for (int k2 = 0; k2 < N2; k2++) {
            const int m = k1 + N1 * k2;
            plane[(m < H ? ra : rb) * Q + (m < H ? m : m - H)] = b[k2];
          }
        }
        __syncthreads();
        for (int ky = tt; ky < N; ky += kThreads * G) {  // separate the packed column: Y0 and Y_H
          const int km = (N - ky) % N;
          const cx<T> w = plane[(ky < H ? 0 : H) * Q + (ky < H ? ky : ky - H)];
          const cx<T> wm = plane[(km < H ? 0 : H) * Q + (km < H ? km : km - H)];
          dst[(long)ky * a.pitch] = cx<T>{T(0.5) * (w.x + wm.x), T(0.5) * (w.y - wm.y)};
          dst[H + (long)ky * a.pitch] = cx<T>{T(0.5) * (w.y + wm.y), T(-0.5) * (w.x - wm.x)};
        }
      }
    }
  }
}

// ---- inverse y and x transforms of one z plane, in place ------------------
// The mirror of k_fft_plane_r2c_ip for the half spectra that the z
// convolutions leave: half spectrum -> inverse y FFT of every column ->
// inverse x FFT of row pairs -> real rows, with one HBM read and one write
// (the separate passes, k_fft_strided + k_fft_x_c2r, make two of each).
// LDS holds the plane as N rows of Q = N/2 + 1 complex values:
//  1. column kk (1 <= kk < H) of the half spectrum goes to plane rows kk
//     (elements ky < H, slot ky) and N - kk (slot ky - H); the columns kx = 0
//     and H, whose inverse a c2r transform keeps only the real part of, travel
//     as one complex column W = S0h + i SHh of their Hermitian parts
//     (S0h[ky] = (S0[ky] + conj S0[N-ky])/2): the inverse of W is then
//     Re x0 + i Re xH, in rows 0 and H;
//  2. inverse y FFT of each column in place in its two rows;
//  3. the rows 2p and 2p+1 of a column come out of the y FFT in lanes l and
//     l ^ 32 of one wave, which swap them and write the packed x input of row
//     pair p into the column's own rows: z_p[kk] = x[2p][kk] + i x[2p+1][kk]
//     at row kk, z_p[N-kk] = conj x[2p][kk] + i conj x[2p+1][kk] at row
//     N - kk, slot p (column 0: z_p[0] = Re w[2p] + i Re w[2p+1] at row 0,
//     z_p[H] = Im w[2p] + i Im w[2p+1] at row H);
//  4. inverse x FFT of each pair column p in place;
//  5. rows 2p and 2p+1 = Re, Im of z_p, stored as aff0 + scale aff1 x.
// The work-group (G x kThreads threads) works on 32 lines at a time, so the
// lines of a wave are 32 columns in each of its halves (step 3 needs that):
// G = N1 / 8 (N >= 24), and the plane fits twice per CU (PlaneIP).
template <typename T, int N> struct PlaneInv {
  static constexpr int G = Shape<N>::N1 >= 8 ? Shape<N>::N1 / 8 : 1;
  static constexpr bool value = PlaneIP<T, N, G>::value && kThreads * G / Shape<N>::N1 == 32;
};
template <typename T, int N>
__global__ __launch_bounds__((kThreads * PlaneInv<T, N>::G)) void k_fft_plane_c2r_ip(PlaneArgs<T> a) {
  constexpr bool fits = PlaneInv<T, N>::value;
  constexpr int G = PlaneInv<T, N>::G;
  constexpr int N1 = Shape<N>::N1, N2 = Shape<N>::N2, H = N / 2, Q = H + 1;
  constexpr int NT = kThreads * G, CW = NT / N1;  // lines per round
  constexpr int R = (H + CW - 1) / CW;  // rounds
  if (!a.active[blockIdx.z]) return;
  __shared__ cx<T> plane[fits ? N * Q : 1];
  if constexpr (fits) {
    const int f = blockIdx.y, iz = blockIdx.x;
    const cx<T>* src = a.c[f] + (long)blockIdx.z * a.cslot + (long)iz * N * a.pitch;
    const cx<T>* wsrc = a.c[f] + (long)blockIdx.z * a.cslot + a.wofs + (long)iz * N;  // packed: W of this plane
    T* dst = a.r[f] + (long)blockIdx.z * a.rslot + (long)iz * N * N;
    const int tt = threadIdx.x;
    // element m of column kk: plane row (m < H ? kk : kk == 0 ? H : N - kk), slot m mod H
    auto pos = [](int kk, int m) {
      const int row = m < H ? kk : (kk == 0 ? H : N - kk);
      return row * Q + (m < H ? m : m - H);
    };
    // 1. a lane per element (ky, kk), consecutive lanes along a half-spectrum
    //    row, and lane ky < N for element ky of the packed column W; every
    //    load issued before the first LDS store
    {
      constexpr int NE = N * H, NL = (NE + NT - 1) / NT;
      static_assert(N <= NT, "one lane per element of W");
      cx<T> w[4];
      const int km = (N - tt) % N;
      if (tt < N) {
        w[0] = a.packed ? wsrc[tt] : src[(long)tt * a.pitch];
        if (!a.packed) {
          w[1] = src[(long)km * a.pitch];
          w[2] = src[(long)tt * a.pitch + H]; w[3] = src[(long)km * a.pitch + H];
        }
      }
      cx<T> v[NL];
#pragma unroll
      for (int k = 0; k < NL; k++) {
        const int e = tt + k * NT;
        if (NE % NT == 0 || e < NE) {
          const int ky = e / H, kk = e - ky * H;
          if (kk > 0) v[k] = src[(long)ky * a.pitch + kk];
        }
      }
#pragma unroll
      for (int k = 0; k < NL; k++) {
        const int e = tt + k * NT;
        if (NE % NT == 0 || e < NE) {
          const int ky = e / H, kk = e - ky * H;
          if (kk > 0) plane[pos(kk, ky)] = v[k];
        }
      }
      if (tt < N) {
        if (a.packed) {  // W = S0 + i SH as the packed z convolution left it
          plane[pos(0, tt)] = w[0];
        } else {
          const cx<T> s0h{T(0.5) * (w[0].x + w[1].x), T(0.5) * (w[0].y - w[1].y)};
          const cx<T> shh{T(0.5) * (w[2].x + w[3].x), T(0.5) * (w[2].y - w[3].y)};
          plane[pos(0, tt)] = cx<T>{s0h.x - shh.y, s0h.y + shh.x};
        }
      }
    }
    __syncthreads();
    const int c = tt % CW;
    // the N1-point first steps of all H lines run at once (NT / N2 lines per
    // round keep every thread busy); the N2-point second steps then run in
    // rounds of CW = 32 lines
    constexpr int CW1 = NT / N2, R1 = (H + CW1 - 1) / CW1;
    // 2-3. inverse y transforms; the rounds of the second step touch disjoint
    //      rows, so only the steps inside a round are separated by barriers
    for (int j = 0; j < R1; j++) {
      const int kk = j * CW1 + tt % CW1, n2 = tt / CW1;
      if (kk < H && n2 < N2) {
        cx<T> v[N1];
#pragma unroll
        for (int n1 = 0; n1 < N1; n1++) v[n1] = plane[pos(kk, N2 * n1 + n2)];
        RegFFT<T, N1>::run(v, +1);
#pragma unroll
        for (int k1 = 0; k1 < N1; k1++) {
          cx<T> w = v[k1];
          if (k1 > 0) w = cmul(w, conjg(a.twy[n2 * k1]));
          plane[pos(kk, k1 * N2 + n2)] = w;  // the positions this thread read
        }
      }
    }
    __syncthreads();
    for (int j = 0; j < R; j++) {
      const int kk = j * CW + c;
      const bool on = kk < H;
      const int k1 = tt / CW;
      cx<T> b[N2];
      if (on) {
#pragma unroll
        for (int m2 = 0; m2 < N2; m2++) b[m2] = plane[pos(kk, k1 * N2 + m2)];
        RegFFT<T, N2>::run(b, +1);  // b[k2] = x[k1 + N1 k2]
      } else {
#pragma unroll
        for (int m2 = 0; m2 < N2; m2++) b[m2] = cx<T>{T(0), T(0)};
      }
      // lane l ^ 32 holds the same column at k1 ^ 1: rows 2p and 2p+1
      cx<T> o[N2];
#pragma unroll
      for (int k2 = 0; k2 < N2; k2++) o[k2] = cx<T>{__shfl_xor(b[k2].x, 32, 64), __shfl_xor(b[k2].y, 32, 64)};
      __syncthreads();  // every element of the column is read before its slots are rewritten
      if (on) {
        const bool odd = k1 & 1;
        const int ra = kk, rb = kk == 0 ? H : N - kk;
#pragma unroll
        for (int k2 = 0; k2 < N2; k2++) {
          const int p = (k1 >> 1) + (N1 / 2) * k2;
          const cx<T> xa = odd ? o[k2] : b[k2], xb = odd ? b[k2] : o[k2];  // x[2p], x[2p+1]
          cx<T> z;
          if (kk == 0) z = odd ? cx<T>{xa.y, xb.y} : cx<T>{xa.x, xb.x};
          else z = odd ? cx<T>{xa.x + xb.y, xb.x - xa.y} : cx<T>{xa.x - xb.y, xa.y + xb.x};
          plane[(odd ? rb : ra) * Q + p] = z;
        }
      }
    }
    __syncthreads();
    // 4. inverse x transforms, in place in pair column p
    for (int j = 0; j < R1; j++) {
      const int p = j * CW1 + tt % CW1, n2 = tt / CW1;
      if (p < H && n2 < N2) {
        cx<T> v[N1];
#pragma unroll
        for (int n1 = 0; n1 < N1; n1++) v[n1] = plane[(N2 * n1 + n2) * Q + p];
        RegFFT<T, N1>::run(v, +1);
#pragma unroll
        for (int k1 = 0; k1 < N1; k1++) {
          cx<T> w = v[k1];
          if (k1 > 0) w = cmul(w, conjg(a.twx[n2 * k1]));
          plane[(k1 * N2 + n2) * Q + p] = w;
        }
      }
    }
    __syncthreads();
    for (int j = 0; j < R; j++) {
      const int p = j * CW + c;
      const bool on = p < H;
      const int k1 = tt / CW;
      cx<T> b[N2];
      if (on) {
#pragma unroll
        for (int m = 0; m < N2; m++) b[m] = plane[(k1 * N2 + m) * Q + p];
        RegFFT<T, N2>::run(b, +1);
      }
      __syncthreads();
      if (on) {
#pragma unroll
        for (int k2 = 0; k2 < N2; k2++) plane[(k1 + N1 * k2) * Q + p] = b[k2];
      }
    }
    __syncthreads();
    // 5. a lane per element i of a row pair: coalesced row stores
    {
      const T off = a.aff0[f], sc = a.scale * a.aff1[f];
      constexpr int NE = N * H;
      for (int e = tt; e < NE; e += NT) {
        const int p = e / N, i = e - p * N;
        const cx<T> z = plane[i * Q + p];
        dst[(long)(2 * p) * N + i] = off + sc * z.x;
        dst[(long)(2 * p + 1) * N + i] = off + sc * z.y;
      }
    }
  }
}

}  // namespace ep
