// C-ABI of the gfx950 Orsay-Trento solver (declared in include/effpot.h).
//
// Host orchestration only: every array operation below is a HIP kernel from
// fft.hpp / stencil.hpp / misc.hpp launched on the context's stream.  The
// psi / H psi history of the integrators is a per-slot ring of buffer indices
// kept in device memory and rotated by a one-thread-per-slot kernel, so that
// slots can run different integrator phases (RK starter vs predictor-
// corrector) and no host round trip is needed between steps.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/effpot.h"
#include "common.hpp"
#include "fft.hpp"
#include "snapshot.hpp"
#include "plus.hpp"
#include "misc.hpp"
#include "stencil.hpp"
#include "state.hpp"
#include "motion.hpp"
#include "noalloc.hpp"

using namespace ep;

namespace {

__global__ void k_ring(int* ring, const int* active, int B, int op) {
  int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= B || !active[s]) return;
  ring_apply(ring + 8 * s, op);
}

// Pointer of ring role r of slot s: all + (s*nbuf + ring[s][r]) * N.
template <typename T> struct Ring {
  cx<T>* psi;  // [B][5][N]
  cx<T>* hp;   // [B][2][N]
  const int* ring;
  long N;
  EP_DEV cx<T>* p(int s, int r) const { return psi + ((long)s * 5 + ring[8 * s + r]) * N; }
  EP_DEV cx<T>* h(int s, int r) const { return hp + ((long)s * 2 + ring[8 * s + r]) * N; }
};

// ---- kernels that need the ring --------------------------------------------
template <typename T, typename V, typename P, int TX, int MODE>  // MODE 0 predict, 1 correct, 2 rk stage
__global__ __launch_bounds__(kThreads) void k_step(StepArgs<T, V, P> a, Ring<T> rg) {
  const int s = blockIdx.z;
  if (!a.active[s]) return;
  // resolve the roles of this slot into the argument block, then run the
  // shared kernel body (the StepArgs pointers are slot-0 based; offsets are
  // re-applied by the body, so pass slot-relative pointers minus the offset).
  const long so = (long)s * a.slot;
  int r8[8];
#pragma unroll
  for (int k = 0; k < 8; k++) r8[k] = rg.ring[8 * s + k];
  // the corrector runs before the ring is rotated: it sees the roles that
  // RING_AFTER_PRED gives (the rotation is applied once per step by the
  // norm reduction that follows)
  if constexpr (MODE == 1) ring_apply(r8, RING_AFTER_PRED);
  auto PP = [&](int r) { return reinterpret_cast<V*>(rg.psi + ((long)s * 5 + r8[r]) * rg.N) - so; };
  auto H = [&](int r) { return reinterpret_cast<V*>(rg.hp + ((long)s * 2 + r8[r]) * rg.N) - so; };
  StepArgs<T, V, P> b = a;
  b.cur = PP(R_CUR);
  b.out = MODE == 1 ? PP(R_SPARE) : PP(R_OUT);
  b.o1 = PP(R_O1); b.o2 = PP(R_O2); b.o3 = PP(R_O3);
  b.h1 = H(R_H1); b.h2 = H(R_H2);
  if constexpr (MODE == 0) k_pc_predict_body<T, V, P, TX>(b);
  else if constexpr (MODE == 1) k_pc_correct_body<T, V, P, TX>(b);
  else k_rk_stage_body<T, V, P, TX>(b);
}

template <typename T, typename V, typename P>
__global__ void k_renorm_ring(Ring<T> rg, P* den, const double* scale, const int* active, long N) {
  const int s = blockIdx.y;
  if (!active[s]) return;
  const T f = (T)scale[s];
  V* psi = reinterpret_cast<V*>(rg.p(s, R_CUR));
  P* d = den ? den + (long)s * N : nullptr;
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < N; i += (long)gridDim.x * blockDim.x) {
    V p = f * psi[i];
    psi[i] = p;
    if (d) d[i] = (P)norm2(p);
  }
}

// partial sums of |psi_cur|^2 in the same [slot][block][2] layout as the steps
template <typename T, typename V>
__global__ __launch_bounds__(kThreads) void k_norm(Ring<T> rg, const int* active, long N, double* partial) {
  __shared__ double red[8];
  const int s = blockIdx.y;
  if (!active[s]) return;
  const V* psi = reinterpret_cast<const V*>(rg.p(s, R_CUR));
  double v[2] = {0.0, 0.0};
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < N; i += (long)gridDim.x * blockDim.x) {
    v[1] += (double)norm2(psi[i]);
  }
  block_sum<2>(v, red);
  if (threadIdx.x == 0) {
    double* q = partial + ((long)s * gridDim.x + blockIdx.x) * 2;
    q[0] = 0.0; q[1] = v[1];
  }
}

template <typename T, typename V, typename P>
__global__ void k_den_from_cur(Ring<T> rg, P* den, const int* active, long N) {
  const int s = blockIdx.y;
  if (!active[s]) return;
  const V* psi = reinterpret_cast<const V*>(rg.p(s, R_CUR));
  P* d = den + (long)s * N;
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < N; i += (long)gridDim.x * blockDim.x) {
    d[i] = (P)norm2(psi[i]);
  }
}

template <typename T, int TX, int ID>
__global__ __launch_bounds__(kThreads) void k_deriv1(Grid g, StenCoef<T> sc, int zc, const T* f, T* df, int axis) {
  __shared__ T lds[2 * Tile<TX>::PL];
  zmarch<T, T, TX, ID == 1, ID == 2>(f, g, sc, TilePos<TX>(g, zc), lds, [](unsigned) { return NoPre{}; },
    [&](int x, int y, int z, bool in, T, T(&d1)[3], T(&d2)[3], const NoPre&) {
      if (!in) return;
      df[g.idx(x, y, z)] = ID == 1 ? d1[axis] : d2[axis];
    });
}

// ---------------------------------------------------------------------------
template <typename T> struct State {
  cx<T>* psi = nullptr;  // [B][5][N]
  cx<T>* hp = nullptr;   // [B][2][N]
  cx<T>* pc = nullptr;   // [B][N]
  cx<T>* q = nullptr;
  T *den = nullptr, *pot4 = nullptr, *uimp = nullptr, *dencg = nullptr, *delj4 = nullptr, *denalf = nullptr;
  T *ua = nullptr, *ualphas = nullptr, *dx = nullptr, *dy = nullptr, *dz = nullptr, *sto1 = nullptr;
  T *intx = nullptr, *inty = nullptr, *intz = nullptr, *div = nullptr, *src[4] = {nullptr, nullptr, nullptr, nullptr};
  cx<T>* spec[8] = {};  // each [B][M]
  T* kern[3] = {};      // wcgk, fvlj4, kalfs: [M]
  T* kw[3] = {};        // the same at kx = 0 and nx/2 as pairs ((K0+KH)/2, (K0-KH)/2): [nz][ny][2]
  cx<T>* tw[3] = {};    // twiddles per axis
  StenCoef<T> sc;
  int* ring = nullptr;  // [B][8]
  T* corr = nullptr;    // [B][N] FP32 fields only: defect correction added to pot4
};

}  // namespace

struct effpot_ctx {
  effpot_params p;
  int dev = 0;
  hipStream_t st = nullptr;
  int nx, ny, nz, nxh, B;
  int nxp;  // half-spectrum row pitch: nx/2+1 rounded up to a multiple of 16
  long N, M;
  double hx, hy, hz, dxyz;
  int npot;
  double delinter;
  State<double> s64;
  State<float> s32;
  bool mixed = false;
  bool cplx = true;  // psi stored as complex (else real: N values per buffer)
  double* imptab = nullptr;  // [B][npot+2]
  int* d_active = nullptr;   // current mask
  int* d_ones = nullptr;
  int* d_mask_tmp = nullptr;
  std::vector<int> h_active;
  std::vector<int> have_imp;
  PlusSlot* d_plus = nullptr;     // [B] X+ potential of the uplus term
  std::vector<PlusSlot> h_plus;
  std::vector<effpot_impurity> h_imp;  // [B] isotropic impurity of each slot
  std::vector<int> h_is_state;    // [B] slot holds a P/D-state impurity
  Motion* d_motion = nullptr;     // [B] moving impurities
  std::vector<Motion> h_motion;
  int any_motion = 0;             // some slot's impurity moves
  int rt = 0;                     // real-time propagation
  InvarDyn* d_dyn = nullptr;      // [B] P/D-state invar dynamics
  std::vector<InvarDyn> h_dyn;
  std::vector<StateArgs> h_sargs; // [B] the P/D-state set-up of each slot
  std::vector<double> h_sals;     // [B] its spin-orbit splitting
  int any_dyn = 0;
  double* d_dynpart = nullptr;    // [B][kStateBlocks][kMomD] moment partials
  int plane = 3;                  // forward x and y FFT passes fused per z plane (plane_mode; EFFPOT_PLANE)
  bool plane_fixed = false;       // EFFPOT_PLANE given: the autotuner keeps it
  int iplane = 1;                 // inverse y and x passes fused per z plane where it fits (EFFPOT_IPLANE)
  int packed = 1;                 // packed half spectra inside poten where both plane kernels run (EFFPOT_PACKED)
  bool packed_fixed = false;      // EFFPOT_PACKED given: the autotuner keeps it
  int zconv_th = 256;             // threads per z-convolution group: 256 or 512 (EFFPOT_ZCONV_TH)
  bool zconv_fixed = false;       // EFFPOT_ZCONV_TH given: the autotuner keeps it
  bool iplane_fixed = false;      // EFFPOT_IPLANE given: the autotuner keeps it
  std::vector<double> h_eso;      // [B] spin-orbit energy of a P/D-state impurity
  double* d_stab = nullptr;       // [B][3][npot+2] P/D-state potential tables (on first use)
  double* d_smom = nullptr;       // [kStateBlocks][kMomD] moment partials
  double* d_partial = nullptr;
  int nblk_max = 0;
  double* d_red = nullptr;    // [B][16]
  double* d_scale = nullptr;  // [B]
  double* d_dscale = nullptr; // [B] factor taking a step's density to n4 atoms
  double* d_n4 = nullptr;     // [B] atoms each slot is normalised to
  double* d_dt = nullptr;     // [B] time step each slot is advanced with
  std::vector<double> h_dt;   // per-slot override (0: the solve options' dt)
  std::vector<double> dt_dev; // what d_dt holds
  std::vector<double> h_n4;
  int* d_bad = nullptr;
  int* d_mask = nullptr;      // [7][B] per-purpose activity masks of effpot_tick
  int* h_mask = nullptr;      // pinned staging for d_mask
  std::vector<int> mask_dev;  // what d_mask holds (-1: unknown)
  std::vector<double> h_relprev;   // |dE/E| per step at each slot's previous check
  std::vector<double> h_reltrend;  // its smoothed change per check
  bool corr_den = true;       // corrector stores the density (false when a renormalisation follows)
  double* h_pin = nullptr;    // pinned host buffer
  int tx = 64, zc = 16;
  bool graphs = true;               // replay captured PC steps (EFFPOT_GRAPHS=0 disables)
  bool tune_log = false;            // EFFPOT_TUNE_LOG=1 prints the autotuner's timings
  hipGraphExec_t pc_exec[2] = {nullptr, nullptr};  // [0] FP64, [1] FP32
  hipGraphExec_t pc_exec2[2] = {nullptr, nullptr}; // the same as two concurrent slot groups
  hipGraphExec_t pck_exec[2] = {nullptr, nullptr};  // pck_k[] consecutive PC steps in one graph
  hipGraphExec_t pck_exec2[2] = {nullptr, nullptr}; // the same, two concurrent slot groups
  int pck_k[2] = {0, 0}, pck_k2[2] = {0, 0};       // steps captured in pck_exec / pck_exec2
  hipStream_t st2 = nullptr;                       // second stream of the two-group graph
  hipEvent_t ev_fork = nullptr, ev_join = nullptr;
  bool split = true;                               // two slot groups when B >= 2 (EFFPOT_SPLIT=0 disables)
  std::vector<void*> allocs;
  std::string err;

  int fail(int code, const std::string& m) {
    err = m;
    return code;
  }
};

#define HIPCHK(x)                                                                        \
  do {                                                                                   \
    hipError_t e_ = (x);                                                                 \
    if (e_ != hipSuccess)                                                                \
      return ctx->fail(EFFPOT_ERR_DEVICE, std::string(#x) + ": " + hipGetErrorString(e_)); \
  } while (0)
#define LAUNCHCHK()                                                                                \
  do {                                                                                             \
    hipError_t e_ = hipGetLastError();                                                             \
    if (e_ != hipSuccess) return ctx->fail(EFFPOT_ERR_DEVICE, std::string("launch: ") + hipGetErrorString(e_)); \
  } while (0)

namespace {

// Device allocation and free, context set-up and tear-down, and every graph
// capture (with its instantiation) hold this lock: HIP treats allocation
// during another thread's stream capture as unsafe (the capture is
// invalidated and the call fails), so a server that creates contexts while
// other contexts capture their step graphs needs them serialised.
std::recursive_mutex g_api_mu;

// Host <-> device copies ordered on the context's own stream (never the null
// stream, which would synchronise with, and break, another thread's capture).
hipError_t copy_sync(effpot_ctx* ctx, void* dst, const void* src, size_t bytes, hipMemcpyKind kind) {
  hipError_t e = hipMemcpyAsync(dst, src, bytes, kind, ctx->st);
  if (e != hipSuccess) return e;
  return hipStreamSynchronize(ctx->st);
}

int tstgrid(int n) {
  if (n < 16 || n > 256) return 1;
  int m = n;
  while (m % 2 == 0) m /= 2;
  while (m % 3 == 0) m /= 3;
  while (m % 5 == 0) m /= 5;
  return m != 1;
}

void fd_coeffs(int id, double* c) {  // same closed forms as the Bickley table rows (mates.f:1117-1125, 1202-1210)
  const double p2 = 518400.0;
  static const double fact[13] = {1, 1, 2, 6, 24, 120, 720, 5040, 40320, 362880, 3628800, 39916800, 479001600};
  for (int m = -6; m <= 6; m++) {
    int am = m < 0 ? -m : m;
    if (m == 0) { c[6] = id == 1 ? 0.0 : -5369.0 / 1800.0; continue; }
    double den = fact[6 - am] * fact[6 + am];
    double sgn = (am % 2 == 1) ? 1.0 : -1.0;
    if (id == 1) c[m + 6] = (m > 0 ? sgn : -sgn) * (p2 / ((double)am * den));
    else c[m + 6] = sgn * (2.0 * p2 / ((double)(am * am) * den));
  }
}

template <typename T> int dalloc(effpot_ctx* ctx, T** p, size_t count) {
  std::lock_guard<std::recursive_mutex> lk(g_api_mu);
  void* v = nullptr;
  if (count == 0) count = 1;
  hipError_t e = hipMalloc(&v, count * sizeof(T));
  if (e != hipSuccess) return ctx->fail(EFFPOT_ERR_NOMEM, std::string("hipMalloc: ") + hipGetErrorString(e));
  ctx->allocs.push_back(v);
  *p = (T*)v;
  return EFFPOT_OK;
}

// fields_only: the FP32 field state of a MIXED context (no wavefunction ring).
template <typename T> int alloc_state(effpot_ctx* ctx, State<T>& s, bool fields_only = false) {
  const long N = ctx->N, M = ctx->M, B = ctx->B;
  const long NP = fields_only ? 0 : N;
  int rc;
#define A_(ptr, cnt) if ((rc = dalloc(ctx, &(ptr), (size_t)(cnt))) != EFFPOT_OK) return rc
  A_(s.psi, B * 5 * NP); A_(s.hp, B * 2 * NP); A_(s.pc, B * NP); A_(s.q, B * NP);
  T** reals[] = {&s.den, &s.pot4, &s.uimp, &s.dencg, &s.delj4, &s.denalf, &s.ua, &s.ualphas, &s.dx, &s.dy,
                 &s.dz, &s.sto1, &s.intx, &s.inty, &s.intz, &s.div, &s.src[0], &s.src[1], &s.src[2], &s.src[3]};
  for (T** r : reals) A_(*r, B * N);
  if (fields_only) A_(s.corr, B * N);
  for (int j = 0; j < 8; j++) A_(s.spec[j], B * M);
  for (int j = 0; j < 3; j++) A_(s.kern[j], M);
  for (int j = 0; j < 3; j++) A_(s.kw[j], 2L * ctx->ny * ctx->nz);
  A_(s.tw[0], ctx->nx); A_(s.tw[1], ctx->ny); A_(s.tw[2], ctx->nz);
  A_(s.ring, B * 8);
#undef A_
  // zero the whole state (history, pc, q, fields)
  if (!fields_only) {
    HIPCHK(hipMemsetAsync(s.psi, 0, sizeof(cx<T>) * B * 5 * N, ctx->st));
    HIPCHK(hipMemsetAsync(s.hp, 0, sizeof(cx<T>) * B * 2 * N, ctx->st));
    HIPCHK(hipMemsetAsync(s.pc, 0, sizeof(cx<T>) * B * N, ctx->st));
    HIPCHK(hipMemsetAsync(s.q, 0, sizeof(cx<T>) * B * N, ctx->st));
  }
  for (T** r : reals) HIPCHK(hipMemsetAsync(*r, 0, sizeof(T) * B * N, ctx->st));
  if (s.corr) HIPCHK(hipMemsetAsync(s.corr, 0, sizeof(T) * B * N, ctx->st));
  // twiddles
  int ns[3] = {ctx->nx, ctx->ny, ctx->nz};
  for (int a = 0; a < 3; a++) {
    std::vector<cx<T>> tw(ns[a]);
    for (int k = 0; k < ns[a]; k++) {
      double ang = -2.0 * M_PI * (double)k / (double)ns[a];
      tw[k] = cx<T>{(T)std::cos(ang), (T)std::sin(ang)};
    }
    HIPCHK(copy_sync(ctx, s.tw[a], tw.data(), sizeof(cx<T>) * ns[a], hipMemcpyHostToDevice));
  }
  // stencil coefficients: c_m / h^id in FP64, then rounded to T
  double c1[13], c2[13];
  fd_coeffs(1, c1);
  fd_coeffs(2, c2);
  double hs[3] = {ctx->hx, ctx->hy, ctx->hz};
  for (int a = 0; a < 3; a++) {
    double s1 = 1.0 / hs[a], s2 = 1.0 / (hs[a] * hs[a]);
    for (int k = 0; k <= 6; k++) {  // c1 is odd and c2 even in the offset, exactly (fd_coeffs)
      s.sc.h1[a][k] = (T)(c1[6 + k] * s1);
      s.sc.h2[a][k] = (T)(c2[6 + k] * s2);
    }
  }
  std::vector<int> ring(B * 8);
  for (int b = 0; b < B; b++) {
    int init[8] = {0, 1, 2, 3, 4, 0, 1, 0};
    for (int k = 0; k < 8; k++) ring[b * 8 + k] = init[k];
  }
  HIPCHK(copy_sync(ctx, s.ring, ring.data(), sizeof(int) * B * 8, hipMemcpyHostToDevice));
  return EFFPOT_OK;
}

dim3 pw_grid(long n, int nslots) {
  long b = (n + 255) / 256;
  if (b > 2048) b = 2048;
  return dim3((unsigned)b, nslots);
}

// ---- FFT launchers -----------------------------------------------------------
#define FFT_SWITCH(n, M_)                                              \
  switch (n) {                                                         \
    case 16: M_(16); break; case 24: M_(24); break; case 32: M_(32); break; \
    case 48: M_(48); break; case 64: M_(64); break; case 96: M_(96); break; \
    case 128: M_(128); break; case 192: M_(192); break; case 256: M_(256); break; \
    default: return ctx->fail(EFFPOT_ERR_GRID, "unsupported FFT length (16 24 32 48 64 96 128 192 256)"); \
  }

int n1_of(int n) {
  switch (n) { case 16: return 4; case 24: case 32: case 48: case 64: return 8; default: return 16; }
}

// rscale: [slot] factor the first field is read with (or null)
template <typename T> int fft_x_r2c(effpot_ctx* ctx, State<T>& s, T* const* in, cx<T>* const* out, int nf, const int* act,
                                    const double* rscale = nullptr, const T* rmul = nullptr) {
  XArgs<T> a{};
  for (int f = 0; f < nf; f++) { a.r[f] = in[f]; a.c[f] = out[f]; }
  a.rslot = ctx->N; a.cslot = ctx->M; a.nrows = ctx->ny * ctx->nz; a.scale = T(1); a.twN = s.tw[0]; a.active = act;
  a.rscale = rscale; a.rmul = rmul;
  a.pitch = ctx->nxp;
  const int C = kThreads / n1_of(ctx->nx);
  dim3 g((ctx->ny * ctx->nz / 2 + C - 1) / C, nf, ctx->B);
#define M_(NN) hipLaunchKernelGGL((k_fft_x_r2c<T, NN>), g, dim3(kThreads), 0, ctx->st, a)
  FFT_SWITCH(ctx->nx, M_)
#undef M_
  LAUNCHCHK();
  return EFFPOT_OK;
}
template <typename T> int fft_x_c2r(effpot_ctx* ctx, State<T>& s, cx<T>* const* in, T* const* out, int nf, T scale,
                                    const int* act, const T* aff0 = nullptr, const T* aff1 = nullptr) {
  XArgs<T> a{};
  for (int f = 0; f < nf; f++) {
    a.r[f] = out[f]; a.c[f] = in[f];
    a.aff0[f] = aff0 ? aff0[f] : T(0);
    a.aff1[f] = aff1 ? aff1[f] : T(1);
  }
  a.rslot = ctx->N; a.cslot = ctx->M; a.nrows = ctx->ny * ctx->nz; a.scale = scale; a.twN = s.tw[0]; a.active = act;
  a.pitch = ctx->nxp;
  const int C = kThreads / n1_of(ctx->nx);
  dim3 g((ctx->ny * ctx->nz / 2 + C - 1) / C, nf, ctx->B);
#define M_(NN) hipLaunchKernelGGL((k_fft_x_c2r<T, NN>), g, dim3(kThreads), 0, ctx->st, a)
  FFT_SWITCH(ctx->nx, M_)
#undef M_
  LAUNCHCHK();
  return EFFPOT_OK;
}
// The fused forward x-y plane transform of square planes (ctx->plane): 1 the
// kernel with per-sub-group scratch (FP64 only up to 64^2, where larger
// planes leave one wave per SIMD), 2 / 3 the in-place kernel with 4 / 2
// sub-groups where the plane fits twice per CU (else mode 1 where it fits);
// 0 separate x and y passes.
template <typename T, int NN> bool plane_fits(int mode) {
  switch (mode) {
    case 1: return PlaneFits<T, NN>::value && !(sizeof(T) == 8 && NN > 64);
    case 2: return PlaneIP<T, NN, 4>::value;
    case 3: return PlaneIP<T, NN, 2>::value;
    case 4: return PlaneIP<T, NN, 4>::value;
  }
  return false;
}
template <typename T> int plane_mode(effpot_ctx* ctx) {
  if (!ctx->plane || ctx->nx != ctx->ny) return 0;
  const int m = ctx->plane;
  auto pick = [m](bool want, bool one) { return want ? m : (m >= 2 && one ? 1 : 0); };
  switch (ctx->nx) {
#define M_(NN) case NN: return pick(plane_fits<T, NN>(m), plane_fits<T, NN>(1));
    M_(16) M_(24) M_(32) M_(48) M_(64) M_(96) M_(128) M_(192) M_(256)
#undef M_
  }
  return 0;
}
template <typename T> bool plane_ok(effpot_ctx* ctx) { return plane_mode<T>(ctx) != 0; }
template <typename T>
PlaneArgs<T> plane_args(effpot_ctx* ctx, State<T>& s, T* const* r, cx<T>* const* c, int nf, const int* act) {
  PlaneArgs<T> a{};
  for (int f = 0; f < nf; f++) { a.r[f] = r[f]; a.c[f] = c[f]; }
  a.rslot = ctx->N; a.cslot = ctx->M; a.pitch = ctx->nxp;
  a.twx = s.tw[0]; a.twy = s.tw[1]; a.active = act;
  return a;
}
// The packed layout (poten with both in-place plane kernels): rows of nx/2
// (column 0 unused), then W = S(kx=0) + i S(kx=nx/2) for every (ky, z), ky
// fastest, at this offset of the slot's spectrum buffer: rows of nx/2 leave at
// least ny nz of the nxp ny nz elements free, since nxp >= nx/2 + 1.
long packed_wofs(effpot_ctx* ctx) { return (long)(ctx->nxh - 1) * ctx->ny * ctx->nz; }
// real fields -> half spectra transformed along x and y (one launch per plane);
// packed: rows of nx/2 with W = S(0) + i S(nx/2) in column 0 (in-place kernel only)
template <typename T> int fft_plane_fwd(effpot_ctx* ctx, State<T>& s, T* const* in, cx<T>* const* out, int nf,
                                        const int* act, bool packed = false, const double* rscale = nullptr) {
  PlaneArgs<T> a = plane_args(ctx, s, in, out, nf, act);
  a.rscale = rscale;
  if (packed) { a.packed = 1; a.pitch = ctx->nxh - 1; a.wofs = packed_wofs(ctx); }
  dim3 g(ctx->nz, nf, ctx->B);
  const int mode = plane_mode<T>(ctx);
  if (mode == 2) {
#define M_(NN) hipLaunchKernelGGL((k_fft_plane_r2c_ip<T, NN, 4>), g, dim3(kThreads * 4), 0, ctx->st, a)
    FFT_SWITCH(ctx->nx, M_)
#undef M_
  } else if (mode == 3) {
#define M_(NN) hipLaunchKernelGGL((k_fft_plane_r2c_ip<T, NN, 2>), g, dim3(kThreads * 2), 0, ctx->st, a)
    FFT_SWITCH(ctx->nx, M_)
#undef M_
  } else if (mode == 4) {  // 4 sub-groups held to 64 VGPRs: two 1024-thread groups per CU
#define M_(NN) hipLaunchKernelGGL((k_fft_plane_r2c_ip<T, NN, 4, 8>), g, dim3(kThreads * 4), 0, ctx->st, a)
    FFT_SWITCH(ctx->nx, M_)
#undef M_
  } else {
#define M_(NN) hipLaunchKernelGGL((k_fft_plane_r2c<T, NN>), g, dim3(kThreads * PlaneFits<T, NN>::G), 0, ctx->st, a)
    FFT_SWITCH(ctx->nx, M_)
#undef M_
  }
  LAUNCHCHK();
  return EFFPOT_OK;
}

// The fused inverse y + x c2r plane transform (ctx->iplane) where its plane
// fits twice per CU with 32 lines per round (PlaneInv).
template <typename T> bool iplane_ok(effpot_ctx* ctx) {
  if (!ctx->iplane || ctx->nx != ctx->ny) return false;
  switch (ctx->nx) {
#define M_(NN) case NN: return PlaneInv<T, NN>::value;
    M_(16) M_(24) M_(32) M_(48) M_(64) M_(96) M_(128) M_(192) M_(256)
#undef M_
  }
  return false;
}
// half spectra (after the z transform) -> real fields, inverse y and x in one
// launch per plane: out = aff0 + scale*aff1*x
template <typename T> int fft_plane_inv(effpot_ctx* ctx, State<T>& s, cx<T>* const* in, T* const* out, int nf, T scale,
                                        const int* act, const T* aff0 = nullptr, const T* aff1 = nullptr,
                                        bool packed = false) {
  PlaneArgs<T> a = plane_args(ctx, s, out, in, nf, act);
  if (packed) { a.packed = 1; a.pitch = ctx->nxh - 1; a.wofs = packed_wofs(ctx); }
  a.scale = scale;
  for (int f = 0; f < nf; f++) {
    a.aff0[f] = aff0 ? aff0[f] : T(0);
    a.aff1[f] = aff1 ? aff1[f] : T(1);
  }
  dim3 g(ctx->nz, nf, ctx->B);
#define M_(NN) hipLaunchKernelGGL((k_fft_plane_c2r_ip<T, NN>), g, dim3(kThreads * PlaneInv<T, NN>::G), 0, ctx->st, a)
  FFT_SWITCH(ctx->nx, M_)
#undef M_
  LAUNCHCHK();
  return EFFPOT_OK;
}

// y lines: (kx, z) runs of nxp, element stride nxp; z lines: (kx, ky) runs of
// nxp, element stride nxp*ny.  Padding lines (kx >= nxh) are skipped.
LineMap ymap(effpot_ctx* ctx) {
  return LineMap{ctx->nxp, (long)ctx->nxp * ctx->ny, (long)ctx->nxp, ctx->nxp * ctx->nz, ctx->nxh};
}
LineMap zmap(effpot_ctx* ctx) {
  return LineMap{ctx->nxp, (long)ctx->nxp, (long)ctx->nxp * ctx->ny, ctx->nxp * ctx->ny, ctx->nxh};
}
template <typename T> int fft_strided(effpot_ctx* ctx, State<T>& s, int axis, cx<T>* const* f, int nf, int sign, const int* act) {
  StridedArgs<T> a{};
  for (int j = 0; j < nf; j++) a.f[j] = f[j];
  a.sstride = ctx->M; a.lm = axis == 1 ? ymap(ctx) : zmap(ctx); a.twN = s.tw[axis]; a.sign = sign; a.active = act;
  const int n = axis == 1 ? ctx->ny : ctx->nz;
  const int C = kThreads / n1_of(n);
  dim3 g((a.lm.nlines + C - 1) / C, nf, ctx->B);
#define M_(NN) hipLaunchKernelGGL((k_fft_strided<T, NN>), g, dim3(kThreads), 0, ctx->st, a)
  FFT_SWITCH(n, M_)
#undef M_
  LAUNCHCHK();
  return EFFPOT_OK;
}
// The packed layout (both in-place plane kernels in poten): rows of nx/2
// holding kx = 1 .. nx/2-1, and W = S(0) + i S(nx/2) in a block of its own
// (packed_wofs).  Where it applies (plane schedules 2-4
// with the in-place inverse) the z convolutions see 64 aligned lines per row at
// 128^3 instead of 65 in rows padded to 80, and the inverse plane reads whole
// 512-byte rows.
// The packed-column groups read their factors from precomputed pairs
// (State::kw, contiguous in ky like W): read from the padded tables instead,
// one cache line per value, they made those groups the longest chains and the
// layout slower where a launch is one wave of groups (64^3 FP64 single slot:
// step 162 -> 186 us); with the pairs it is as fast or faster at every shape
// measured (64^3 FP64 single 164 / 165 us, 96^3 FP64 single 324 -> 316 us,
// 96^3 FP32 x 8 950 -> 888 us, 128^3 FP32 x 4 1063 -> 1001 us).  The autotuner
// still times both.
template <typename T> bool packed_ok(effpot_ctx* ctx) {
  return ctx->packed != 0 && plane_mode<T>(ctx) >= 2 && iplane_ok<T>(ctx);
}
template <typename T> int fft_zconv(effpot_ctx* ctx, State<T>& s, ConvArgs<T>& a, int ngroups, const int* act,
                                    bool packed = false) {
  a.sstride = ctx->M; a.lm = zmap(ctx); a.km = a.lm; a.twN = s.tw[2]; a.active = act;
  a.scale = (T)(1.0 / ((double)ctx->nx * ctx->ny * ctx->nz));
  const int P = ctx->nxh - 1;
  a.nw = 0;
  if (packed) {  // lines kx' = 1 .. P-1 of rows of P; the factor tables keep their padded rows
    a.lm = LineMap{P, (long)P, (long)P * ctx->ny, P * ctx->ny, P, 1};
    a.km = LineMap{P, (long)ctx->nxp, (long)ctx->nxp * ctx->ny, P * ctx->ny, P, 1};
    a.P = P; a.ny = ctx->ny; a.wofs = packed_wofs(ctx);
    for (int g = 0; g < ngroups; g++)  // each factor table's pairs for the packed column
      for (int j = 0; j < a.nout[g]; j++)
        for (int k = 0; k < 3; k++)
          if (a.kern[g][j] == s.kern[k]) a.kw[g][j] = s.kw[k];
  }
  const int th = ctx->zconv_th;
  const int C = th / n1_of(ctx->nz);
  if (packed) a.nw = (ctx->ny + C - 1) / C;  // groups of the packed column, launched first
  dim3 g(a.nw + (a.lm.nlines + C - 1) / C, ngroups, ctx->B);
  // One output per input (stages B and C): the instance that keeps no spectrum
  // or next factors between outputs holds 91 VGPRs instead of 162 at 128 FP32
  // (occupancy 5 instead of 3; stage B 81 -> 77 us, C 22 -> 19 us at 128^3 x 4).
  // In FP64 it measured slower at 128^3 (step 544 -> 547 us) and is not used.
  bool one = true;
  for (int k = 0; k < ngroups; k++) one = one && a.nout[k] == 1;
  if constexpr (sizeof(T) == 4) {
    if (one) {
      if (th == 512) {
#define M_(NN) hipLaunchKernelGGL((k_fft_zconv<T, NN, 512, true>), g, dim3(512), 0, ctx->st, a)
        FFT_SWITCH(ctx->nz, M_)
#undef M_
      } else {
#define M_(NN) hipLaunchKernelGGL((k_fft_zconv<T, NN, kThreads, true>), g, dim3(kThreads), 0, ctx->st, a)
        FFT_SWITCH(ctx->nz, M_)
#undef M_
      }
      LAUNCHCHK();
      return EFFPOT_OK;
    }
  }
  if (th == 512) {
#define M_(NN) hipLaunchKernelGGL((k_fft_zconv<T, NN, 512, false>), g, dim3(512), 0, ctx->st, a)
    FFT_SWITCH(ctx->nz, M_)
#undef M_
  } else {
#define M_(NN) hipLaunchKernelGGL((k_fft_zconv<T, NN, kThreads, false>), g, dim3(kThreads), 0, ctx->st, a)
    FFT_SWITCH(ctx->nz, M_)
#undef M_
  }
  LAUNCHCHK();
  return EFFPOT_OK;
}

// ---- stencil launchers -----------------------------------------------------------
#define TX_SWITCH(tx, M_)                         \
  switch (tx) {                                   \
    case 64: M_(64); break; case 32: M_(32); break; \
    case 16: M_(16); break; default: M_(8); break;  \
  }

template <int TX> int nblk_of(effpot_ctx* ctx) {
  dim3 g = zmarch_grid<TX>(ctx->nx, ctx->ny, ctx->nz, ctx->zc, 1);
  return (int)(g.x * g.y);
}
int nblk(effpot_ctx* ctx) {
  int r = 0;
#define M_(TXX) r = nblk_of<TXX>(ctx)
  TX_SWITCH(ctx->tx, M_)
#undef M_
  return r;
}
dim3 sgrid(effpot_ctx* ctx) {
  dim3 g;
#define M_(TXX) g = zmarch_grid<TXX>(ctx->nx, ctx->ny, ctx->nz, ctx->zc, ctx->B)
  TX_SWITCH(ctx->tx, M_)
#undef M_
  return g;
}
Grid grid_of(effpot_ctx* ctx) {
  return Grid{ctx->nx, ctx->ny, ctx->nz, ctx->p.icon == EFFPOT_ICON_PERIODIC ? 1 : 0, ctx->N};
}

template <typename T> Ring<T> ring_of(effpot_ctx* ctx, State<T>& s) { return Ring<T>{s.psi, s.hp, s.ring, ctx->N}; }

template <typename T> int ring_op(effpot_ctx* ctx, State<T>& s, int op, const int* act) {
  hipLaunchKernelGGL(k_ring, dim3((ctx->B + 63) / 64), dim3(64), 0, ctx->st, s.ring, act, ctx->B, op);
  LAUNCHCHK();
  return EFFPOT_OK;
}

// Runge-Kutta-Gill stage coefficients (unused by the PC kernels).
struct StageCoef { double arun = 0, brun = 0, crun = 0; int first = 0; };

template <typename T, typename V, typename P, int MODE>
int launch_step_v(effpot_ctx* ctx, State<T>& s, State<P>& f, const int* act, StageCoef c) {
  StepArgs<T, V, P> a{};
  a.g = grid_of(ctx); a.sc = s.sc; a.zc = ctx->zc; a.h2o2m4 = (T)ctx->p.h2o2m4; a.dtv = ctx->d_dt;
  a.active = act; a.slot = ctx->N;
  // pc and q are allocated (and reset per slot) as complex arrays of N per
  // slot; a real wavefunction uses the first half of its slot's range.
  a.pc = reinterpret_cast<V*>(s.pc); a.q = reinterpret_cast<V*>(s.q);
  a.pcq_slot = ctx->N * (long)(sizeof(cx<T>) / sizeof(V));
  a.pot4 = f.pot4; a.uimp = f.uimp; a.den = f.den;
  a.partial = ctx->d_partial; a.nblk = nblk(ctx);
  a.arun = (T)c.arun; a.brun = (T)c.brun; a.crun = (T)c.crun; a.first_stage = c.first;
  a.write_den = ctx->corr_den ? 1 : 0;
  a.rt = ctx->rt;
  Ring<T> rg = ring_of(ctx, s);
  dim3 g = sgrid(ctx);
#define M_(TXX) hipLaunchKernelGGL((k_step<T, V, P, TXX, MODE>), g, dim3(kThreads), 0, ctx->st, a, rg)
  TX_SWITCH(ctx->tx, M_)
#undef M_
  LAUNCHCHK();
  return EFFPOT_OK;
}
// f32: the step reads/writes the FP32 mean fields of a MIXED context.
template <typename T, int MODE>
int launch_step(effpot_ctx* ctx, State<T>& s, const int* act, bool f32, StageCoef c = StageCoef()) {
  if (f32) {
    if (ctx->cplx) return launch_step_v<T, cx<T>, float, MODE>(ctx, s, ctx->s32, act, c);
    return launch_step_v<T, T, float, MODE>(ctx, s, ctx->s32, act, c);
  }
  if (ctx->cplx) return launch_step_v<T, cx<T>, T, MODE>(ctx, s, s, act, c);
  return launch_step_v<T, T, T, MODE>(ctx, s, s, act, c);
}

// d_dt[s] = per-slot override, else `dt` (uploaded only when it changes).
int push_dt(effpot_ctx* ctx, double dt) {
  bool change = false;
  std::vector<double> v(ctx->B);
  for (int b = 0; b < ctx->B; b++) {
    v[b] = ctx->h_dt[b] > 0 ? ctx->h_dt[b] : dt;
    if (v[b] != ctx->dt_dev[b]) change = true;
  }
  if (!change) return EFFPOT_OK;
  HIPCHK(hipStreamSynchronize(ctx->st));  // kernels in flight read d_dt
  HIPCHK(copy_sync(ctx, ctx->d_dt, v.data(), sizeof(double) * ctx->B, hipMemcpyHostToDevice));
  ctx->dt_dev = v;
  return EFFPOT_OK;
}
// Launch a ring kernel templated on <T, V> with V chosen by the context's psi type.
#define PSI_LAUNCH(KERNEL, T_, GRID, ...)                                                          \
  do {                                                                                               \
    if (ctx->cplx) hipLaunchKernelGGL((KERNEL<T_, cx<T_>>), GRID, dim3(256), 0, ctx->st, __VA_ARGS__); \
    else hipLaunchKernelGGL((KERNEL<T_, T_>), GRID, dim3(256), 0, ctx->st, __VA_ARGS__);              \
    LAUNCHCHK();                                                                                     \
  } while (0)
// Same for <T, V, P> kernels writing fields of precision P.
#define PSI_LAUNCH_P(KERNEL, T_, P_, GRID, ...)                                                          \
  do {                                                                                                     \
    if (ctx->cplx) hipLaunchKernelGGL((KERNEL<T_, cx<T_>, P_>), GRID, dim3(256), 0, ctx->st, __VA_ARGS__); \
    else hipLaunchKernelGGL((KERNEL<T_, T_, P_>), GRID, dim3(256), 0, ctx->st, __VA_ARGS__);              \
    LAUNCHCHK();                                                                                           \
  } while (0)
// den (precision of the fields in use) = |psi_cur|^2 for the slots in `act`.
int den_from_cur(effpot_ctx* ctx, const int* act, bool f32) {
  if (f32) PSI_LAUNCH_P(k_den_from_cur, double, float, pw_grid(ctx->N, ctx->B), ring_of(ctx, ctx->s64), ctx->s32.den, act, ctx->N);
  else PSI_LAUNCH_P(k_den_from_cur, double, double, pw_grid(ctx->N, ctx->B), ring_of(ctx, ctx->s64), ctx->s64.den, act, ctx->N);
  return EFFPOT_OK;
}

template <int K> int reduce(effpot_ctx* ctx, int mode, const int* act, int* ring = nullptr, int op1 = -1,
                            int op2 = -1) {
  hipLaunchKernelGGL((k_reduce<K>), dim3(ctx->B), dim3(kThreads), 0, ctx->st, ctx->d_partial, nblk(ctx), ctx->d_red,
                     act, mode, (const double*)ctx->d_n4, ctx->dxyz, mode == 2 ? ctx->d_dscale : ctx->d_scale,
                     ring, op1, op2);
  LAUNCHCHK();
  return EFFPOT_OK;
}

#define RC(x) do { int rc_ = (x); if (rc_ != EFFPOT_OK) return rc_; } while (0)

// ---- poten (poten.f90:5-131 + term_alfa.f90:5-109) ----------------------------
// `only` >= 0 runs just launch number `only` of the sequence (benchmark hook:
// 0-4 stage A x r2c, y, z convolution, y, x c2r; 5 k_grad; 6-10 stage B;
// 11 k_div; 12-16 stage C; 17 k_pot4; with the fused plane transforms 0/6/12
// are the x+y forward transforms and 1/7/13 are empty).
// dscale: [slot] factor the density is read with (the step's density scaled to
// n4 atoms, from step_den_scale), or null.  grads: store the density gradient
// (dx, dy, dz) also where k_div forms its own (FP32 field downloads).
template <typename T> int poten_t(effpot_ctx* ctx, State<T>& s, const int* act, int only = -1,
                                  const double* dscale = nullptr, bool grads = false) {
  const bool otc = ctx->p.core == EFFPOT_CORE_OTC;
  const bool pl = plane_ok<T>(ctx);  // forward x and y passes fused: launches 1, 7, 13 are empty
  // FP32 fields with the plane transforms: k_div forms the density gradient
  // from den and k_grad stores 4 fields instead of 7 (the FP64 gradient is
  // kept: the energy and the impurity force read it; with separate x passes
  // stage B's x transform reads it)
  const bool regrad = sizeof(T) == 4 && otc && pl && !grads;
  const bool ipl = iplane_ok<T>(ctx);  // inverse y and x passes fused: launches 3, 9, 15 are empty
  const bool pk = packed_ok<T>(ctx);   // packed half spectra between the plane kernels (implies pl and ipl)
  auto sel = [only](int k) { return only < 0 || only == k; };
  // stage A: FFT(den) -> dencg, delj4 [, denalf]
  {
    T* in[1] = {s.den};
    cx<T>* sp[1] = {s.spec[0]};
    if (sel(0)) RC(pl ? fft_plane_fwd(ctx, s, in, sp, 1, act, pk, dscale) : fft_x_r2c(ctx, s, in, sp, 1, act, dscale));
    if (sel(1) && !pl) RC(fft_strided(ctx, s, 1, sp, 1, -1, act));
    ConvArgs<T> ca{};
    ca.in[0] = s.spec[0];
    ca.nout[0] = otc ? 3 : 2;
    ca.out[0][0] = s.spec[1]; ca.kern[0][0] = s.kern[0];
    ca.out[0][1] = s.spec[2]; ca.kern[0][1] = s.kern[1];
    ca.out[0][2] = s.spec[3]; ca.kern[0][2] = s.kern[2];
    if (sel(2)) RC(fft_zconv(ctx, s, ca, 1, act, pk));
    cx<T>* o[3] = {s.spec[1], s.spec[2], s.spec[3]};
    if (sel(3) && !ipl) RC(fft_strided(ctx, s, 1, o, ca.nout[0], +1, act));
    // the alpha_s convolution is stored as falfs = 1 - denalf/den0s (term_alfa.f90:157-159),
    // the only form its consumers use
    T* r[3] = {s.dencg, s.delj4, s.denalf};
    const T a0[3] = {T(0), T(0), T(1)};
    const T a1[3] = {T(1), T(1), T(-1.0 / ctx->p.den0s)};
    if (sel(4))
      RC(ipl ? fft_plane_inv(ctx, s, o, r, ca.nout[0], T(1), act, a0, a1, pk)
             : fft_x_c2r(ctx, s, o, r, ca.nout[0], T(1), act, a0, a1));
  }
  // derden + FFT sources
  if (sel(5)) {
    PotenArgs<T> a{};
    a.g = grid_of(ctx); a.sc = s.sc; a.zc = ctx->zc; a.active = act; a.slot = ctx->N; a.otc = otc;
    a.cp4 = (T)ctx->p.cp4; a.cpp4 = (T)ctx->p.cpp4; a.den0s = (T)ctx->p.den0s;
    a.den = s.den; a.dencg = s.dencg; a.denalf = s.denalf; a.dx = s.dx; a.dy = s.dy; a.dz = s.dz;
    // with separate x and y passes falfs * d{x,y,z}den are formed by stage B's
    // x transform from dx, dy, dz and not stored
    a.s0 = s.src[0]; a.dscale = dscale;
    if (pl) { a.s1 = s.src[1]; a.s2 = s.src[2]; a.s3 = s.src[3]; }
    if (regrad) a.dx = a.dy = a.dz = nullptr;
    dim3 g = sgrid(ctx);
#define M_(TXX) hipLaunchKernelGGL((k_grad<T, TXX>), g, dim3(kThreads), 0, ctx->st, a)
    TX_SWITCH(ctx->tx, M_)
#undef M_
    LAUNCHCHK();
  }
  // stage B: correlation convolution [+ alpha_s integrals]
  {
    const int nf = otc ? 4 : 1;
    // sources: the correlation source and falfs * d{x,y,z}den.  The x pass
    // forms the products as it reads dx, dy, dz (k_grad then stores 4 fields
    // instead of 7: 128^3 FP64 step 545 -> 538 us); the plane kernel, whose
    // loads are its latency chain, reads the stored products (with the
    // products formed on load its stage-B launch took 99-120 us against 67 at
    // 128^3 x 4 FP32, more than k_grad saved)
    T* in[4] = {s.src[0], s.src[1], s.src[2], s.src[3]};
    if (!pl) { in[1] = s.dx; in[2] = s.dy; in[3] = s.dz; }
    cx<T>* sp[4] = {s.spec[4], s.spec[5], s.spec[6], s.spec[7]};
    const T* fa = otc ? s.denalf : nullptr;  // denalf holds falfs
    if (sel(6))
      RC(pl ? fft_plane_fwd(ctx, s, in, sp, nf, act, pk) : fft_x_r2c(ctx, s, in, sp, nf, act, (const double*)nullptr, fa));
    if (sel(7) && !pl) RC(fft_strided(ctx, s, 1, sp, nf, -1, act));
    ConvArgs<T> ca{};
    for (int g = 0; g < nf; g++) {
      ca.in[g] = sp[g];
      ca.nout[g] = 1;
      ca.out[g][0] = s.spec[g];
      ca.kern[g][0] = g == 0 ? s.kern[0] : s.kern[2];
    }
    if (sel(8)) RC(fft_zconv(ctx, s, ca, nf, act, pk));
    cx<T>* o[4] = {s.spec[0], s.spec[1], s.spec[2], s.spec[3]};
    if (sel(9) && !ipl) RC(fft_strided(ctx, s, 1, o, nf, +1, act));
    T* r[4] = {s.sto1, s.intx, s.inty, s.intz};
    if (sel(10))
      RC(ipl ? fft_plane_inv(ctx, s, o, r, nf, T(1), act, (const T*)nullptr, (const T*)nullptr, pk) : fft_x_c2r(ctx, s, o, r, nf, T(1), act));
  }
  if (otc) {
    DivArgs<T> a{};
    a.g = grid_of(ctx); a.sc = s.sc; a.zc = ctx->zc; a.active = act; a.slot = ctx->N; a.den0s = (T)ctx->p.den0s;
    a.denalf = s.denalf; a.ix = s.intx; a.iy = s.inty; a.iz = s.intz; a.dx = s.dx; a.dy = s.dy; a.dz = s.dz;
    a.src = s.src[0]; a.div = s.div; a.den = s.den; a.dscale = dscale;
    dim3 g = sgrid(ctx);
    if (sel(11)) {
      bool done = false;
      if constexpr (sizeof(T) == 4) {
        if (regrad) {
#define M_(TXX) hipLaunchKernelGGL((k_div<T, TXX, true>), g, dim3(kThreads), 0, ctx->st, a)
          TX_SWITCH(ctx->tx, M_)
#undef M_
          done = true;
        }
      }
      if (!done) {
#define M_(TXX) hipLaunchKernelGGL((k_div<T, TXX>), g, dim3(kThreads), 0, ctx->st, a)
        TX_SWITCH(ctx->tx, M_)
#undef M_
      }
      LAUNCHCHK();
    }
    // stage C: FFT(sum_i d_i den int_i) * kalfs
    T* in[1] = {s.src[0]};
    cx<T>* sp[1] = {s.spec[4]};
    if (sel(12)) RC(pl ? fft_plane_fwd(ctx, s, in, sp, 1, act, pk) : fft_x_r2c(ctx, s, in, sp, 1, act));
    if (sel(13) && !pl) RC(fft_strided(ctx, s, 1, sp, 1, -1, act));
    ConvArgs<T> ca{};
    ca.in[0] = s.spec[4]; ca.nout[0] = 1; ca.out[0][0] = s.spec[0]; ca.kern[0][0] = s.kern[2];
    if (sel(14)) RC(fft_zconv(ctx, s, ca, 1, act, pk));
    cx<T>* o[1] = {s.spec[0]};
    if (sel(15) && !ipl) RC(fft_strided(ctx, s, 1, o, 1, +1, act));
    T* r[1] = {s.ua};
    if (sel(16))
      RC(ipl ? fft_plane_inv(ctx, s, o, r, 1, T(1), act, (const T*)nullptr, (const T*)nullptr, pk) : fft_x_c2r(ctx, s, o, r, 1, T(1), act));
  }
  if (sel(17)) {
    Pot4Args<T> a{};
    a.N = ctx->N; a.slot = ctx->N; a.active = act; a.otc = otc;
    a.a0 = (T)(ctx->p.cp4 / 2.0); a.a1 = (T)(ctx->p.cpp4 / 3.0); a.den0s = (T)ctx->p.den0s;
    a.h2o2m4 = (T)ctx->p.h2o2m4; a.alphas = (T)ctx->p.alphas; a.C = (T)ctx->p.C; a.beta = (T)ctx->p.beta;
    a.den_m = (T)ctx->p.den_m;
    a.delj4 = s.delj4; a.dencg = s.dencg; a.sto1 = s.sto1; a.ua = s.ua; a.div = s.div; a.den = s.den;
    a.pot4 = s.pot4; a.ualphas = s.ualphas; a.corr = s.corr; a.dscale = dscale;
    hipLaunchKernelGGL((k_pot4<T>), pw_grid(ctx->N, ctx->B), dim3(256), 0, ctx->st, a);
    LAUNCHCHK();
  }
  return EFFPOT_OK;
}

// ---- integrators ----------------------------------------------------------------
// With FP32 fields the density is not stored: every path that follows (the
// next predictor or RK stage, an energy check) recomputes it before reading.
template <typename T> int renorm_t(effpot_ctx* ctx, State<T>& s, const int* act, bool f32 = false) {
  if (f32)
    PSI_LAUNCH_P(k_renorm_ring, T, float, pw_grid(ctx->N, ctx->B), ring_of(ctx, s), (float*)nullptr,
                 (const double*)ctx->d_scale, act, ctx->N);
  else
    PSI_LAUNCH_P(k_renorm_ring, T, T, pw_grid(ctx->N, ctx->B), ring_of(ctx, s), s.den, (const double*)ctx->d_scale,
                 act, ctx->N);
  return EFFPOT_OK;
}

// scale[s] = sqrt(n4 / (sum |psi|^2 dxyz)) for the live psi
template <typename T> int norm_t(effpot_ctx* ctx, State<T>& s, const int* act) {
  PSI_LAUNCH(k_norm, T, dim3(nblk(ctx), ctx->B), ring_of(ctx, s), act, ctx->N, ctx->d_partial);
  return reduce<2>(ctx, 1, act);
}

// The slot's impurity stops moving (a new impurity was set, or mass 0).
int stop_motion(effpot_ctx* ctx, int slot) {
  if (!ctx->h_motion[slot].on) return EFFPOT_OK;
  HIPCHK(hipStreamSynchronize(ctx->st));
  ctx->h_motion[slot] = Motion{};
  HIPCHK(copy_sync(ctx, ctx->d_motion + slot, &ctx->h_motion[slot], sizeof(Motion), hipMemcpyHostToDevice));
  ctx->any_motion = 0;
  for (const Motion& q : ctx->h_motion) ctx->any_motion |= q.on;
  return EFFPOT_OK;
}

// The slot's invar stops evolving (a new impurity was set, or switched off).
int stop_dyn(effpot_ctx* ctx, int slot) {
  if (!ctx->h_dyn[slot].on) return EFFPOT_OK;
  if (ctx->h_is_state[slot]) RC(stop_motion(ctx, slot));  // a P/D impurity moves only with its invar
  HIPCHK(hipStreamSynchronize(ctx->st));
  ctx->h_dyn[slot] = InvarDyn{};
  HIPCHK(copy_sync(ctx, ctx->d_dyn + slot, &ctx->h_dyn[slot], sizeof(InvarDyn), hipMemcpyHostToDevice));
  ctx->any_dyn = 0;
  for (const InvarDyn& q : ctx->h_dyn) ctx->any_dyn |= q.on;
  return EFFPOT_OK;
}

// uimp of the moving slots at their current positions (FP64 fields).
int move_field(effpot_ctx* ctx, const int* act) {
  hipLaunchKernelGGL(k_imp_field_moving, dim3(256, ctx->B), dim3(256), 0, ctx->st, (const Motion*)ctx->d_motion, act,
                     ctx->nx, ctx->ny, ctx->nz, -ctx->p.xmax, -ctx->p.ymax, -ctx->p.zmax, ctx->hx, ctx->hy, ctx->hz,
                     ctx->npot, (double)(ctx->npot - 1) * ctx->delinter, (const double*)ctx->imptab, ctx->s64.uimp,
                     ctx->N);
  LAUNCHCHK();
  return EFFPOT_OK;
}
// potenimp of the slots whose invar evolves, with their current invar.
int dyn_field(effpot_ctx* ctx, const int* act) {
  hipLaunchKernelGGL(k_dyn_uimp, dim3(256, ctx->B), dim3(256), 0, ctx->st, (const InvarDyn*)ctx->d_dyn,
                     (const Motion*)ctx->d_motion, act, ctx->s64.uimp, ctx->N);
  LAUNCHCHK();
  return EFFPOT_OK;
}
// potinvar after poten: Hinvar from the moments of the current density.
int dyn_hiv(effpot_ctx* ctx, const int* act) {
  if (!ctx->d_dynpart) RC(dalloc(ctx, &ctx->d_dynpart, (size_t)ctx->B * kStateBlocks * kMomD));
  hipLaunchKernelGGL(k_dyn_moments, dim3(kStateBlocks, ctx->B), dim3(256), 0, ctx->st, (const InvarDyn*)ctx->d_dyn,
                     (const Motion*)ctx->d_motion, act, (const double*)ctx->s64.den, ctx->N, ctx->d_dynpart);
  LAUNCHCHK();
  hipLaunchKernelGGL(k_dyn_hiv, dim3((ctx->B + 63) / 64), dim3(64), 0, ctx->st, ctx->d_dyn, act, ctx->B,
                     (const double*)ctx->d_dynpart, ctx->dxyz);
  LAUNCHCHK();
  return EFFPOT_OK;
}
// forceimp (potenimp.f90:193-258) after poten: a = -sum grad(den) uimp dV / m.
int force_t(effpot_ctx* ctx, const int* act) {
  hipLaunchKernelGGL(k_force_partials, dim3(nblk(ctx), ctx->B), dim3(256), 0, ctx->st, (const Motion*)ctx->d_motion,
                     (const PlusSlot*)ctx->d_plus, act, (const double*)ctx->s64.dx, (const double*)ctx->s64.dy,
                     (const double*)ctx->s64.dz, (const double*)ctx->s64.uimp, ctx->N, (long)ctx->nx * ctx->ny,
                     ctx->nz, -ctx->p.zmax, ctx->hz, ctx->d_partial);
  LAUNCHCHK();
  RC(reduce<3>(ctx, 0, act));
  hipLaunchKernelGGL(k_motion_force, dim3((ctx->B + 63) / 64), dim3(64), 0, ctx->st, ctx->d_motion, act, ctx->B,
                     (const double*)ctx->d_red, ctx->dxyz);
  LAUNCHCHK();
  return EFFPOT_OK;
}

// Imaginary time: the density of the stage or predicted state (written with
// its block partial sums by the step kernel) scaled to n4 atoms before the
// step's poten.  The step multiplies psi by about 1 - mu dt and only the
// renormalisation after the step restores n4, so the field of the unscaled
// density would be that of ~(1 - 2 mu dt) n4 atoms and the relaxation's fixed
// point would miss the eigenvalue equation by O(dt).  Real time keeps
// den = |psi|^2 as the reference's stepper has it.
int step_den_scale(effpot_ctx* ctx, const int* act, bool f32, const double** dscale) {
  *dscale = nullptr;
  if (ctx->rt) return EFFPOT_OK;
  RC(reduce<2>(ctx, 2, act));
  if (!ctx->any_motion && !ctx->any_dyn) {  // poten reads the density through the factor
    *dscale = ctx->d_dscale;
    return EFFPOT_OK;
  }
  // the impurity's force and the internal-state moments read den after poten: scale it in place
  if (f32)
    hipLaunchKernelGGL((k_scale_den<float>), pw_grid(ctx->N, ctx->B), dim3(256), 0, ctx->st, ctx->s32.den,
                       (const double*)ctx->d_dscale, act, ctx->N);
  else
    hipLaunchKernelGGL((k_scale_den<double>), pw_grid(ctx->N, ctx->B), dim3(256), 0, ctx->st, ctx->s64.den,
                       (const double*)ctx->d_dscale, act, ctx->N);
  LAUNCHCHK();
  return EFFPOT_OK;
}

// RK-Gill step (steprk.f90:286-405).  Leaves the norm reduction in d_red/d_scale.
template <typename T> int step_rk_t(effpot_ctx* ctx, State<T>& s, const int* act, bool f32 = false) {
  const double r2 = std::sqrt(2.0);
  const double arun[4] = {0.5, 1.0 - 1.0 / r2, 1.0 + 1.0 / r2, 1.0 / 6.0};
  const double brun[4] = {2.0, 1.0, 1.0, 2.0};
  const double crun[4] = {0.5, 1.0 - 1.0 / r2, 1.0 + 1.0 / r2, 0.5};
  for (int j = 0; j < 4; j++) {
    StageCoef c;
    c.arun = arun[j]; c.brun = brun[j]; c.crun = crun[j]; c.first = j == 0;
    RC(ring_op(ctx, s, j == 0 ? RING_BEFORE_RK0 : RING_BEFORE_RKJ, act));
    RC((launch_step<T, 2>(ctx, s, act, f32, c)));
    RC(ring_op(ctx, s, j == 0 ? RING_AFTER_RK0 : RING_AFTER_RKJ, act));
    if (ctx->any_motion) {  // r, v stage; the field at the new position (steprk.f90:117-170)
      hipLaunchKernelGGL(k_motion_rk, dim3((ctx->B + 63) / 64), dim3(64), 0, ctx->st, ctx->d_motion, act, ctx->B, j,
                         (const double*)ctx->d_dt, c.arun, c.brun, c.crun);
      LAUNCHCHK();
      if (j <= 2) RC(move_field(ctx, act));
    }
    if (ctx->any_dyn) {  // invar stage; potenimp with the new invar (steprk.f90:152-170)
      hipLaunchKernelGGL(k_dyn_rk, dim3((ctx->B + 63) / 64), dim3(64), 0, ctx->st, ctx->d_dyn, act, ctx->B, j,
                         (const double*)ctx->d_dt, c.arun, c.brun, c.crun);
      LAUNCHCHK();
      if (j <= 2) RC(dyn_field(ctx, act));
    }
    const double* ds = nullptr;
    if (j <= 2) RC(step_den_scale(ctx, act, f32, &ds));
    if (j <= 2) RC(f32 ? poten_t(ctx, ctx->s32, act, -1, ds) : poten_t(ctx, s, act, -1, ds));
    if (ctx->any_motion && j <= 2) RC(force_t(ctx, act));
    if (ctx->any_dyn && j <= 2) RC(dyn_hiv(ctx, act));
  }
  RC(reduce<2>(ctx, 1, act));
  return EFFPOT_OK;
}

// Hamming predictor-modifier-corrector (steppc.f90:35-159).
template <typename T> int step_pc_t(effpot_ctx* ctx, State<T>& s, const int* act, bool f32 = false) {
  RC((launch_step<T, 0>(ctx, s, act, f32)));
  if (ctx->any_motion) {  // r, v predictor-modifier and the field there (steppc.f90:68-113)
    hipLaunchKernelGGL(k_motion_pred, dim3((ctx->B + 63) / 64), dim3(64), 0, ctx->st, ctx->d_motion, act, ctx->B,
                       (const double*)ctx->d_dt);
    LAUNCHCHK();
    RC(move_field(ctx, act));
  }
  if (ctx->any_dyn) {  // invar predictor-modifier and potenimp with it (steppc.f90:84-112)
    hipLaunchKernelGGL(k_dyn_pred, dim3((ctx->B + 63) / 64), dim3(64), 0, ctx->st, ctx->d_dyn, act, ctx->B,
                       (const double*)ctx->d_dt);
    LAUNCHCHK();
    RC(dyn_field(ctx, act));
  }
  const double* ds = nullptr;
  RC(step_den_scale(ctx, act, f32, &ds));
  RC(f32 ? poten_t(ctx, ctx->s32, act, -1, ds) : poten_t(ctx, s, act, -1, ds));
  if (ctx->any_motion) RC(force_t(ctx, act));
  if (ctx->any_dyn) RC(dyn_hiv(ctx, act));  // potinvar
  RC((launch_step<T, 1>(ctx, s, act, f32)));  // roles as after RING_AFTER_PRED
  // norm sums, then both role rotations of the step in the same launch
  RC(reduce<2>(ctx, 1, act, s.ring, RING_AFTER_PRED, RING_AFTER_CORR));
  if (ctx->any_motion) {
    hipLaunchKernelGGL(k_motion_corr, dim3((ctx->B + 63) / 64), dim3(64), 0, ctx->st, ctx->d_motion, act, ctx->B,
                       (const double*)ctx->d_dt);
    LAUNCHCHK();
  }
  if (ctx->any_dyn) {
    hipLaunchKernelGGL(k_dyn_corr, dim3((ctx->B + 63) / 64), dim3(64), 0, ctx->st, ctx->d_dyn, act, ctx->B,
                       (const double*)ctx->d_dt);
    LAUNCHCHK();
  }
  return EFFPOT_OK;
}

// energy reductions -> host energies
template <typename T> int energy_t(effpot_ctx* ctx, State<T>& s, const int* act, effpot_energies* out) {
  EnergyArgs<T> a{};
  const effpot_params& p = ctx->p;
  a.g = grid_of(ctx); a.sc = s.sc; a.zc = ctx->zc; a.active = act; a.slot = ctx->N;
  a.otc = p.core == EFFPOT_CORE_OTC;
  a.h2o2m4 = (T)p.h2o2m4; a.aux1 = (T)(0.5 * p.cp4); a.aux2 = (T)((1.0 / 3.0) * p.cpp4);
  a.beta = (T)p.beta; a.den_m = (T)p.den_m; a.den0s = (T)p.den0s;
  a.x0 = -p.xmax; a.y0 = -p.ymax; a.z0 = -p.zmax; a.hx = ctx->hx; a.hy = ctx->hy; a.hz = ctx->hz;
  a.den = s.den; a.dencg = s.dencg; a.delj4 = s.delj4; a.denalf = s.denalf;
  a.dx = s.dx; a.dy = s.dy; a.dz = s.dz; a.ix = s.intx; a.iy = s.inty; a.iz = s.intz;
  a.uimp = s.uimp; a.pot4 = s.pot4; a.partial = ctx->d_partial; a.nblk = nblk(ctx);
  a.psi_all = s.psi; a.ring = s.ring; a.plus = ctx->d_plus;
  if (ctx->any_dyn) {  // current invar for eso
    HIPCHK(hipStreamSynchronize(ctx->st));
    HIPCHK(copy_sync(ctx, ctx->h_dyn.data(), ctx->d_dyn, sizeof(InvarDyn) * ctx->B, hipMemcpyDeviceToHost));
  }
  if (ctx->any_motion) {  // current positions for uplus, velocities for ekinx
    HIPCHK(hipStreamSynchronize(ctx->st));
    HIPCHK(copy_sync(ctx, ctx->h_motion.data(), ctx->d_motion, sizeof(Motion) * ctx->B, hipMemcpyDeviceToHost));
    for (int b = 0; b < ctx->B; b++)
      if (ctx->h_motion[b].on)
        for (int c = 0; c < 3; c++) ctx->h_plus[b].rimp[c] = ctx->h_motion[b].r[c];
    HIPCHK(hipMemcpyAsync(ctx->d_plus, ctx->h_plus.data(), sizeof(PlusSlot) * ctx->B, hipMemcpyHostToDevice, ctx->st));
  }
  dim3 g = sgrid(ctx);
#define M_(TXX)                                                                                   \
  if (ctx->cplx) hipLaunchKernelGGL((k_energy<T, cx<T>, TXX>), g, dim3(kThreads), 0, ctx->st, a); \
  else hipLaunchKernelGGL((k_energy<T, T, TXX>), g, dim3(kThreads), 0, ctx->st, a)
  TX_SWITCH(ctx->tx, M_)
#undef M_
  LAUNCHCHK();
  RC(reduce<kEnergySums>(ctx, 0, act));
  HIPCHK(hipMemcpyAsync(ctx->h_pin, ctx->d_red, sizeof(double) * ctx->B * kEnergySums, hipMemcpyDeviceToHost, ctx->st));
  HIPCHK(hipStreamSynchronize(ctx->st));
  const double dv = ctx->dxyz;
  for (int b = 0; b < ctx->B; b++) {
    const double* S = ctx->h_pin + b * kEnergySums;
    effpot_energies e;
    std::memset(&e, 0, sizeof(e));
    e.natoms = S[6] * dv;
    e.ekin4 = p.h2o2m4 * S[0] * dv;
    e.elj4 = 0.5 * S[1] * dv;
    e.ecor4 = S[2] * dv;
    e.esolid = p.C * S[4] * dv;
    e.etot4 = e.ekin4 + e.elj4 + e.ecor4;
    if (p.core == EFFPOT_CORE_OTC) {
      e.ealphas = -p.h2o2m4 * 0.5 * p.alphas * S[3] * dv;
      e.etot4 += e.ealphas;
    }
    e.eHeX = S[5] * dv;
    e.uplus = S[11] * dv;
    e.eso = ctx->h_eso[b];
    if (ctx->any_dyn && ctx->h_dyn[b].on) e.eso = state_eso(ctx->h_dyn[b].state, ctx->h_dyn[b].iv, ctx->h_dyn[b].als);
    if (ctx->any_motion && ctx->h_motion[b].on) {  // energy.f90:112
      const Motion& q = ctx->h_motion[b];
      e.ekinx = 0.5 * q.m * (q.v[0] * q.v[0] + q.v[1] * q.v[1] + q.v[2] * q.v[2]);
    }
    e.eimpu = e.ekinx + e.eHeX + e.eso;
    e.etot = e.etot4 + e.eimpu;
    e.mu = e.natoms > 0 ? S[10] * dv / e.natoms : 0.0;
    // |H psi - mu psi|^2 = |H psi|^2 - mu^2 N (both weighted by dV)
    const double r2 = S[12] * dv - e.mu * e.mu * e.natoms;
    e.resid = (e.natoms > 0 && e.mu != 0.0) ? std::sqrt(std::max(r2, 0.0)) / (std::fabs(e.mu) * std::sqrt(e.natoms)) : 0.0;
    const double n4 = ctx->h_n4[b];
    e.xcm = S[7] * dv / n4; e.ycm = S[8] * dv / n4; e.zcm = S[9] * dv / n4;
    out[b] = e;
  }
  return EFFPOT_OK;
}

template <typename T> State<T>& st_of(effpot_ctx* ctx);
template <> State<double>& st_of<double>(effpot_ctx* ctx) { return ctx->s64; }
template <> State<float>& st_of<float>(effpot_ctx* ctx) { return ctx->s32; }

int check_ctx(effpot_ctx* ctx) { return ctx ? EFFPOT_OK : EFFPOT_ERR_ARG; }

}  // namespace

// ===========================================================================
// C ABI (linkage comes from the extern "C" declarations in effpot.h)
// ===========================================================================

int64_t effpot_footprint_bytes(const effpot_params* p) {
  if (!p) return -1;
  const int64_t N = (int64_t)p->nx * p->ny * p->nz, M = (int64_t)((p->nx / 2 + 1 + 15) / 16 * 16) * p->ny * p->nz;
  const int64_t per64 = (2 * 9 + 20) * N * 8 + 8 * M * 16;
  int64_t per = per64;
  if (p->precision == EFFPOT_PREC_MIXED) per += per64 / 2;
  return per * (int64_t)p->batch + 3 * M * 12 + 3 * 2 * (int64_t)p->ny * p->nz * 12 +
         (int64_t)(100 * 256 + 3) * 8 * p->batch;
}

int effpot_ctx_create(const effpot_params* p, int device, effpot_ctx** out) {
  if (!p || !out) return EFFPOT_ERR_ARG;
  std::lock_guard<std::recursive_mutex> lk(g_api_mu);
  *out = nullptr;
  {
    const int n3[3] = {p->nx, p->ny, p->nz};
    for (int a = 0; a < 3; a++) {
      if (tstgrid(n3[a])) {
        snap::g_err = "grid: n" + std::string(1, "xyz"[a]) + " = " + std::to_string(n3[a]) +
                      " is not 2^a 3^b 5^c in 16..256 (tstgrid)";
        return EFFPOT_ERR_GRID;
      }
      if (!fft_size_supported(n3[a])) {
        snap::g_err = "grid: n" + std::string(1, "xyz"[a]) + " = " + std::to_string(n3[a]) +
                      " has no FFT here; supported sizes: 16 24 32 48 64 96 128 192 256";
        return EFFPOT_ERR_GRID;
      }
    }
  }
  if (p->core != EFFPOT_CORE_OT && p->core != EFFPOT_CORE_OTC) return EFFPOT_ERR_CORE;
  if (p->npd != 13 || (p->icon != 13 && p->icon != 8) || p->batch < 1 || p->batch > 4096) return EFFPOT_ERR_ARG;
  if (p->precision != EFFPOT_PREC_FP64 && p->precision != EFFPOT_PREC_MIXED) return EFFPOT_ERR_ARG;
  effpot_ctx* ctx = new effpot_ctx();
  ctx->p = *p;
  ctx->dev = device;
  if (hipSetDevice(device) != hipSuccess) { delete ctx; return EFFPOT_ERR_DEVICE; }
  if (hipStreamCreateWithFlags(&ctx->st, hipStreamNonBlocking) != hipSuccess) { delete ctx; return EFFPOT_ERR_DEVICE; }
  if (hipStreamCreateWithFlags(&ctx->st2, hipStreamNonBlocking) != hipSuccess ||
      hipEventCreateWithFlags(&ctx->ev_fork, hipEventDisableTiming) != hipSuccess ||
      hipEventCreateWithFlags(&ctx->ev_join, hipEventDisableTiming) != hipSuccess) {
    effpot_ctx_destroy(ctx);
    return EFFPOT_ERR_DEVICE;
  }
  ctx->nx = p->nx; ctx->ny = p->ny; ctx->nz = p->nz; ctx->nxh = p->nx / 2 + 1; ctx->B = p->batch;
  ctx->nxp = (ctx->nxh + 15) / 16 * 16;
  ctx->N = (long)p->nx * p->ny * p->nz; ctx->M = (long)ctx->nxp * p->ny * p->nz;
  ctx->hx = 2.0 * std::fabs(p->xmax) / p->nx;
  ctx->hy = 2.0 * std::fabs(p->ymax) / p->ny;
  ctx->hz = 2.0 * std::fabs(p->zmax) / p->nz;
  ctx->dxyz = ctx->hx * ctx->hy * ctx->hz;
  int nmax = std::max(p->nx, std::max(p->ny, p->nz));
  double xm = std::max(p->xmax, std::max(p->ymax, p->zmax));
  ctx->npot = 100 * nmax + 1;
  ctx->delinter = 4.0 * xm / (double)(ctx->npot - 1);
  ctx->tx = p->nx % 64 == 0 ? 64 : (p->nx % 32 == 0 ? 32 : (p->nx % 16 == 0 ? 16 : 8));
  ctx->zc = std::min(16, p->nz);
  ctx->mixed = p->precision == EFFPOT_PREC_MIXED;
  ctx->cplx = p->psi_complex != 0;
  if (const char* g = std::getenv("EFFPOT_GRAPHS")) ctx->graphs = std::atoi(g) != 0;
  if (const char* g = std::getenv("EFFPOT_SPLIT")) ctx->split = std::atoi(g) != 0;
  if (const char* g = std::getenv("EFFPOT_TUNE_LOG")) ctx->tune_log = std::atoi(g) != 0;
  ctx->h_active.assign(ctx->B, 1);
  ctx->have_imp.assign(ctx->B, 0);
  int rc;
  auto bail = [&](int r) { effpot_ctx_destroy(ctx); return r; };
  if ((rc = alloc_state(ctx, ctx->s64)) != EFFPOT_OK) return bail(rc);
  if (ctx->mixed && (rc = alloc_state(ctx, ctx->s32, true)) != EFFPOT_OK) return bail(rc);
  if ((rc = dalloc(ctx, &ctx->imptab, (size_t)ctx->B * (ctx->npot + 2))) != EFFPOT_OK) return bail(rc);
  if ((rc = dalloc(ctx, &ctx->d_active, ctx->B)) != EFFPOT_OK) return bail(rc);
  if ((rc = dalloc(ctx, &ctx->d_plus, ctx->B)) != EFFPOT_OK) return bail(rc);
  ctx->h_plus.assign(ctx->B, PlusSlot{});
  ctx->h_eso.assign(ctx->B, 0.0);
  if (const char* e = std::getenv("EFFPOT_PLANE")) { ctx->plane = std::atoi(e); ctx->plane_fixed = true; }
  if (const char* e = std::getenv("EFFPOT_IPLANE")) { ctx->iplane = std::atoi(e) != 0; ctx->iplane_fixed = true; }
  if (const char* e = std::getenv("EFFPOT_PACKED")) {
    ctx->packed = std::atoi(e) != 0 ? 1 : 0;
    ctx->packed_fixed = true;
  }
  if (const char* e = std::getenv("EFFPOT_ZCONV_TH")) {
    ctx->zconv_th = std::atoi(e) == 512 ? 512 : 256;
    ctx->zconv_fixed = true;
  }
  ctx->h_imp.assign(ctx->B, effpot_impurity{});
  ctx->h_is_state.assign(ctx->B, 0);
  if ((rc = dalloc(ctx, &ctx->d_motion, ctx->B)) != EFFPOT_OK) return bail(rc);
  ctx->h_motion.assign(ctx->B, Motion{});
  if ((rc = dalloc(ctx, &ctx->d_dyn, ctx->B)) != EFFPOT_OK) return bail(rc);
  ctx->h_dyn.assign(ctx->B, InvarDyn{});
  ctx->h_sargs.assign(ctx->B, StateArgs{});
  ctx->h_sals.assign(ctx->B, 0.0);
  if (copy_sync(ctx, ctx->d_dyn, ctx->h_dyn.data(), sizeof(InvarDyn) * ctx->B, hipMemcpyHostToDevice) != hipSuccess)
    return bail(EFFPOT_ERR_DEVICE);
  if (copy_sync(ctx, ctx->d_motion, ctx->h_motion.data(), sizeof(Motion) * ctx->B, hipMemcpyHostToDevice) != hipSuccess)
    return bail(EFFPOT_ERR_DEVICE);
  if (copy_sync(ctx, ctx->d_plus, ctx->h_plus.data(), sizeof(PlusSlot) * ctx->B, hipMemcpyHostToDevice) != hipSuccess)
    return bail(EFFPOT_ERR_DEVICE);
  if ((rc = dalloc(ctx, &ctx->d_ones, ctx->B)) != EFFPOT_OK) return bail(rc);
  if ((rc = dalloc(ctx, &ctx->d_mask_tmp, ctx->B)) != EFFPOT_OK) return bail(rc);
  {  // partial sums sized for the largest block count any tile choice gives
    int tx0 = ctx->tx, zc0 = ctx->zc;
    ctx->nblk_max = 0;
    for (int tx : {64, 32, 16, 8})
      for (int zc : {4, 8, 16, 32}) {
        if (p->nx % tx) continue;
        ctx->tx = tx; ctx->zc = std::min(zc, p->nz);
        ctx->nblk_max = std::max(ctx->nblk_max, nblk(ctx));
      }
    ctx->tx = tx0; ctx->zc = zc0;
  }
  if ((rc = dalloc(ctx, &ctx->d_partial, (size_t)ctx->B * ctx->nblk_max * 16)) != EFFPOT_OK) return bail(rc);
  if ((rc = dalloc(ctx, &ctx->d_red, (size_t)ctx->B * 16)) != EFFPOT_OK) return bail(rc);
  if ((rc = dalloc(ctx, &ctx->d_scale, (size_t)ctx->B)) != EFFPOT_OK) return bail(rc);
  if ((rc = dalloc(ctx, &ctx->d_dscale, (size_t)ctx->B)) != EFFPOT_OK) return bail(rc);
  if ((rc = dalloc(ctx, &ctx->d_n4, (size_t)ctx->B)) != EFFPOT_OK) return bail(rc);
  if ((rc = dalloc(ctx, &ctx->d_dt, (size_t)ctx->B)) != EFFPOT_OK) return bail(rc);
  ctx->h_dt.assign(ctx->B, 0.0);
  ctx->mask_dev.assign((size_t)ctx->B * 7, -1);
  ctx->h_relprev.assign(ctx->B, 0.0);
  ctx->h_reltrend.assign(ctx->B, 0.0);
  ctx->dt_dev.assign(ctx->B, -1.0);
  ctx->h_n4.assign(ctx->B, p->n4);
  if (copy_sync(ctx, ctx->d_n4, ctx->h_n4.data(), sizeof(double) * ctx->B, hipMemcpyHostToDevice) != hipSuccess)
    return bail(EFFPOT_ERR_DEVICE);
  if ((rc = dalloc(ctx, &ctx->d_bad, (size_t)ctx->B)) != EFFPOT_OK) return bail(rc);
  if ((rc = dalloc(ctx, &ctx->d_mask, (size_t)ctx->B * 7)) != EFFPOT_OK) return bail(rc);
  if (hipHostMalloc((void**)&ctx->h_mask, sizeof(int) * ctx->B * 7, hipHostMallocDefault) != hipSuccess)
    return bail(EFFPOT_ERR_NOMEM);
  if (hipHostMalloc((void**)&ctx->h_pin, sizeof(double) * ctx->B * 16 + 64, hipHostMallocDefault) != hipSuccess)
    return bail(EFFPOT_ERR_NOMEM);
  std::vector<int> ones(ctx->B, 1);
  if (copy_sync(ctx, ctx->d_active, ones.data(), sizeof(int) * ctx->B, hipMemcpyHostToDevice) != hipSuccess ||
      copy_sync(ctx, ctx->d_ones, ones.data(), sizeof(int) * ctx->B, hipMemcpyHostToDevice) != hipSuccess)
    return bail(EFFPOT_ERR_DEVICE);
  // one-time k-space kernels in FP64 (fforma / initcg / kalfs)
  KernParams kp{p->nx, p->ny, p->nz, ctx->nxh, ctx->hx, ctx->hy, ctx->hz, p->eps4, p->sigma4, p->h_core, p->l,
                ctx->nxp};
  hipLaunchKernelGGL(k_kspace, dim3(512), dim3(256), 0, ctx->st, kp, ctx->s64.kern[0], ctx->s64.kern[1],
                     ctx->s64.kern[2]);
  if (hipGetLastError() != hipSuccess) return bail(EFFPOT_ERR_DEVICE);
  if (ctx->mixed) {
    for (int j = 0; j < 3; j++)
      hipLaunchKernelGGL((k_convert<double, float>), dim3(512), dim3(256), 0, ctx->st, ctx->s64.kern[j],
                         ctx->s32.kern[j], ctx->M);
  }
  // the factor pairs of the packed column, in each precision from that precision's tables
  for (int j = 0; j < 3; j++) {
    hipLaunchKernelGGL((k_pair_tables<double>), dim3(256), dim3(256), 0, ctx->st, (const double*)ctx->s64.kern[j],
                       ctx->s64.kw[j], ctx->nxp, ctx->nxh - 1, ctx->ny, ctx->nz);
    if (ctx->mixed)
      hipLaunchKernelGGL((k_pair_tables<float>), dim3(256), dim3(256), 0, ctx->st, (const float*)ctx->s32.kern[j],
                         ctx->s32.kw[j], ctx->nxp, ctx->nxh - 1, ctx->ny, ctx->nz);
  }
  if (hipGetLastError() != hipSuccess) return bail(EFFPOT_ERR_DEVICE);
  if (hipStreamSynchronize(ctx->st) != hipSuccess) return bail(EFFPOT_ERR_DEVICE);
  *out = ctx;
  return EFFPOT_OK;
}

int effpot_ctx_destroy(effpot_ctx* ctx) {
  if (!ctx) return EFFPOT_ERR_ARG;
  std::lock_guard<std::recursive_mutex> lk(g_api_mu);
  hipSetDevice(ctx->dev);
  if (ctx->st) hipStreamSynchronize(ctx->st);
  for (int k = 0; k < 2; k++) {
    if (ctx->pc_exec[k]) hipGraphExecDestroy(ctx->pc_exec[k]);
    if (ctx->pc_exec2[k]) hipGraphExecDestroy(ctx->pc_exec2[k]);
    if (ctx->pck_exec[k]) hipGraphExecDestroy(ctx->pck_exec[k]);
    if (ctx->pck_exec2[k]) hipGraphExecDestroy(ctx->pck_exec2[k]);
  }
  for (void* v : ctx->allocs) hipFree(v);
  if (ctx->h_pin) hipHostFree(ctx->h_pin);
  if (ctx->h_mask) hipHostFree(ctx->h_mask);
  if (ctx->ev_fork) hipEventDestroy(ctx->ev_fork);
  if (ctx->ev_join) hipEventDestroy(ctx->ev_join);
  if (ctx->st2) hipStreamDestroy(ctx->st2);
  if (ctx->st) hipStreamDestroy(ctx->st);
  delete ctx;
  return EFFPOT_OK;
}

const char* effpot_last_error(const effpot_ctx* ctx) { return ctx ? ctx->err.c_str() : snap::g_err.c_str(); }

int effpot_snapshot_read(const char* path, effpot_snapshot* hdr, double* psi) { return snap::read(path, hdr, psi); }

int effpot_snapshot_write(const char* path, const effpot_snapshot* hdr, const double* psi, int binary) {
  return snap::write(path, hdr, psi, binary);
}

int effpot_sync(effpot_ctx* ctx) {
  RC(check_ctx(ctx));
  HIPCHK(hipStreamSynchronize(ctx->st));
  return EFFPOT_OK;
}

int effpot_set_active(effpot_ctx* ctx, int slot, int active) {
  RC(check_ctx(ctx));
  if (slot < 0 || slot >= ctx->B) return ctx->fail(EFFPOT_ERR_ARG, "slot out of range");
  ctx->h_active[slot] = active ? 1 : 0;
  HIPCHK(hipMemcpyAsync(ctx->d_active, ctx->h_active.data(), sizeof(int) * ctx->B, hipMemcpyHostToDevice, ctx->st));
  HIPCHK(hipStreamSynchronize(ctx->st));
  return EFFPOT_OK;
}

// Reset one slot's integrator state (ring, pc, q, history) in the given precision.
template <typename T> static int reset_slot(effpot_ctx* ctx, State<T>& s, int slot) {
  const long N = ctx->N;
  HIPCHK(hipMemsetAsync(s.psi + (long)slot * 5 * N + N, 0, sizeof(cx<T>) * 4 * N, ctx->st));
  HIPCHK(hipMemsetAsync(s.hp + (long)slot * 2 * N, 0, sizeof(cx<T>) * 2 * N, ctx->st));
  HIPCHK(hipMemsetAsync(s.pc + (long)slot * N, 0, sizeof(cx<T>) * N, ctx->st));
  HIPCHK(hipMemsetAsync(s.q + (long)slot * N, 0, sizeof(cx<T>) * N, ctx->st));
  int init[8] = {0, 1, 2, 3, 4, 0, 1, 0};
  HIPCHK(hipMemcpyAsync(s.ring + 8 * slot, init, sizeof(init), hipMemcpyHostToDevice, ctx->st));
  if (ctx->s32.corr) HIPCHK(hipMemsetAsync(ctx->s32.corr + (long)slot * N, 0, sizeof(float) * N, ctx->st));
  HIPCHK(hipStreamSynchronize(ctx->st));
  return EFFPOT_OK;
}

static int den_of_slot(effpot_ctx* ctx, int slot) {
  std::vector<int> m(ctx->B, 0);
  m[slot] = 1;
  HIPCHK(hipMemcpyAsync(ctx->d_mask_tmp, m.data(), sizeof(int) * ctx->B, hipMemcpyHostToDevice, ctx->st));
  RC(den_from_cur(ctx, ctx->d_mask_tmp, false));
  if (ctx->mixed) RC(den_from_cur(ctx, ctx->d_mask_tmp, true));
  HIPCHK(hipStreamSynchronize(ctx->st));
  return EFFPOT_OK;
}

int effpot_upload_psi(effpot_ctx* ctx, int slot, const double* psi) {
  RC(check_ctx(ctx));
  if (!psi || slot < 0 || slot >= ctx->B) return ctx->fail(EFFPOT_ERR_ARG, "bad slot or pointer");
  hipSetDevice(ctx->dev);
  RC(reset_slot(ctx, ctx->s64, slot));
  cx<double>* dst = ctx->s64.psi + (long)slot * 5 * ctx->N;  // buffer 0 = CUR after reset
  const long nval = ctx->cplx ? 2 * ctx->N : ctx->N;        // doubles per state
  if (ctx->cplx) {
    HIPCHK(copy_sync(ctx, dst, psi, sizeof(double) * 2 * ctx->N, hipMemcpyHostToDevice));
  } else {
    std::vector<double> re(ctx->N);
    for (long i = 0; i < ctx->N; i++) {
      if (psi[2 * i + 1] != 0.0)
        return ctx->fail(EFFPOT_ERR_ARG, "complex wavefunction uploaded to a context with psi_complex = 0");
      re[i] = psi[2 * i];
    }
    HIPCHK(copy_sync(ctx, dst, re.data(), sizeof(double) * ctx->N, hipMemcpyHostToDevice));
  }
  (void)nval;
  return den_of_slot(ctx, slot);
}

template <typename T> static int cur_index(effpot_ctx* ctx, State<T>& s, int slot, int* idx) {
  int r[8];
  HIPCHK(hipMemcpyAsync(r, s.ring + 8 * slot, sizeof(r), hipMemcpyDeviceToHost, ctx->st));
  HIPCHK(hipStreamSynchronize(ctx->st));
  *idx = r[R_CUR];
  return EFFPOT_OK;
}

// Which precision holds the live state of a slot: FP64 unless this is a mixed
// context that has not been polished (the FP32 phase is the live one).

int effpot_download_psi(effpot_ctx* ctx, int slot, double* psi) {
  RC(check_ctx(ctx));
  if (!psi || slot < 0 || slot >= ctx->B) return ctx->fail(EFFPOT_ERR_ARG, "bad slot or pointer");
  hipSetDevice(ctx->dev);
  int idx;
  RC(cur_index(ctx, ctx->s64, slot, &idx));
  const cx<double>* src = ctx->s64.psi + ((long)slot * 5 + idx) * ctx->N;
  if (ctx->cplx) {
    HIPCHK(copy_sync(ctx, psi, src, sizeof(double) * 2 * ctx->N, hipMemcpyDeviceToHost));
  } else {
    std::vector<double> re(ctx->N);
    HIPCHK(copy_sync(ctx, re.data(), src, sizeof(double) * ctx->N, hipMemcpyDeviceToHost));
    for (long i = 0; i < ctx->N; i++) { psi[2 * i] = re[i]; psi[2 * i + 1] = 0.0; }
  }
  return EFFPOT_OK;
}

int effpot_download_den(effpot_ctx* ctx, int slot, double* den) {
  RC(check_ctx(ctx));
  if (!den || slot < 0 || slot >= ctx->B) return ctx->fail(EFFPOT_ERR_ARG, "bad slot or pointer");
  HIPCHK(hipMemcpyAsync(den, ctx->s64.den + (long)slot * ctx->N, sizeof(double) * ctx->N, hipMemcpyDeviceToHost, ctx->st));
  HIPCHK(hipStreamSynchronize(ctx->st));
  return EFFPOT_OK;
}

int effpot_download_field(effpot_ctx* ctx, int slot, int which, double* out) {
  RC(check_ctx(ctx));
  if (!out || slot < 0 || slot >= ctx->B) return ctx->fail(EFFPOT_ERR_ARG, "bad slot or pointer");
  State<double>& s = ctx->s64;
  double* src[13] = {s.pot4, s.dencg, s.delj4, s.denalf, s.ualphas, s.uimp, s.dx, s.dy, s.dz, s.sto1, s.intx, s.inty, s.intz};
  if (which < 0 || which > 12) return ctx->fail(EFFPOT_ERR_ARG, "unknown field");
  HIPCHK(hipMemcpyAsync(out, src[which] + (long)slot * ctx->N, sizeof(double) * ctx->N, hipMemcpyDeviceToHost, ctx->st));
  HIPCHK(hipStreamSynchronize(ctx->st));
  if (which == 3)  // the device keeps falfs = 1 - denalf/den0s
    for (long i = 0; i < ctx->N; i++) out[i] = (1.0 - out[i]) * ctx->p.den0s;
  return EFFPOT_OK;
}

int effpot_poten_fp32(effpot_ctx* ctx) {
  RC(check_ctx(ctx));
  if (!ctx->mixed) return ctx->fail(EFFPOT_ERR_ARG, "FP32 fields exist in a mixed-precision context only");
  hipSetDevice(ctx->dev);
  RC(den_from_cur(ctx, ctx->d_active, true));
  RC(poten_t(ctx, ctx->s32, ctx->d_active, -1, nullptr, true));  // the gradient stored for downloads
  HIPCHK(hipStreamSynchronize(ctx->st));
  return EFFPOT_OK;
}

int effpot_download_field_fp32(effpot_ctx* ctx, int slot, int which, float* out) {
  RC(check_ctx(ctx));
  if (!ctx->mixed) return ctx->fail(EFFPOT_ERR_ARG, "FP32 fields exist in a mixed-precision context only");
  if (!out || slot < 0 || slot >= ctx->B) return ctx->fail(EFFPOT_ERR_ARG, "bad slot or pointer");
  State<float>& s = ctx->s32;
  float* src[14] = {s.pot4, s.dencg, s.delj4, s.denalf, s.ualphas, s.uimp, s.dx, s.dy, s.dz,
                    s.sto1, s.intx, s.inty, s.intz, s.corr};
  if (which < 0 || which > 13) return ctx->fail(EFFPOT_ERR_ARG, "unknown field");
  HIPCHK(hipMemcpyAsync(out, src[which] + (long)slot * ctx->N, sizeof(float) * ctx->N, hipMemcpyDeviceToHost, ctx->st));
  HIPCHK(hipStreamSynchronize(ctx->st));
  if (which == 3)  // the device keeps falfs = 1 - denalf/den0s
    for (long i = 0; i < ctx->N; i++) out[i] = (float)((1.0 - (double)out[i]) * ctx->p.den0s);
  return EFFPOT_OK;
}

// Counter-based uniform in [0,1): splitmix64 finaliser of the seed and index.
static uint64_t mix64(uint64_t z) {
  z += 0x9E3779B97F4A7C15ULL;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}

int effpot_init_fermi(effpot_ctx* ctx, int slot, double rho0, double R, double afermi, double noise, uint64_t seed,
                      const double centre[3]) {
  RC(check_ctx(ctx));
  if (slot < 0 || slot >= ctx->B || !centre || !(afermi > 0)) return ctx->fail(EFFPOT_ERR_ARG, "bad Fermi arguments");
  const int nx = ctx->nx, ny = ctx->ny, nz = ctx->nz;
  std::vector<double> rho(ctx->N), psi(2 * ctx->N);
  double sum = 0.0;
  for (int iz = 0; iz < nz; iz++)
    for (int iy = 0; iy < ny; iy++)
      for (int ix = 0; ix < nx; ix++) {
        long i = (long)ix + (long)nx * ((long)iy + (long)ny * iz);
        double dx = (-ctx->p.xmax + ctx->hx * ix) - centre[0];
        double dy = (-ctx->p.ymax + ctx->hy * iy) - centre[1];
        double dz = (-ctx->p.zmax + ctx->hz * iz) - centre[2];
        double r = std::sqrt(dx * dx + dy * dy + dz * dz);
        double u = (double)(mix64(seed ^ mix64((uint64_t)i)) >> 11) * 0x1.0p-53;
        double v = rho0 / (1.0 + std::exp((r - R) / afermi)) + noise * u;
        rho[i] = v;
        sum += v;
      }
  double f = std::sqrt(ctx->h_n4[slot] / (sum * ctx->dxyz));
  for (long i = 0; i < ctx->N; i++) { psi[2 * i] = std::sqrt(rho[i]) * f; psi[2 * i + 1] = 0.0; }
  return effpot_upload_psi(ctx, slot, psi.data());
}

int effpot_set_dt(effpot_ctx* ctx, int slot, double dt) {
  RC(check_ctx(ctx));
  if (slot < 0 || slot >= ctx->B || dt < 0 || !std::isfinite(dt)) return ctx->fail(EFFPOT_ERR_DT, "bad slot time step");
  ctx->h_dt[slot] = dt;
  return EFFPOT_OK;
}

int effpot_set_natoms(effpot_ctx* ctx, int slot, double n4) {
  RC(check_ctx(ctx));
  if (slot < 0 || slot >= ctx->B || !(n4 > 0)) return ctx->fail(EFFPOT_ERR_ARG, "bad slot or atom number");
  ctx->h_n4[slot] = n4;
  HIPCHK(hipStreamSynchronize(ctx->st));  // kernels in flight read d_n4
  HIPCHK(copy_sync(ctx, ctx->d_n4 + slot, &n4, sizeof(double), hipMemcpyHostToDevice));
  return EFFPOT_OK;
}

int effpot_set_impurity(effpot_ctx* ctx, int slot, const effpot_impurity* imp) {
  RC(check_ctx(ctx));
  if (!imp || slot < 0 || slot >= ctx->B) return ctx->fail(EFFPOT_ERR_ARG, "bad slot or pointer");
  if (imp->kind != EFFPOT_POT_MORSE && imp->kind != EFFPOT_POT_NULL)
    return ctx->fail(EFFPOT_ERR_POT, "unknown impurity potential");
  hipSetDevice(ctx->dev);
  ImpParams ip{imp->kind, imp->D, imp->re, imp->a, imp->r_cutoff, imp->umax, imp->cap, imp->rimp[0], imp->rimp[1], imp->rimp[2]};
  double* tab = ctx->imptab + (long)slot * (ctx->npot + 2);
  hipLaunchKernelGGL(k_imp_table, dim3((ctx->npot + 257) / 256), dim3(256), 0, ctx->st, ip, ctx->npot, ctx->delinter, tab);
  LAUNCHCHK();
  const double rmax = (double)(ctx->npot - 1) * ctx->delinter;
  hipLaunchKernelGGL((k_imp_field<double>), dim3(1024), dim3(256), 0, ctx->st, ip, ctx->nx, ctx->ny, ctx->nz,
                     -ctx->p.xmax, -ctx->p.ymax, -ctx->p.zmax, ctx->hx, ctx->hy, ctx->hz, ctx->npot, rmax, (const double*)tab,
                     ctx->s64.uimp + (long)slot * ctx->N);
  LAUNCHCHK();
  if (ctx->mixed) {
    hipLaunchKernelGGL((k_convert<double, float>), dim3(1024), dim3(256), 0, ctx->st,
                       (const double*)(ctx->s64.uimp + (long)slot * ctx->N), ctx->s32.uimp + (long)slot * ctx->N, ctx->N);
    LAUNCHCHK();
  }
  PlusSlot& ps = ctx->h_plus[slot];
  for (int i = 0; i < 3; i++) ps.rimp[i] = imp->rimp[i];
  HIPCHK(hipMemcpyAsync(ctx->d_plus + slot, &ps, sizeof(PlusSlot), hipMemcpyHostToDevice, ctx->st));
  HIPCHK(hipStreamSynchronize(ctx->st));
  ctx->have_imp[slot] = 1;
  ctx->h_eso[slot] = 0.0;
  ctx->h_imp[slot] = *imp;
  ctx->h_is_state[slot] = 0;
  RC(stop_dyn(ctx, slot));
  return stop_motion(ctx, slot);
}

int effpot_instates_select(int state, const double* moments, double als, int instate, const effpot_select* sel,
                           double* invar, double* eigen, double* eso) {
  if (state != EFFPOT_STATE_P && state != EFFPOT_STATE_D) return EFFPOT_ERR_ARG;
  const int n = state == EFFPOT_STATE_P ? 6 : 10;
  if (!moments || instate >= n / 2 || instate < -1) return EFFPOT_ERR_ARG;
  const int mode = sel ? sel->mode : EFFPOT_SELECT_PAIR;
  if (mode < EFFPOT_SELECT_PAIR || mode > EFFPOT_SELECT_MIXTURE) return EFFPOT_ERR_ARG;
  if (mode >= EFFPOT_SELECT_FIX && state != EFFPOT_STATE_P) return EFFPOT_ERR_ARG;
  if (mode == EFFPOT_SELECT_MIXTURE && !(sel->a_pi_3o2 >= 0.0 && sel->a_sig_1o2 >= 0.0)) return EFFPOT_ERR_ARG;
  zd H[100], V[100];
  double w[10];
  state_matrix(state, moments, als, H);
  herm_eigen(n, H, w, V);
  if (eigen) std::memcpy(eigen, w, sizeof(double) * n);
  if (instate >= 0 && invar) {
    switch (mode) {
      case EFFPOT_SELECT_JZ: RC(state_select_jz(state, n, V, instate, sel->jz2, invar)); break;
      case EFFPOT_SELECT_FIX: state_select_fix(instate, invar); break;
      case EFFPOT_SELECT_AVERAGE: state_select_average(n, V, invar); break;
      case EFFPOT_SELECT_MIXTURE: state_select_mixture(n, V, sel->a_pi_3o2, sel->a_sig_1o2, invar); break;
      default: state_select(n, V, instate, invar);
    }
  }
  if (eso && invar) *eso = state_eso(state, invar, als);
  return EFFPOT_OK;
}

int effpot_instates(int state, const double* moments, double als, int instate, double* invar, double* eigen,
                    double* eso) {
  return effpot_instates_select(state, moments, als, instate, nullptr, invar, eigen, eso);
}

int effpot_exciplex_moments(const effpot_plus* v_pi, const effpot_plus* v_sigma, int n_he, int axis, double r_exc,
                            double* moments) {
  if (!v_pi || !v_sigma || !moments || (n_he != 2 && n_he != 6) || axis < 0 || axis > 2 || !(r_exc > 0.0))
    return EFFPOT_ERR_ARG;
  for (const effpot_plus* v : {v_pi, v_sigma})
    if ((v->kind != EFFPOT_POT_NULL && v->kind != EFFPOT_POT_SERIES && v->kind != EFFPOT_POT_KOUTSELOS) ||
        (v->kind == EFFPOT_POT_SERIES && (v->nterms < 1 || v->nterms > EFFPOT_SERIES_MAX || v->k0 < 0)))
      return EFFPOT_ERR_POT;
  // the two axes of the ring plane, lower first (instates.f90:72-81)
  const int ax = axis == 0 ? 1 : 0, ay = axis == 2 ? 1 : 2;
  double m[7] = {};
  for (int i = 0; i < n_he; i++) {
    double r3[3] = {0.0, 0.0, 0.0};
    if (n_he == 2) {
      r3[axis] = i == 0 ? r_exc : -r_exc;
    } else {
      const double t = (double)i * (2.0 * M_PI / (double)n_he);
      r3[ax] = r_exc * std::cos(t);
      r3[ay] = r_exc * std::sin(t);
    }
    const double r2 = r3[0] * r3[0] + r3[1] * r3[1] + r3[2] * r3[2], r = std::sqrt(r2);
    const double vpi = plus_pot(*v_pi, r), dv = plus_pot(*v_sigma, r) - vpi;
    for (int c = 0; c < 3; c++) m[1 + c] += vpi + dv * (r3[c] / r) * (r3[c] / r);
    m[4] += dv * r3[0] * r3[1] / r2;
    m[5] += dv * r3[0] * r3[2] / r2;
    m[6] += dv * r3[1] * r3[2] / r2;
  }
  std::memcpy(moments, m, sizeof(m));
  return EFFPOT_OK;
}

int effpot_set_impurity_state(effpot_ctx* ctx, int slot, effpot_state_imp* imp) {
  RC(check_ctx(ctx));
  if (!imp || slot < 0 || slot >= ctx->B) return ctx->fail(EFFPOT_ERR_ARG, "bad slot or pointer");
  const int st = imp->state;
  if (st != EFFPOT_STATE_P && st != EFFPOT_STATE_D) return ctx->fail(EFFPOT_ERR_ARG, "state must be P or D");
  const int n = st == EFFPOT_STATE_P ? 6 : 10;
  if (imp->instate < -1 || imp->instate >= n / 2)
    return ctx->fail(EFFPOT_ERR_ARG, "instate out of range (P: 0..2, D: 0..4, or -1 for a given invar)");
  for (const effpot_plus* v : {&imp->v_pi, &imp->v_sigma, &imp->v_delta})
    if ((v->kind != EFFPOT_POT_NULL && v->kind != EFFPOT_POT_SERIES && v->kind != EFFPOT_POT_KOUTSELOS) ||
        (v->kind == EFFPOT_POT_SERIES && (v->nterms < 1 || v->nterms > EFFPOT_SERIES_MAX || v->k0 < 0)))
      return ctx->fail(EFFPOT_ERR_POT, "bad state potential");
  hipSetDevice(ctx->dev);
  const long L = ctx->npot + 2;
  if (!ctx->d_stab) {
    int rc;
    if ((rc = dalloc(ctx, &ctx->d_stab, (size_t)ctx->B * 3 * L)) != EFFPOT_OK) return rc;
    if ((rc = dalloc(ctx, &ctx->d_smom, (size_t)kStateBlocks * kMomD)) != EFFPOT_OK) return rc;
  }
  double* tab = ctx->d_stab + (long)slot * 3 * L;
  const dim3 tg((unsigned)((L + 255) / 256)), tb(256);
  const double dl = ctx->delinter;
  if (st == EFFPOT_STATE_P) {  // potpi, potdel = V_sigma - V_pi
    hipLaunchKernelGGL(k_state_table, tg, tb, 0, ctx->st, imp->v_pi, imp->v_pi, 0, ctx->npot, dl, tab);
    hipLaunchKernelGGL(k_state_table, tg, tb, 0, ctx->st, imp->v_pi, imp->v_sigma, 1, ctx->npot, dl, tab + L);
  } else {  // potdel = V_delta, potpidel = V_pi - V_delta, potsigdel = V_sigma - V_delta
    hipLaunchKernelGGL(k_state_table, tg, tb, 0, ctx->st, imp->v_delta, imp->v_delta, 0, ctx->npot, dl, tab);
    hipLaunchKernelGGL(k_state_table, tg, tb, 0, ctx->st, imp->v_delta, imp->v_pi, 1, ctx->npot, dl, tab + L);
    hipLaunchKernelGGL(k_state_table, tg, tb, 0, ctx->st, imp->v_delta, imp->v_sigma, 1, ctx->npot, dl, tab + 2 * L);
  }
  LAUNCHCHK();
  StateArgs a{};
  a.state = st; a.nx = ctx->nx; a.ny = ctx->ny; a.nz = ctx->nz;
  a.x0 = -ctx->p.xmax; a.y0 = -ctx->p.ymax; a.z0 = -ctx->p.zmax; a.hx = ctx->hx; a.hy = ctx->hy; a.hz = ctx->hz;
  a.npot = ctx->npot;
  a.h = ((double)(ctx->npot - 1) * dl) / (double)(ctx->npot - 1);
  a.c1oh = 1.0 / a.h; a.c1o2h2 = a.c1oh * a.c1oh * 0.5;
  a.tab = tab; a.umax_pi = imp->v_pi.umax; a.umax_sigma = imp->v_sigma.umax;
  a.rx = imp->rimp[0]; a.ry = imp->rimp[1]; a.rz = imp->rimp[2];
  a.cap = 1000.0;  // potenimp.f90:184-186
  const bool fixed = imp->select.fixed_moments != 0;
  if (fixed && st != EFFPOT_STATE_P) return ctx->fail(EFFPOT_ERR_ARG, "fixed moments (exciplex) are for P states");
  if (!fixed) std::memset(imp->moments, 0, sizeof(imp->moments));
  if (imp->instate >= 0 && !fixed) {  // instates on the slot's current density
    const int K = st == EFFPOT_STATE_P ? kMomP : kMomD;
    // the slot's current psi buffer (ring roles live on the device)
    std::vector<int> ring(8);
    HIPCHK(hipStreamSynchronize(ctx->st));
    HIPCHK(copy_sync(ctx, ring.data(), ctx->s64.ring + 8 * slot, sizeof(int) * 8, hipMemcpyDeviceToHost));
    const long off = ((long)slot * 5 + ring[R_CUR]) * ctx->N;  // in complex elements, as Ring::p
#define M_(KK)                                                                                                 \
  if (ctx->cplx)                                                                                             \
    hipLaunchKernelGGL((k_state_moments<cx<double>, KK>), dim3(kStateBlocks), dim3(256), 0, ctx->st,           \
                       (const cx<double>*)(ctx->s64.psi + off), a, ctx->d_smom);                              \
  else                                                                                                       \
    hipLaunchKernelGGL((k_state_moments<double, KK>), dim3(kStateBlocks), dim3(256), 0, ctx->st,              \
                       (const double*)(ctx->s64.psi + off), a, ctx->d_smom)
    if (K == kMomP) { M_(kMomP); } else { M_(kMomD); }
#undef M_
    LAUNCHCHK();
    std::vector<double> part((size_t)kStateBlocks * K);
    HIPCHK(hipMemcpyAsync(part.data(), ctx->d_smom, sizeof(double) * part.size(), hipMemcpyDeviceToHost, ctx->st));
    HIPCHK(hipStreamSynchronize(ctx->st));
    for (int k = 0; k < K; k++) {
      double acc = 0.0;
      for (int b = 0; b < kStateBlocks; b++) acc += part[(size_t)b * K + k];
      imp->moments[k] = acc * ctx->dxyz;
    }
  }
  {
    const int rc = effpot_instates_select(st, imp->moments, imp->als, imp->instate, &imp->select, imp->invar,
                                          imp->eigen, &imp->eso);
    if (rc != EFFPOT_OK) return ctx->fail(rc, "bad invar selection (mode, Ljz or mixture amplitudes)");
  }
  if (imp->instate < 0) imp->eso = state_eso(st, imp->invar, imp->als);
  std::memcpy(a.inv, imp->invar, sizeof(a.inv));
  hipLaunchKernelGGL((k_state_uimp<double>), dim3(1024), dim3(256), 0, ctx->st, a, ctx->s64.uimp + (long)slot * ctx->N);
  LAUNCHCHK();
  if (ctx->mixed) {
    hipLaunchKernelGGL((k_convert<double, float>), dim3(1024), dim3(256), 0, ctx->st,
                       (const double*)(ctx->s64.uimp + (long)slot * ctx->N), ctx->s32.uimp + (long)slot * ctx->N, ctx->N);
    LAUNCHCHK();
  }
  PlusSlot& ps = ctx->h_plus[slot];
  for (int i = 0; i < 3; i++) ps.rimp[i] = imp->rimp[i];
  HIPCHK(hipMemcpyAsync(ctx->d_plus + slot, &ps, sizeof(PlusSlot), hipMemcpyHostToDevice, ctx->st));
  HIPCHK(hipStreamSynchronize(ctx->st));
  ctx->h_eso[slot] = imp->eso;
  ctx->have_imp[slot] = 1;
  ctx->h_is_state[slot] = 1;
  ctx->h_sargs[slot] = a;
  ctx->h_sals[slot] = imp->als;
  RC(stop_dyn(ctx, slot));
  return stop_motion(ctx, slot);
}

int effpot_set_state_dynamics(effpot_ctx* ctx, int slot, int on, double ev0) {
  RC(check_ctx(ctx));
  if (slot < 0 || slot >= ctx->B) return ctx->fail(EFFPOT_ERR_ARG, "bad slot");
  if (!on) return stop_dyn(ctx, slot);
  if (ctx->mixed) return ctx->fail(EFFPOT_ERR_ARG, "invar dynamics need an FP64 context");
  if (!ctx->h_is_state[slot]) return ctx->fail(EFFPOT_ERR_STATE, "set a P/D-state impurity first");
  hipSetDevice(ctx->dev);
  HIPCHK(hipStreamSynchronize(ctx->st));
  std::vector<int> one(ctx->B, 0);
  one[slot] = 1;
  HIPCHK(copy_sync(ctx, ctx->d_mask_tmp, one.data(), sizeof(int) * ctx->B, hipMemcpyHostToDevice));
  RC(den_from_cur(ctx, ctx->d_mask_tmp, false));
  RC(poten_t(ctx, ctx->s64, ctx->d_mask_tmp));
  if (ev0 == 0.0) {  // Ev0 = eHeX + eso of the starting state (DFT4He3d.f90:444)
    std::vector<effpot_energies> e(ctx->B);
    RC(energy_t(ctx, ctx->s64, ctx->d_mask_tmp, e.data()));
    ev0 = e[slot].eHeX + e[slot].eso;
  }
  const StateArgs& a = ctx->h_sargs[slot];
  InvarDyn q{};
  q.on = 1;
  q.state = a.state;
  q.n = a.state == EFFPOT_STATE_P ? 6 : 10;
  q.als = ctx->h_sals[slot];
  q.ev0 = ev0;
  for (int k = 0; k < 20; k++) q.iv[k] = a.inv[k];
  for (int k = 0; k < 3; k++) q.iiv[k] = k;
  q.ihiv[0] = 0; q.ihiv[1] = 1;
  q.sa = a;
  ctx->h_dyn[slot] = q;
  HIPCHK(copy_sync(ctx, ctx->d_dyn + slot, &q, sizeof(InvarDyn), hipMemcpyHostToDevice));
  ctx->any_dyn = 1;
  RC(dyn_hiv(ctx, ctx->d_mask_tmp));  // the initial Hinvar (DFT4He3d.f90:454)
  HIPCHK(hipStreamSynchronize(ctx->st));
  return EFFPOT_OK;
}

int effpot_get_invar(effpot_ctx* ctx, int slot, double* invar) {
  RC(check_ctx(ctx));
  if (slot < 0 || slot >= ctx->B || !invar) return ctx->fail(EFFPOT_ERR_ARG, "bad slot or pointer");
  hipSetDevice(ctx->dev);
  HIPCHK(hipStreamSynchronize(ctx->st));
  InvarDyn q;
  HIPCHK(copy_sync(ctx, &q, ctx->d_dyn + slot, sizeof(InvarDyn), hipMemcpyDeviceToHost));
  const double* src = q.on ? q.iv : ctx->h_sargs[slot].inv;
  for (int k = 0; k < 20; k++) invar[k] = src[k];
  return EFFPOT_OK;
}

int effpot_set_real_time(effpot_ctx* ctx, int on) {
  RC(check_ctx(ctx));
  if (on && (!ctx->cplx || ctx->mixed))
    return ctx->fail(EFFPOT_ERR_ARG, "real time needs complex psi (psi_complex = 1) and FP64 precision");
  hipSetDevice(ctx->dev);
  HIPCHK(hipStreamSynchronize(ctx->st));
  ctx->rt = on ? 1 : 0;  // effpot_solve / effpot_tick (the captured graphs) refuse real time
  return EFFPOT_OK;
}

int effpot_set_impurity_motion(effpot_ctx* ctx, int slot, const double* vimp, double mass_u) {
  RC(check_ctx(ctx));
  if (slot < 0 || slot >= ctx->B) return ctx->fail(EFFPOT_ERR_ARG, "bad slot");
  if (!(mass_u > 0)) return stop_motion(ctx, slot);
  if (!vimp) return ctx->fail(EFFPOT_ERR_ARG, "null velocity");
  if (ctx->mixed) return ctx->fail(EFFPOT_ERR_ARG, "a moving impurity needs an FP64 context");
  if (!ctx->have_imp[slot]) return ctx->fail(EFFPOT_ERR_STATE, "set an impurity first");
  if (ctx->h_is_state[slot] && !ctx->h_dyn[slot].on)
    return ctx->fail(EFFPOT_ERR_STATE, "a P/D-state impurity moves with its invar: effpot_set_state_dynamics first");
  hipSetDevice(ctx->dev);
  HIPCHK(hipStreamSynchronize(ctx->st));
  Motion q{};
  q.on = 1;
  q.iso = ctx->h_is_state[slot] ? 0 : 1;
  q.m = mass_u * 0.0207569277;  // mp_u, modules.f90:333
  const effpot_impurity& im = ctx->h_imp[slot];
  const StateArgs& sa = ctx->h_sargs[slot];
  const double r0[3] = {q.iso ? im.rimp[0] : sa.rx, q.iso ? im.rimp[1] : sa.ry, q.iso ? im.rimp[2] : sa.rz};
  for (int c = 0; c < 3; c++) { q.r[c] = r0[c]; q.v[c] = vimp[c]; }
  for (int k = 0; k < 3; k++) { q.ir[k] = k; q.iv[k] = k; }
  q.ia[0] = 0; q.ia[1] = 1;
  q.ip = ImpParams{im.kind, im.D, im.re, im.a, im.r_cutoff, im.umax, im.cap, im.rimp[0], im.rimp[1], im.rimp[2]};
  ctx->h_motion[slot] = q;
  HIPCHK(copy_sync(ctx, ctx->d_motion + slot, &q, sizeof(Motion), hipMemcpyHostToDevice));
  ctx->any_motion = 1;
  // the initial acceleration: poten and forceimp (DFT4He3d.f90:437-453)
  std::vector<int> one(ctx->B, 0);
  one[slot] = 1;
  HIPCHK(copy_sync(ctx, ctx->d_mask_tmp, one.data(), sizeof(int) * ctx->B, hipMemcpyHostToDevice));
  RC(den_from_cur(ctx, ctx->d_mask_tmp, false));
  RC(poten_t(ctx, ctx->s64, ctx->d_mask_tmp));
  RC(force_t(ctx, ctx->d_mask_tmp));
  HIPCHK(hipStreamSynchronize(ctx->st));
  return EFFPOT_OK;
}

int effpot_get_impurity(effpot_ctx* ctx, int slot, double* r, double* v, double* a) {
  RC(check_ctx(ctx));
  if (slot < 0 || slot >= ctx->B) return ctx->fail(EFFPOT_ERR_ARG, "bad slot");
  hipSetDevice(ctx->dev);
  HIPCHK(hipStreamSynchronize(ctx->st));
  Motion q;
  HIPCHK(copy_sync(ctx, &q, ctx->d_motion + slot, sizeof(Motion), hipMemcpyDeviceToHost));
  for (int c = 0; c < 3; c++) {
    if (r) r[c] = q.on ? q.r[c] : ctx->h_plus[slot].rimp[c];
    if (v) v[c] = q.on ? q.v[c] : 0.0;
    if (a) a[c] = q.on ? q.a[c] : 0.0;
  }
  return EFFPOT_OK;
}

int effpot_set_plus(effpot_ctx* ctx, int slot, const effpot_plus* plus) {
  RC(check_ctx(ctx));
  if (slot < 0 || slot >= ctx->B) return ctx->fail(EFFPOT_ERR_ARG, "bad slot");
  PlusSlot& ps = ctx->h_plus[slot];
  if (plus) {
    if (plus->kind != EFFPOT_POT_NULL && plus->kind != EFFPOT_POT_SERIES && plus->kind != EFFPOT_POT_KOUTSELOS)
      return ctx->fail(EFFPOT_ERR_POT, "unknown X+ potential kind");
    if (plus->kind == EFFPOT_POT_SERIES &&
        (plus->nterms < 1 || plus->nterms > EFFPOT_SERIES_MAX || plus->k0 < 0 || plus->k0 > 64))
      return ctx->fail(EFFPOT_ERR_ARG, "X+ series: nterms in 1..16 and k0 in 0..64");
    ps.p = *plus;
    ps.on = plus->kind != EFFPOT_POT_NULL;
  } else {
    std::memset(&ps.p, 0, sizeof(ps.p));
    ps.on = 0;
  }
  hipSetDevice(ctx->dev);
  HIPCHK(hipStreamSynchronize(ctx->st));  // an energy launch may still read the slot's entry
  HIPCHK(hipMemcpyAsync(ctx->d_plus + slot, &ps, sizeof(PlusSlot), hipMemcpyHostToDevice, ctx->st));
  HIPCHK(hipStreamSynchronize(ctx->st));
  return EFFPOT_OK;
}

int effpot_set_exclusion(effpot_ctx* ctx, int slot, double r) {
  RC(check_ctx(ctx));
  if (slot < 0 || slot >= ctx->B) return ctx->fail(EFFPOT_ERR_ARG, "bad slot");
  if (!(r >= 0.0) || !std::isfinite(r)) return ctx->fail(EFFPOT_ERR_ARG, "exclusion radius must be finite and >= 0");
  PlusSlot& ps = ctx->h_plus[slot];
  ps.excl = r;
  hipSetDevice(ctx->dev);
  HIPCHK(hipStreamSynchronize(ctx->st));  // an energy launch may still read the slot's entry
  HIPCHK(hipMemcpyAsync(ctx->d_plus + slot, &ps, sizeof(PlusSlot), hipMemcpyHostToDevice, ctx->st));
  HIPCHK(hipStreamSynchronize(ctx->st));
  return EFFPOT_OK;
}

int effpot_set_force_filter(effpot_ctx* ctx, int slot, int on) {
  RC(check_ctx(ctx));
  if (slot < 0 || slot >= ctx->B) return ctx->fail(EFFPOT_ERR_ARG, "bad slot");
  PlusSlot& ps = ctx->h_plus[slot];
  ps.force_filter = on ? 1 : 0;
  hipSetDevice(ctx->dev);
  HIPCHK(hipStreamSynchronize(ctx->st));
  HIPCHK(hipMemcpyAsync(ctx->d_plus + slot, &ps, sizeof(PlusSlot), hipMemcpyHostToDevice, ctx->st));
  HIPCHK(hipStreamSynchronize(ctx->st));
  return EFFPOT_OK;
}

int effpot_plus_preset(const char* selec, double r_cutoff, double umax, effpot_plus* out) {
  if (!selec || !out) return EFFPOT_ERR_ARG;
  return plus_preset(selec, r_cutoff, umax, out) ? EFFPOT_OK : EFFPOT_ERR_POT;
}

double effpot_plus_eval(const effpot_plus* plus, double r) { return plus ? plus_pot(*plus, r) : 0.0; }

int effpot_poten(effpot_ctx* ctx) {
  RC(check_ctx(ctx));
  hipSetDevice(ctx->dev);
  RC(poten_t(ctx, ctx->s64, ctx->d_active));
  return EFFPOT_OK;
}

int effpot_energy(effpot_ctx* ctx, effpot_energies* out) {
  RC(check_ctx(ctx));
  if (!out) return ctx->fail(EFFPOT_ERR_ARG, "null output");
  hipSetDevice(ctx->dev);
  return energy_t(ctx, ctx->s64, ctx->d_active, out);
}

int effpot_step_rk(effpot_ctx* ctx, double dt) {
  RC(check_ctx(ctx));
  if (!(dt > 0)) return ctx->fail(EFFPOT_ERR_DT, "non-positive time step");
  hipSetDevice(ctx->dev);
  RC(push_dt(ctx, dt));
  return step_rk_t(ctx, ctx->s64, ctx->d_active);
}

int effpot_step_pc(effpot_ctx* ctx, double dt, double* err_he, double* err_imp, double* err_vimp, double* err_iv) {
  RC(check_ctx(ctx));
  if (!(dt > 0)) return ctx->fail(EFFPOT_ERR_DT, "non-positive time step");
  hipSetDevice(ctx->dev);
  RC(push_dt(ctx, dt));
  RC(step_pc_t(ctx, ctx->s64, ctx->d_active));
  if (err_he) {
    HIPCHK(hipMemcpyAsync(ctx->h_pin, ctx->d_red, sizeof(double) * ctx->B * 2, hipMemcpyDeviceToHost, ctx->st));
    HIPCHK(hipStreamSynchronize(ctx->st));
    for (int b = 0; b < ctx->B; b++) err_he[b] = ctx->h_pin[2 * b] / (double)ctx->N;
  }
  if (err_imp || err_vimp || err_iv) {
    // steppc.f90:179,193,213-220, from the corrector differences the step left
    // in the slots' motion and invar records (0 for a fixed impurity, a frozen invar)
    HIPCHK(hipStreamSynchronize(ctx->st));
    std::vector<Motion> mo(ctx->B);
    std::vector<InvarDyn> dy(ctx->B);
    HIPCHK(copy_sync(ctx, mo.data(), ctx->d_motion, sizeof(Motion) * ctx->B, hipMemcpyDeviceToHost));
    HIPCHK(copy_sync(ctx, dy.data(), ctx->d_dyn, sizeof(InvarDyn) * ctx->B, hipMemcpyDeviceToHost));
    const double c9 = 9.0 / 121.0;
    for (int b = 0; b < ctx->B; b++) {
      double er = 0.0, ev = 0.0, ei = 0.0;
      if (mo[b].on) {
        for (int c = 0; c < 3; c++) { er += std::fabs(c9 * mo[b].pcr[c]); ev += std::fabs(c9 * mo[b].pcv[c]); }
        er *= 0.3333333333;
        ev *= 0.3333333333;
      }
      if (dy[b].on) {
        for (int k = 0; k < dy[b].n; k++) ei += c9 * std::hypot(dy[b].pciv[2 * k], dy[b].pciv[2 * k + 1]);
        ei *= dy[b].state == EFFPOT_STATE_P ? 0.1666666666666 : 0.1;
      }
      if (err_imp) err_imp[b] = er;
      if (err_vimp) err_vimp[b] = ev;
      if (err_iv) err_iv[b] = ei;
    }
  }
  return EFFPOT_OK;
}

int effpot_renormalize(effpot_ctx* ctx) {
  RC(check_ctx(ctx));
  hipSetDevice(ctx->dev);
  RC(norm_t(ctx, ctx->s64, ctx->d_active));
  RC(renorm_t(ctx, ctx->s64, ctx->d_active));
  HIPCHK(hipStreamSynchronize(ctx->st));
  return EFFPOT_OK;
}

int effpot_renormalize_noalloc(double* psi, double* den, int64_t npoints, int nslots, int psi_complex, double dxyz,
                               const double* n4, double* natoms_out, void* stream) {
  return renormalize_noalloc(psi, den, npoints, nslots, psi_complex, dxyz, n4, natoms_out, stream, snap::g_err);
}

int effpot_derivnd(effpot_ctx* ctx, int id, int iv, const double* f, double* df) {
  RC(check_ctx(ctx));
  if (!f || !df || id < 1 || id > 2 || iv < 1 || iv > 3) return ctx->fail(EFFPOT_ERR_ARG, "bad derivnd arguments");
  hipSetDevice(ctx->dev);
  State<double>& s = ctx->s64;
  HIPCHK(hipMemcpyAsync(s.src[0], f, sizeof(double) * ctx->N, hipMemcpyHostToDevice, ctx->st));
  Grid g = grid_of(ctx);
  dim3 gr = sgrid(ctx);
  gr.z = 1;
#define M_(TXX)                                                                                                   \
  if (id == 1) hipLaunchKernelGGL((k_deriv1<double, TXX, 1>), gr, dim3(kThreads), 0, ctx->st, g, s.sc, ctx->zc, \
                                  (const double*)s.src[0], s.src[1], iv - 1);                                     \
  else hipLaunchKernelGGL((k_deriv1<double, TXX, 2>), gr, dim3(kThreads), 0, ctx->st, g, s.sc, ctx->zc,         \
                          (const double*)s.src[0], s.src[1], iv - 1)
  TX_SWITCH(ctx->tx, M_)
#undef M_
  LAUNCHCHK();
  HIPCHK(hipMemcpyAsync(df, s.src[1], sizeof(double) * ctx->N, hipMemcpyDeviceToHost, ctx->st));
  HIPCHK(hipStreamSynchronize(ctx->st));
  return EFFPOT_OK;
}

// The 3-D transforms of effpot_fft_r2c/c2r in the precision T of state s
// (slot 0's source field and first spectrum as staging).
template <typename T> static int fft_r2c_t(effpot_ctx* ctx, State<T>& s, const T* in, T* out) {
  std::vector<int> m(ctx->B, 0);
  m[0] = 1;
  HIPCHK(hipMemcpyAsync(ctx->d_mask_tmp, m.data(), sizeof(int) * ctx->B, hipMemcpyHostToDevice, ctx->st));
  HIPCHK(hipMemcpyAsync(s.src[0], in, sizeof(T) * ctx->N, hipMemcpyHostToDevice, ctx->st));
  T* r[1] = {s.src[0]};
  cx<T>* c[1] = {s.spec[0]};
  if (plane_ok<T>(ctx)) {
    RC(fft_plane_fwd(ctx, s, r, c, 1, ctx->d_mask_tmp));
  } else {
    RC(fft_x_r2c(ctx, s, r, c, 1, ctx->d_mask_tmp));
    RC(fft_strided(ctx, s, 1, c, 1, -1, ctx->d_mask_tmp));
  }
  RC(fft_strided(ctx, s, 2, c, 1, -1, ctx->d_mask_tmp));
  HIPCHK(hipMemcpy2DAsync(out, sizeof(cx<T>) * ctx->nxh, s.spec[0], sizeof(cx<T>) * ctx->nxp,
                          sizeof(cx<T>) * ctx->nxh, (size_t)ctx->ny * ctx->nz, hipMemcpyDeviceToHost, ctx->st));
  HIPCHK(hipStreamSynchronize(ctx->st));
  return EFFPOT_OK;
}
template <typename T> static int fft_c2r_t(effpot_ctx* ctx, State<T>& s, const T* in, T* out) {
  std::vector<int> m(ctx->B, 0);
  m[0] = 1;
  HIPCHK(hipMemcpyAsync(ctx->d_mask_tmp, m.data(), sizeof(int) * ctx->B, hipMemcpyHostToDevice, ctx->st));
  HIPCHK(hipMemcpy2DAsync(s.spec[0], sizeof(cx<T>) * ctx->nxp, in, sizeof(cx<T>) * ctx->nxh,
                          sizeof(cx<T>) * ctx->nxh, (size_t)ctx->ny * ctx->nz, hipMemcpyHostToDevice, ctx->st));
  cx<T>* c[1] = {s.spec[0]};
  T* r[1] = {s.src[0]};
  RC(fft_strided(ctx, s, 2, c, 1, +1, ctx->d_mask_tmp));
  const T scale = (T)(1.0 / ((double)ctx->nx * ctx->ny * ctx->nz));
  if (iplane_ok<T>(ctx)) {
    RC(fft_plane_inv(ctx, s, c, r, 1, scale, ctx->d_mask_tmp));
  } else {
    RC(fft_strided(ctx, s, 1, c, 1, +1, ctx->d_mask_tmp));
    RC(fft_x_c2r(ctx, s, c, r, 1, scale, ctx->d_mask_tmp));
  }
  HIPCHK(hipMemcpyAsync(out, s.src[0], sizeof(T) * ctx->N, hipMemcpyDeviceToHost, ctx->st));
  HIPCHK(hipStreamSynchronize(ctx->st));
  return EFFPOT_OK;
}

int effpot_fft_r2c_fp32(effpot_ctx* ctx, const float* in, float* out) {
  RC(check_ctx(ctx));
  if (!in || !out) return ctx->fail(EFFPOT_ERR_ARG, "null pointer");
  if (!ctx->mixed) return ctx->fail(EFFPOT_ERR_ARG, "FP32 transforms need a mixed-precision context");
  hipSetDevice(ctx->dev);
  return fft_r2c_t(ctx, ctx->s32, in, out);
}

int effpot_fft_c2r_fp32(effpot_ctx* ctx, const float* in, float* out) {
  RC(check_ctx(ctx));
  if (!in || !out) return ctx->fail(EFFPOT_ERR_ARG, "null pointer");
  if (!ctx->mixed) return ctx->fail(EFFPOT_ERR_ARG, "FP32 transforms need a mixed-precision context");
  hipSetDevice(ctx->dev);
  return fft_c2r_t(ctx, ctx->s32, in, out);
}

int effpot_fft_r2c(effpot_ctx* ctx, const double* in, double* out) {
  RC(check_ctx(ctx));
  if (!in || !out) return ctx->fail(EFFPOT_ERR_ARG, "null pointer");
  hipSetDevice(ctx->dev);
  return fft_r2c_t(ctx, ctx->s64, in, out);
}

int effpot_fft_c2r(effpot_ctx* ctx, const double* in, double* out) {
  RC(check_ctx(ctx));
  if (!in || !out) return ctx->fail(EFFPOT_ERR_ARG, "null pointer");
  hipSetDevice(ctx->dev);
  return fft_c2r_t(ctx, ctx->s64, in, out);
}

// ---- continuous-batching engine ---------------------------------------------------
// Device masks: 0 start, 1 rk, 2 pc, 3 energy check, 4 polish conversion.
// Masks change only at phase transitions.  The tick does not wait for the
// device, so a queued copy may still read the pinned staging buffer: wait for
// the stream before rewriting it.
static int push_mask(effpot_ctx* ctx, int k, const std::vector<int>& m) {
  int* last = ctx->mask_dev.data() + (long)k * ctx->B;
  if (std::equal(m.begin(), m.end(), last)) return EFFPOT_OK;
  HIPCHK(hipStreamSynchronize(ctx->st));
  std::copy(m.begin(), m.end(), last);
  int* h = ctx->h_mask + (long)k * ctx->B;
  std::memcpy(h, m.data(), sizeof(int) * ctx->B);
  HIPCHK(hipMemcpyAsync(ctx->d_mask + (long)k * ctx->B, h, sizeof(int) * ctx->B, hipMemcpyHostToDevice, ctx->st));
  return EFFPOT_OK;
}

static void drop_graphs(effpot_ctx* ctx) {
  std::lock_guard<std::recursive_mutex> lk(g_api_mu);
  for (int k = 0; k < 2; k++) {
    if (ctx->pc_exec[k]) { hipGraphExecDestroy(ctx->pc_exec[k]); ctx->pc_exec[k] = nullptr; }
    if (ctx->pc_exec2[k]) { hipGraphExecDestroy(ctx->pc_exec2[k]); ctx->pc_exec2[k] = nullptr; }
    if (ctx->pck_exec[k]) { hipGraphExecDestroy(ctx->pck_exec[k]); ctx->pck_exec[k] = nullptr; }
    if (ctx->pck_exec2[k]) { hipGraphExecDestroy(ctx->pck_exec2[k]); ctx->pck_exec2[k] = nullptr; }
    ctx->pck_k[k] = ctx->pck_k2[k] = 0;
  }
}

// The PC step as two independent chains, one per slot group (masks dA, dB),
// on two streams inside one graph: the groups touch disjoint slot data, and
// each chain's launch tails and barriers overlap the other's work.  k > 1
// captures k consecutive steps of each chain in the same graph (one launch
// for k steps; the same kernels in the same order as k single replays).
template <typename T>
static int pc_step_graph2(effpot_ctx* ctx, State<T>& s, bool f32, const int* dA, const int* dB, int k = 1) {
  const int which = f32 ? 1 : 0;
  hipGraphExec_t& ex = k > 1 ? ctx->pck_exec2[which] : ctx->pc_exec2[which];
  if (ex && k > 1 && ctx->pck_k2[which] != k) {
    HIPCHK(hipStreamSynchronize(ctx->st));  // a replay of the old length may still run
    hipGraphExecDestroy(ex);
    ex = nullptr;
  }
  if (!ex) {
    std::lock_guard<std::recursive_mutex> lk(g_api_mu);
    hipGraph_t g;
    HIPCHK(hipStreamBeginCapture(ctx->st, hipStreamCaptureModeThreadLocal));
    hipError_t e1 = hipEventRecord(ctx->ev_fork, ctx->st);
    hipError_t e2 = hipStreamWaitEvent(ctx->st2, ctx->ev_fork, 0);
    int rc = (e1 == hipSuccess && e2 == hipSuccess) ? EFFPOT_OK : EFFPOT_ERR_DEVICE;
    ctx->corr_den = false;
    for (int part = 0; part < 2 && rc == EFFPOT_OK; part++) {
      if (part == 1) std::swap(ctx->st, ctx->st2);  // launches of group B go to the second stream
      const int* d = part == 0 ? dA : dB;
      for (int j = 0; j < k && rc == EFFPOT_OK; j++) {
        rc = step_pc_t(ctx, s, d, f32);
        if (rc == EFFPOT_OK) rc = renorm_t(ctx, s, d, f32);
      }
      if (part == 1) std::swap(ctx->st, ctx->st2);
    }
    ctx->corr_den = true;
    if (rc == EFFPOT_OK && (hipEventRecord(ctx->ev_join, ctx->st2) != hipSuccess ||
                            hipStreamWaitEvent(ctx->st, ctx->ev_join, 0) != hipSuccess))
      rc = EFFPOT_ERR_DEVICE;
    hipError_t e = hipStreamEndCapture(ctx->st, &g);
    if (rc != EFFPOT_OK) return ctx->fail(rc, "two-group graph capture failed");
    if (e != hipSuccess) return ctx->fail(EFFPOT_ERR_DEVICE, std::string("graph capture: ") + hipGetErrorString(e));
    e = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
    hipGraphDestroy(g);
    if (e != hipSuccess) return ctx->fail(EFFPOT_ERR_DEVICE, std::string("graph instantiate: ") + hipGetErrorString(e));
    if (k > 1) ctx->pck_k2[which] = k;
  }
  HIPCHK(hipGraphLaunch(ex, ctx->st));
  return EFFPOT_OK;
}

// One predictor-corrector step + renormalisation of the slots in mask dP (k
// consecutive steps when k > 1).  The launch sequence only references device
// memory (rings, masks, fields), so it is captured once per precision and
// replayed as a HIP graph (the time step of each slot is read from device
// memory, so changing it needs no re-capture).
template <typename T>
static int pc_step_graph(effpot_ctx* ctx, State<T>& s, bool f32, const int* dP, int k = 1) {
  const int which = f32 ? 1 : 0;
  if (!ctx->graphs) {
    for (int j = 0; j < k; j++) {
      ctx->corr_den = false;
      int rc = step_pc_t(ctx, s, dP, f32);
      ctx->corr_den = true;
      RC(rc);
      RC(renorm_t(ctx, s, dP, f32));
    }
    return EFFPOT_OK;
  }
  hipGraphExec_t& ex = k > 1 ? ctx->pck_exec[which] : ctx->pc_exec[which];
  if (ex && k > 1 && ctx->pck_k[which] != k) {
    HIPCHK(hipStreamSynchronize(ctx->st));  // a replay of the old length may still run
    hipGraphExecDestroy(ex);
    ex = nullptr;
  }
  if (!ex) {
    std::lock_guard<std::recursive_mutex> lk(g_api_mu);
    hipGraph_t g;
    HIPCHK(hipStreamBeginCapture(ctx->st, hipStreamCaptureModeThreadLocal));
    ctx->corr_den = false;
    int rc = EFFPOT_OK;
    for (int j = 0; j < k && rc == EFFPOT_OK; j++) {
      rc = step_pc_t(ctx, s, dP, f32);
      if (rc == EFFPOT_OK) rc = renorm_t(ctx, s, dP, f32);
    }
    ctx->corr_den = true;
    hipError_t e = hipStreamEndCapture(ctx->st, &g);
    if (rc != EFFPOT_OK) return rc;
    if (e != hipSuccess) return ctx->fail(EFFPOT_ERR_DEVICE, std::string("graph capture: ") + hipGetErrorString(e));
    e = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
    hipGraphDestroy(g);
    if (e != hipSuccess) return ctx->fail(EFFPOT_ERR_DEVICE, std::string("graph instantiate: ") + hipGetErrorString(e));
    if (k > 1) ctx->pck_k[which] = k;
  }
  HIPCHK(hipGraphLaunch(ex, ctx->st));
  return EFFPOT_OK;
}

// Advance the slots whose field precision matches f32 (FP32 mean fields of a
// MIXED context, else FP64) by one step of their own phase.  The wavefunction,
// its history and the integrator always live in the FP64 state.
// A mixed slot hands over to FP64 fields this many energy checks before its
// |dE/E| trend reaches tol: the FP32-phase rate of decrease is biased by a few
// checks' worth near the end of a relaxation.
constexpr double kSwitchChecks = 10.0;

static int tick_phase(effpot_ctx* ctx, bool f32, const effpot_solve_opts* opt, effpot_slot* sl) {
  State<double>& s = ctx->s64;
  const int B = ctx->B;
  std::vector<int> ms(B, 0), mr(B, 0), mp(B, 0);
  bool any_s = false, any_r = false, any_p = false;
  for (int b = 0; b < B; b++) {
    if (!ctx->h_active[b] || ((sl[b].fp32 != 0) != f32)) continue;
    if (sl[b].phase == EFFPOT_SLOT_START) { ms[b] = 1; any_s = true; }
    else if (sl[b].phase == EFFPOT_SLOT_RK) { mr[b] = 1; any_r = true; }
    else if (sl[b].phase == EFFPOT_SLOT_PC) { mp[b] = 1; any_p = true; }
  }
  const int* dS = ctx->d_mask;
  const int* dR = ctx->d_mask + B;
  const int* dP = ctx->d_mask + 2 * B;
  const int* dE = ctx->d_mask + 3 * B;
  std::vector<effpot_energies> e(B);
  // Every energy is FP64 (FP64 state, FP64 poten at the check), but the rate
  // at which it falls depends on the fields that advanced the state, and near
  // the end of a relaxation |dE/E| per step decreases so slowly that a 0.2 %
  // bias moves its crossing of tol by several checks.  So convergence is
  // decided in the FP64 phase only: an FP32 slot switches to FP64 fields (the
  // polish) when its |dE/E| per step is below tol_fp32, or when it is below
  // 2 tol and its trend (smoothed change per check) would reach tol within
  // kSwitchChecks checks, or at max_steps; it is DONE at the first FP64 check
  // below tol, or after max_steps + polish_steps.  (The early, fast decay is
  // not extrapolated: it would hand almost the whole relaxation to FP64.)
  const int cap_f32 = opt->max_steps;
  const int cap = ctx->mixed ? opt->max_steps + std::max(opt->polish_steps, 0) : opt->max_steps;
  // START: fields of the new states; they take their first RK step below.
  if (any_s) {
    RC(push_mask(ctx, 0, ms));
    if (f32) {
      // a new solve starts from the defect correction of its starting state,
      // so the RK starter already runs on FP32 fields that match FP64
      hipLaunchKernelGGL((k_zero_slots<float>), pw_grid(ctx->N, B), dim3(256), 0, ctx->st, ctx->s32.corr, dS, ctx->N);
      LAUNCHCHK();
      RC(den_from_cur(ctx, dS, false));
      RC(poten_t(ctx, s, dS));
    }
    RC(den_from_cur(ctx, dS, f32));
    RC(f32 ? poten_t(ctx, ctx->s32, dS) : poten_t(ctx, s, dS));
    if (f32) {
      hipLaunchKernelGGL(k_defect, pw_grid(ctx->N, B), dim3(256), 0, ctx->st, (const double*)s.pot4,
                         (const double*)s.uimp, ctx->s32.pot4, (const float*)ctx->s32.uimp, ctx->s32.corr, dS, ctx->N);
      LAUNCHCHK();
    }
    for (int b = 0; b < B; b++)
      if (ms[b]) {  // a new solve: counters start from zero
        sl[b].nrk = 0; sl[b].since = 0; sl[b].steps = 0; sl[b].converged = 0;
        sl[b].eprev = 0.0; sl[b].rel_de = 0.0;
        ctx->h_relprev[b] = 0.0;
        ctx->h_reltrend[b] = 0.0;
        if (opt->n_rk > 0) { sl[b].phase = EFFPOT_SLOT_RK; mr[b] = 1; any_r = true; }
        else { sl[b].phase = EFFPOT_SLOT_PC; }
      }
  }
  std::vector<int> me(B, 0);
  bool any_e = false;
  if (any_r) {
    RC(push_mask(ctx, 1, mr));
    RC(step_rk_t(ctx, s, dR, f32));
    RC(renorm_t(ctx, s, dR, f32));
    for (int b = 0; b < B; b++)
      if (mr[b]) {
        sl[b].nrk++; sl[b].steps++;
        if (sl[b].nrk >= opt->n_rk) { me[b] = 1; any_e = true; }  // baseline energy before PC
      }
  }
  if (any_p) {
    // k > 1 PC steps in one graph launch when no slot of this phase reaches an
    // energy check (hence a phase change) or its step cap within them
    int k = 1;
    if (ctx->graphs && opt->graph_steps > 1) {
      k = opt->graph_steps;
      for (int b = 0; b < B; b++)
        if (mp[b] && (sl[b].since + k >= opt->pener || sl[b].steps + k >= (f32 ? cap_f32 : cap))) k = 1;
    }
    RC(push_mask(ctx, 2, mp));
    if (ctx->graphs && ctx->split && B >= 2) {
      std::vector<int> ma(B, 0), mb(B, 0);
      for (int b = 0; b < B; b++) (b < B / 2 ? ma : mb)[b] = mp[b];
      RC(push_mask(ctx, 5, ma));
      RC(push_mask(ctx, 6, mb));
      RC(pc_step_graph2(ctx, s, f32, ctx->d_mask + 5 * B, ctx->d_mask + 6 * B, k));
    } else {
      RC(pc_step_graph(ctx, s, f32, dP, k));
    }
    for (int b = 0; b < B; b++)
      if (mp[b]) {
        sl[b].steps += k; sl[b].since += k;
        if (sl[b].since >= opt->pener || sl[b].steps >= cap) { me[b] = 1; any_e = true; }
      }
  }
  if (any_e) {
    // energy checks always in FP64: density of the current psi, FP64 poten, energy
    RC(push_mask(ctx, 3, me));
    if (f32) RC(den_from_cur(ctx, dE, false));
    if (opt->check_finite) {
      HIPCHK(hipMemsetAsync(ctx->d_bad, 0, sizeof(int) * B, ctx->st));
      hipLaunchKernelGGL((k_check_den<double>), pw_grid(ctx->N, B), dim3(256), 0, ctx->st, (const double*)s.den,
                         ctx->N, ctx->N, dE, ctx->d_bad);
      LAUNCHCHK();
    }
    RC(poten_t(ctx, s, dE));
    RC(energy_t(ctx, s, dE, e.data()));
    if (f32) {
      // Defect correction of the FP32 mean field: evaluate it at the same
      // state, add the difference to the FP64 field (pot4 + uimp) to the
      // correction carried by every later FP32 poten, and continue from the
      // FP64 field of this state (as the FP64 path does after its energy check).
      RC(den_from_cur(ctx, dE, true));
      RC(poten_t(ctx, ctx->s32, dE));
      hipLaunchKernelGGL(k_defect, pw_grid(ctx->N, B), dim3(256), 0, ctx->st, (const double*)s.pot4,
                         (const double*)s.uimp, ctx->s32.pot4, (const float*)ctx->s32.uimp, ctx->s32.corr, dE, ctx->N);
      LAUNCHCHK();
    }
    std::vector<int> bad(B, 0);
    if (opt->check_finite) {
      HIPCHK(hipMemcpyAsync(bad.data(), ctx->d_bad, sizeof(int) * B, hipMemcpyDeviceToHost, ctx->st));
      HIPCHK(hipStreamSynchronize(ctx->st));
    }
    for (int b = 0; b < B; b++) {
      if (!me[b]) continue;
      effpot_slot& q = sl[b];
      if (bad[b] || !std::isfinite(e[b].etot)) { q.phase = EFFPOT_SLOT_FAILED; continue; }
      if (q.phase == EFFPOT_SLOT_RK) {  // end of the RK starter
        q.phase = EFFPOT_SLOT_PC; q.eprev = e[b].etot; q.e = e[b]; q.since = 0;
        continue;
      }
      double rel = std::fabs(e[b].etot - q.eprev) / std::fabs(e[b].etot) / (double)std::max(q.since, 1);
      q.rel_de = rel; q.eprev = e[b].etot; q.e = e[b]; q.since = 0;
      const double prev = ctx->h_relprev[b];
      ctx->h_relprev[b] = rel;
      if (prev > 0.0) {
        const double d = rel - prev;
        ctx->h_reltrend[b] = ctx->h_reltrend[b] == 0.0 ? d : 0.7 * ctx->h_reltrend[b] + 0.3 * d;
      }
      const double trend = ctx->h_reltrend[b];
      if (!f32) {
        if (rel < opt->tol || q.steps >= cap) { q.phase = EFFPOT_SLOT_DONE; q.converged = rel < opt->tol; }
      } else if (rel < opt->tol_fp32 || (rel < 2.0 * opt->tol && trend < 0.0 && rel + kSwitchChecks * trend < opt->tol) ||
                 q.steps >= cap_f32) {
        q.fp32 = 0;  // FP64 polish: same state and history, FP64 fields from here on
      }
    }
  }
  return EFFPOT_OK;
}

int effpot_tick(effpot_ctx* ctx, const effpot_solve_opts* opt, effpot_slot* slots) {
  RC(check_ctx(ctx));
  if (!opt || !slots) return ctx->fail(EFFPOT_ERR_ARG, "null pointer");
  if (ctx->rt || ctx->any_motion || ctx->any_dyn)
    return ctx->fail(EFFPOT_ERR_STATE,
                     "relaxation with real time, a moving impurity or evolving invar: step with effpot_step_rk/pc");
  if (!(opt->dt > 0)) return ctx->fail(EFFPOT_ERR_DT, "non-positive time step");
  if (opt->pener < 1 || opt->max_steps < 1 || opt->n_rk < 0) return ctx->fail(EFFPOT_ERR_ARG, "bad solve options");
  hipSetDevice(ctx->dev);
  for (int b = 0; b < ctx->B; b++) {
    int ph = slots[b].phase;
    if (ctx->h_active[b] && ph >= EFFPOT_SLOT_START && ph <= EFFPOT_SLOT_PC && !ctx->have_imp[b])
      return ctx->fail(EFFPOT_ERR_STATE, "impurity not set for slot " + std::to_string(b));
    if (!ctx->mixed) slots[b].fp32 = 0;
  }
  RC(push_dt(ctx, opt->dt));
  if (ctx->mixed) RC(tick_phase(ctx, true, opt, slots));
  RC(tick_phase(ctx, false, opt, slots));
  // No wait here: the slot phases are decided on the host and only energy
  // checks read the device (they wait themselves), so the host queues the
  // next steps while the device runs these.
  HIPCHK(hipGetLastError());
  return EFFPOT_OK;
}

// The corrector timed inside whole predictor-corrector steps: the caches hold
// what a relaxation leaves in them (poten runs in between), not the
// corrector's own previous launch.
// whole = false: the corrector inside `reps` whole PC steps; true: the whole
// steps (predictor, poten, corrector, norm reduction, renormalisation).
template <typename T> static int time_corrector_in_steps(effpot_ctx* ctx, State<T>& s, int reps, double* ms,
                                                         bool f32, bool whole = false) {
  const int* act = ctx->d_active;
  std::vector<hipEvent_t> ev(2 * reps);
  for (auto& e : ev) HIPCHK(hipEventCreate(&e));
  int rc = EFFPOT_OK;
  for (int r = 0; r < reps && rc == EFFPOT_OK; r++) {
    if (whole) HIPCHK(hipEventRecord(ev[2 * r], ctx->st));
    rc = launch_step<T, 0>(ctx, s, act, f32);
    const double* ds = nullptr;
    if (rc == EFFPOT_OK) rc = step_den_scale(ctx, act, f32, &ds);
    if (rc == EFFPOT_OK) rc = f32 ? poten_t(ctx, ctx->s32, act, -1, ds) : poten_t(ctx, s, act, -1, ds);
    if (rc != EFFPOT_OK) break;
    if (!whole) HIPCHK(hipEventRecord(ev[2 * r], ctx->st));
    ctx->corr_den = false;
    rc = launch_step<T, 1>(ctx, s, act, f32);
    ctx->corr_den = true;
    if (!whole) HIPCHK(hipEventRecord(ev[2 * r + 1], ctx->st));
    if (rc == EFFPOT_OK) rc = reduce<2>(ctx, 1, act, s.ring, RING_AFTER_PRED, RING_AFTER_CORR);
    if (rc == EFFPOT_OK) rc = renorm_t(ctx, s, act, f32);
    if (whole) HIPCHK(hipEventRecord(ev[2 * r + 1], ctx->st));
  }
  HIPCHK(hipStreamSynchronize(ctx->st));
  double sum = 0.0;
  for (int r = 0; r < reps && rc == EFFPOT_OK; r++) {
    float t = 0.f;
    HIPCHK(hipEventElapsedTime(&t, ev[2 * r], ev[2 * r + 1]));
    sum += t;
  }
  for (auto& e : ev) hipEventDestroy(e);
  *ms = sum / reps;
  return rc;
}

template <typename T> static int time_kernel_t(effpot_ctx* ctx, State<T>& s, int which, int reps, double* ms,
                                               bool f32 = false) {
  if (which == 3 || which == 4) return time_corrector_in_steps(ctx, s, reps, ms, f32, which == 4);
  hipEvent_t e0, e1;
  HIPCHK(hipEventCreate(&e0));
  HIPCHK(hipEventCreate(&e1));
  auto once = [&]() -> int {
    switch (which) {
      case 0: return launch_step<T, 0>(ctx, s, ctx->d_active, f32);
      case 1: {  // as in a relaxation: the renormalisation that follows stores the density
        ctx->corr_den = false;
        int rc = launch_step<T, 1>(ctx, s, ctx->d_active, f32);
        ctx->corr_den = true;
        return rc;
      }
      case 2: return f32 ? poten_t(ctx, ctx->s32, ctx->d_active) : poten_t(ctx, s, ctx->d_active);
      default:
        if (which >= 100 && which < 118)
          return f32 ? poten_t(ctx, ctx->s32, ctx->d_active, which - 100) : poten_t(ctx, s, ctx->d_active, which - 100);
        return ctx->fail(EFFPOT_ERR_ARG, "unknown kernel id");
    }
  };
  RC(once());
  HIPCHK(hipEventRecord(e0, ctx->st));
  for (int r = 0; r < reps; r++) RC(once());
  HIPCHK(hipEventRecord(e1, ctx->st));
  HIPCHK(hipEventSynchronize(e1));
  float t = 0.f;
  HIPCHK(hipEventElapsedTime(&t, e0, e1));
  *ms = (double)t / reps;
  hipEventDestroy(e0);
  hipEventDestroy(e1);
  return EFFPOT_OK;
}

int effpot_autotune(effpot_ctx* ctx, int reps, int* tx_out, int* zc_out) {
  RC(check_ctx(ctx));
  hipSetDevice(ctx->dev);
  RC(push_dt(ctx, 1e-4));
  int best_tx = ctx->tx, best_zc = ctx->zc;
  double best = 1e30;
  for (int tx : {64, 32, 16, 8}) {
    if (ctx->nx % tx) continue;
    for (int zc : {4, 8, 16, 32}) {
      if (zc > ctx->nz) continue;
      ctx->tx = tx; ctx->zc = zc;
      double t0 = 0, t1 = 0, t2 = 0;
      RC(time_kernel_t(ctx, ctx->s64, 0, reps, &t0, ctx->mixed));
      RC(time_kernel_t(ctx, ctx->s64, 1, reps, &t1, ctx->mixed));
      RC(time_kernel_t(ctx, ctx->s64, 2, reps, &t2, ctx->mixed));
      if (ctx->tune_log) std::fprintf(stderr, "autotune tx=%d zc=%d pred=%.4f corr=%.4f poten=%.4f ms\n", tx, zc, t0, t1, t2);
      if (t0 + t1 + t2 < best) { best = t0 + t1 + t2; best_tx = tx; best_zc = zc; }
    }
  }
  ctx->tx = best_tx; ctx->zc = best_zc;
  // the forward x-y transform schedule: separate passes (0), the plane kernel
  // with sub-group scratch (1) or in place (3), timed on whole poten launches
  // of the live precision (EFFPOT_PLANE set in the environment fixes it)
  if (!ctx->plane_fixed) {
    const bool f32 = ctx->mixed;
    int best_plane = ctx->plane;
    double best_p = 1e30;
    for (int mode : {0, 1, 3}) {
      ctx->plane = mode;
      const int eff = f32 ? plane_mode<float>(ctx) : plane_mode<double>(ctx);
      if (eff != mode) continue;  // not applicable at this size and precision
      double t = 0;
      RC(time_kernel_t(ctx, ctx->s64, 2, reps, &t, f32));
      if (ctx->tune_log) std::fprintf(stderr, "autotune plane=%d poten=%.4f ms\n", mode, t);
      if (t < best_p) { best_p = t; best_plane = mode; }
    }
    ctx->plane = best_plane;
  }
  // the inverse y + x transform: separate passes (0) or the in-place plane kernel (1)
  if (!ctx->iplane_fixed) {
    const bool f32 = ctx->mixed;
    int best_i = ctx->iplane;
    double best_p = 1e30;
    for (int on : {0, 1}) {
      ctx->iplane = on;
      if (on && !(f32 ? iplane_ok<float>(ctx) : iplane_ok<double>(ctx))) continue;
      double t = 0;
      RC(time_kernel_t(ctx, ctx->s64, 2, reps, &t, f32));
      if (ctx->tune_log) std::fprintf(stderr, "autotune iplane=%d poten=%.4f ms\n", on, t);
      if (t < best_p) { best_p = t; best_i = on; }
    }
    ctx->iplane = best_i;
  }
  // the half-spectrum layout inside poten: packed or padded (where the packed
  // one applies), timed on whole poten launches of the live precision
  if (!ctx->packed_fixed) {
    const bool f32 = ctx->mixed;
    int best_k = ctx->packed;
    double best_p = 1e30;
    for (int k : {0, 1}) {
      ctx->packed = k;
      if (k == 1 && !(f32 ? packed_ok<float>(ctx) : packed_ok<double>(ctx))) continue;
      double t = 0;
      RC(time_kernel_t(ctx, ctx->s64, 2, reps, &t, f32));
      if (ctx->tune_log) std::fprintf(stderr, "autotune packed=%d poten=%.4f ms\n", k, t);
      if (t < best_p) { best_p = t; best_k = k; }
    }
    ctx->packed = best_k;
  }
  // the z-convolution group: 16 or 32 lines (256 or 512 threads), timed on
  // whole poten launches of the live precision
  if (!ctx->zconv_fixed) {
    const bool f32 = ctx->mixed;
    int best_t = ctx->zconv_th;
    double best_p = 1e30;
    for (int th : {256, 512}) {
      ctx->zconv_th = th;
      double t = 0;
      RC(time_kernel_t(ctx, ctx->s64, 2, reps, &t, f32));
      if (ctx->tune_log) std::fprintf(stderr, "autotune zconv_th=%d poten=%.4f ms\n", th, t);
      if (t < best_p) { best_p = t; best_t = th; }
    }
    ctx->zconv_th = best_t;
  }
  drop_graphs(ctx);  // captured graphs hold the old launch shapes
  if (tx_out) *tx_out = best_tx;
  if (zc_out) *zc_out = best_zc;
  return EFFPOT_OK;
}

int effpot_set_tiles(effpot_ctx* ctx, int tx, int zc) {
  RC(check_ctx(ctx));
  if ((tx != 64 && tx != 32 && tx != 16 && tx != 8) || ctx->nx % tx || zc < 1 || zc > ctx->nz)
    return ctx->fail(EFFPOT_ERR_ARG, "bad tile");
  int tx0 = ctx->tx, zc0 = ctx->zc;
  ctx->tx = tx; ctx->zc = zc;
  if (nblk(ctx) > ctx->nblk_max) { ctx->tx = tx0; ctx->zc = zc0; return ctx->fail(EFFPOT_ERR_ARG, "tile exceeds partial buffer"); }
  HIPCHK(hipStreamSynchronize(ctx->st));  // a replay may still be running
  drop_graphs(ctx);
  return EFFPOT_OK;
}

int effpot_kernel_time(effpot_ctx* ctx, int which, int reps, double* ms_per_launch) {
  RC(check_ctx(ctx));
  if (!ms_per_launch || reps < 1) return ctx->fail(EFFPOT_ERR_ARG, "bad arguments");
  hipSetDevice(ctx->dev);
  RC(push_dt(ctx, 1e-4));
  return time_kernel_t(ctx, ctx->s64, which, reps, ms_per_launch, ctx->mixed);
}

int effpot_plane_fused(effpot_ctx* ctx, int fp32_fields) {
  if (check_ctx(ctx) != EFFPOT_OK) return 0;
  return fp32_fields ? plane_mode<float>(ctx) : plane_mode<double>(ctx);
}

int effpot_packed_layout(effpot_ctx* ctx, int fp32_fields) {
  if (check_ctx(ctx) != EFFPOT_OK) return 0;
  return (fp32_fields ? packed_ok<float>(ctx) : packed_ok<double>(ctx)) ? 1 : 0;
}

int effpot_zconv_threads(effpot_ctx* ctx) {
  if (check_ctx(ctx) != EFFPOT_OK) return 0;
  return ctx->zconv_th;
}

int effpot_plane_inverse_fused(effpot_ctx* ctx, int fp32_fields) {
  if (check_ctx(ctx) != EFFPOT_OK) return 0;
  return (fp32_fields ? iplane_ok<float>(ctx) : iplane_ok<double>(ctx)) ? 1 : 0;
}

int effpot_solve(effpot_ctx* ctx, const effpot_solve_opts* opt, effpot_result* out) {
  RC(check_ctx(ctx));
  if (!opt || !out) return ctx->fail(EFFPOT_ERR_ARG, "null pointer");
  const int B = ctx->B;
  std::vector<effpot_slot> sl(B);
  std::memset(sl.data(), 0, sizeof(effpot_slot) * B);
  for (int b = 0; b < B; b++) {
    if (!ctx->h_active[b]) continue;
    sl[b].phase = EFFPOT_SLOT_START;
    sl[b].fp32 = ctx->mixed ? 1 : 0;
  }
  for (;;) {
    bool busy = false;
    for (int b = 0; b < B; b++)
      if (ctx->h_active[b] && sl[b].phase >= EFFPOT_SLOT_START && sl[b].phase <= EFFPOT_SLOT_PC) busy = true;
    if (!busy) break;
    RC(effpot_tick(ctx, opt, sl.data()));
  }
  int rc = EFFPOT_OK;
  for (int b = 0; b < B; b++) {
    std::memset(&out[b], 0, sizeof(effpot_result));
    out[b].e = sl[b].e;
    out[b].steps = sl[b].steps;
    out[b].converged = sl[b].converged;
    out[b].rel_de = sl[b].rel_de;
    if (sl[b].phase == EFFPOT_SLOT_FAILED) rc = ctx->fail(EFFPOT_ERR_NONFINITE, "non-finite density or energy (slot " + std::to_string(b) + ")");
  }
  HIPCHK(hipStreamSynchronize(ctx->st));
  return rc;
}
