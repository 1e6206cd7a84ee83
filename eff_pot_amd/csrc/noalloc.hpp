// Renormalisation to n4 atoms on caller-owned device memory with no device
// allocation: the variant of effpot_renormalize for callers that cannot call
// hipMalloc (code under HIP graph capture, serving loops, harnesses that
// forbid allocation).  The norm's block partial sums and the per-slot scale
// live in one file-scope device array (`scratch`), not in context memory.
//
// The array's name is a host-side stand-in: host code never passes it to a
// kernel.  The first call on a device, with that device current, looks up its
// device address with hipGetSymbolAddress and keeps it per device; every
// launch takes that pointer.  Every partial sum and scale a call reads is
// written earlier in the same call, so no part of the array has to start at
// zero, and nothing is cleared with a memset (which a captured graph must not
// hold).  All calls on one device share the array: calls that overlap on one
// device (different streams, or host threads) race.
#pragma once
#include <hip/hip_runtime.h>

#include <cmath>
#include <mutex>
#include <string>

#include "common.hpp"
#include "stencil.hpp"

namespace ep {

constexpr int kNaMaxSlots = 64;                         // slots per call
constexpr int kNaBlocks = 512;                          // norm partials per slot
constexpr long kNaMaxPoints = (long)kMaxN * kMaxN * kMaxN;  // points per slot (256^3)
constexpr int kNaMaxDevices = 64;
// partials [slot][kNaBlocks], then the scale of each slot: 264 KiB per device
constexpr long kNaScratch = (long)kNaMaxSlots * kNaBlocks + kNaMaxSlots;

static __device__ __align__(256) double scratch[kNaScratch];

struct NaArgs {
  void* psi;          // [nslots][npoints] complex (interleaved) or real
  double* den;        // [nslots][npoints] or null
  double* natoms;     // [nslots] sum |psi|^2 dxyz before the scaling, or null
  double* part;       // the scratch array's device address
  long npoints;
  int nslots, nblk;
  double dxyz;
  double n4[kNaMaxSlots];
};

template <typename V>
__global__ __launch_bounds__(kThreads) void k_na_partial(NaArgs a) {
  __shared__ double red[4];
  const int s = blockIdx.y;
  const V* psi = reinterpret_cast<const V*>(a.psi) + (long)s * a.npoints;
  double v[1] = {0.0};
  for (long i = blockIdx.x * (long)kThreads + threadIdx.x; i < a.npoints; i += (long)gridDim.x * kThreads)
    v[0] += (double)norm2(psi[i]);
  block_sum<1>(v, red);
  if (threadIdx.x == 0) a.part[(long)s * kNaBlocks + blockIdx.x] = v[0];
}

// one block per slot: sum the slot's nblk partials in a fixed order
__global__ __launch_bounds__(kThreads) void k_na_finish(NaArgs a) {
  __shared__ double red[4];
  const int s = blockIdx.x;
  double v[1] = {0.0};
  for (int b = threadIdx.x; b < a.nblk; b += kThreads) v[0] += a.part[(long)s * kNaBlocks + b];
  block_sum<1>(v, red);
  if (threadIdx.x == 0) {
    const double nat = v[0] * a.dxyz;
    a.part[(long)kNaMaxSlots * kNaBlocks + s] = sqrt(a.n4[s] / nat);
    if (a.natoms) a.natoms[s] = nat;
  }
}

template <typename V>
__global__ __launch_bounds__(kThreads) void k_na_scale(NaArgs a) {
  const int s = blockIdx.y;
  const double f = a.part[(long)kNaMaxSlots * kNaBlocks + s];
  V* psi = reinterpret_cast<V*>(a.psi) + (long)s * a.npoints;
  double* d = a.den ? a.den + (long)s * a.npoints : nullptr;
  for (long i = blockIdx.x * (long)kThreads + threadIdx.x; i < a.npoints; i += (long)gridDim.x * kThreads) {
    const V p = f * psi[i];
    psi[i] = p;
    if (d) d[i] = norm2(p);
  }
}

namespace {

std::mutex g_na_mu;
double* g_na_ptr[kNaMaxDevices] = {};

// Device address of `scratch` on the current device (looked up once per device).
hipError_t na_scratch(int dev, double** out) {
  std::lock_guard<std::mutex> lk(g_na_mu);
  if (!g_na_ptr[dev]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    g_na_ptr[dev] = static_cast<double*>(p);
  }
  *out = g_na_ptr[dev];
  return hipSuccess;
}

}  // namespace

// Host side of effpot_renormalize_noalloc; `err` receives the message of a refusal.
inline int renormalize_noalloc(double* psi, double* den, int64_t npoints, int nslots, int psi_complex, double dxyz,
                               const double* n4, double* natoms_out, void* stream, std::string& err) {
  if (!psi || !n4) { err = "renormalize_noalloc: null psi or n4"; return EFFPOT_ERR_ARG; }
  if (nslots < 1 || nslots > kNaMaxSlots) {
    err = "renormalize_noalloc: nslots must be 1.." + std::to_string(kNaMaxSlots) + " (the scratch array's size)";
    return EFFPOT_ERR_ARG;
  }
  if (npoints < 1 || npoints > kNaMaxPoints) {
    err = "renormalize_noalloc: npoints must be 1.." + std::to_string(kNaMaxPoints) + " (256^3) per slot";
    return EFFPOT_ERR_ARG;
  }
  if (!(dxyz > 0.0) || !std::isfinite(dxyz)) { err = "renormalize_noalloc: dxyz must be finite and > 0"; return EFFPOT_ERR_ARG; }
  NaArgs a{};
  for (int s = 0; s < nslots; s++) {
    if (!(n4[s] > 0.0) || !std::isfinite(n4[s])) { err = "renormalize_noalloc: n4 must be finite and > 0"; return EFFPOT_ERR_ARG; }
    a.n4[s] = n4[s];
  }
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) { err = std::string("hipGetDevice: ") + hipGetErrorString(e); return EFFPOT_ERR_DEVICE; }
  if (dev < 0 || dev >= kNaMaxDevices) { err = "renormalize_noalloc: device index beyond 63"; return EFFPOT_ERR_DEVICE; }
  e = na_scratch(dev, &a.part);
  if (e != hipSuccess) { err = std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e); return EFFPOT_ERR_DEVICE; }
  a.psi = psi; a.den = den; a.natoms = natoms_out; a.npoints = npoints; a.nslots = nslots; a.dxyz = dxyz;
  a.nblk = (int)std::min<long>(kNaBlocks, (npoints + kThreads - 1) / kThreads);
  hipStream_t st = static_cast<hipStream_t>(stream);
  const dim3 gp((unsigned)a.nblk, (unsigned)nslots), gs((unsigned)std::min<long>(2048, (npoints + kThreads - 1) / kThreads), (unsigned)nslots);
  if (psi_complex) hipLaunchKernelGGL((k_na_partial<cx<double>>), gp, dim3(kThreads), 0, st, a);
  else hipLaunchKernelGGL((k_na_partial<double>), gp, dim3(kThreads), 0, st, a);
  hipLaunchKernelGGL(k_na_finish, dim3((unsigned)nslots), dim3(kThreads), 0, st, a);
  if (psi_complex) hipLaunchKernelGGL((k_na_scale<cx<double>>), gs, dim3(kThreads), 0, st, a);
  else hipLaunchKernelGGL((k_na_scale<double>), gs, dim3(kThreads), 0, st, a);
  e = hipGetLastError();
  if (e != hipSuccess) { err = std::string("launch: ") + hipGetErrorString(e); return EFFPOT_ERR_DEVICE; }
  return EFFPOT_OK;
}

}  // namespace ep
