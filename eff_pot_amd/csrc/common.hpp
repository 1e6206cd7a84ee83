// Shared device primitives for the gfx950 Orsay-Trento solver.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define EP_DEV __device__ __forceinline__
#define EP_HD __host__ __device__ __forceinline__

namespace ep {

// Interleaved complex value (matches Fortran complex(8)/complex(4) layout).
template <typename T> struct alignas(2 * sizeof(T)) cx { T x, y; };

template <typename T> EP_HD cx<T> mk(T a, T b) { return cx<T>{a, b}; }
template <typename T> EP_HD cx<T> operator+(cx<T> a, cx<T> b) { return {a.x + b.x, a.y + b.y}; }
template <typename T> EP_HD cx<T> operator-(cx<T> a, cx<T> b) { return {a.x - b.x, a.y - b.y}; }
template <typename T> EP_HD cx<T> operator*(T s, cx<T> a) { return {s * a.x, s * a.y}; }
template <typename T> EP_HD cx<T> cmul(cx<T> a, cx<T> b) {
  return {a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
}
template <typename T> EP_HD cx<T> conjg(cx<T> a) { return {a.x, -a.y}; }
// multiply by sign*i  (sign = -1 forward, +1 backward)
template <typename T> EP_HD cx<T> mul_si(cx<T> a, int sign) {
  return sign < 0 ? cx<T>{a.y, -a.x} : cx<T>{-a.y, a.x};
}

// Element i of an array whose elements all lie within 4 GiB of its base (one
// field of one slot: at most 256^3 complex FP64 values, 268 MB), addressed by
// a 32-bit byte offset: the loads and stores then take the uniform-base +
// 32-bit-offset form instead of a 64-bit address per access held in two VGPRs.
template <typename V> EP_DEV V ld32(const V* base, unsigned i) {
  return *reinterpret_cast<const V*>(reinterpret_cast<const char*>(base) + i * (unsigned)sizeof(V));
}
template <typename V> EP_DEV void st32(V* base, unsigned i, V v) {
  *reinterpret_cast<V*>(reinterpret_cast<char*>(base) + i * (unsigned)sizeof(V)) = v;
}

// Value helpers so stencil code is generic over real and complex fields.
template <typename T> EP_HD T vzero(T*) { return T(0); }
template <typename T> EP_HD cx<T> vzero(cx<T>*) { return cx<T>{T(0), T(0)}; }
template <typename T> EP_HD T vfma(T c, T v, T acc) { return acc + c * v; }
template <typename T> EP_HD cx<T> vfma(T c, cx<T> v, cx<T> acc) { return {acc.x + c * v.x, acc.y + c * v.y}; }

// Largest tile widths etc.
constexpr int kThreads = 256;
constexpr int kMaxN = 256;

// ---- block / wave reductions (deterministic order) ----------------------
EP_DEV double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// Sum K values over a 256-thread block; result valid in thread 0.
template <int K> EP_DEV void block_sum(double (&v)[K], double* lds /* >= 4*K */) {
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
  for (int k = 0; k < K; k++) v[k] = wave_sum(v[k]);
  if (lane == 0) {
#pragma unroll
    for (int k = 0; k < K; k++) lds[wid * K + k] = v[k];
  }
  __syncthreads();
  if (threadIdx.x == 0) {
#pragma unroll
    for (int k = 0; k < K; k++) v[k] = ((lds[k] + lds[K + k]) + lds[2 * K + k]) + lds[3 * K + k];
  }
  __syncthreads();
}

}  // namespace ep
