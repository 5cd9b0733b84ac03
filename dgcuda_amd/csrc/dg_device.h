// dg_device.h -- device helpers shared by the forward (dg_kernels.hip) and
// adjoint (dg_adjoint.hip) kernels: template dispatch over (physics, order),
// the XCD-aware block remap, reductions and global-address-space accesses.
#pragma once

#include <type_traits>
#include <utility>

#include "dg_internal.h"

namespace dgk {

// ------------------------------------------------------------ dispatch

// DG_ONLY_PHYS / DG_ONLY_P restrict the instantiations (quick experiments)
template <int PH, int PP>
constexpr bool enabled() {
#if defined(DG_ONLY_PHYS) && defined(DG_ONLY_P)
  return PH == DG_ONLY_PHYS && PP == DG_ONLY_P;
#else
  return true;
#endif
}

template <typename F>
static void dispatch(int phys, int p, F &&f) {
  auto call = [&](auto PH, auto PP) {
    if constexpr (enabled<decltype(PH)::value, decltype(PP)::value>()) f(PH, PP);
  };
  auto withp = [&](auto PH) {
    switch (p) {
      case 0: call(PH, std::integral_constant<int, 0>{}); break;
      case 1: call(PH, std::integral_constant<int, 1>{}); break;
      case 2: call(PH, std::integral_constant<int, 2>{}); break;
      case 3: call(PH, std::integral_constant<int, 3>{}); break;
      case 4: call(PH, std::integral_constant<int, 4>{}); break;
      case 5: call(PH, std::integral_constant<int, 5>{}); break;
    }
  };
  switch (phys) {
    case 0: withp(std::integral_constant<int, 0>{}); break;
    case 1: withp(std::integral_constant<int, 1>{}); break;
    case 2: withp(std::integral_constant<int, 2>{}); break;
    case 3: withp(std::integral_constant<int, 3>{}); break;
    case 4: withp(std::integral_constant<int, 4>{}); break;
  }
}

static inline int nblocks(long long n, int blk) { return (int)((n + blk - 1) / blk); }

// Workgroups are dispatched round-robin over the 8 XCDs; give XCD x the
// contiguous logical range [x*q + min(x, r), ...) of the nblk blocks.
__device__ __forceinline__ int remap_block(int b, int nblk) {
  int q = nblk >> 3, r = nblk & 7, x = b & 7;
  return x * q + (x < r ? x : r) + (b >> 3);
}

// Order-preserving 64-bit key of a double: unsigned comparison of keys is
// numeric comparison, with +NaN above +inf, so a max over keys keeps a NaN
// (the reference's host max keeps a NaN in element 0, time_integrator.cu:208-211).
__host__ __device__ __forceinline__ unsigned long long ord_key(double x) {
  unsigned long long b;
  __builtin_memcpy(&b, &x, sizeof b);
  return (b >> 63) ? ~b : (b | 0x8000000000000000ull);
}
__host__ __device__ __forceinline__ double ord_val(unsigned long long k) {
  unsigned long long b = (k >> 63) ? (k & 0x7fffffffffffffffull) : ~k;
  double x;
  __builtin_memcpy(&x, &b, sizeof x);
  return x;
}

__device__ __forceinline__ unsigned long long wave_max_key(unsigned long long v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    unsigned long long w = __shfl_xor(v, o, 64);
    v = v > w ? v : w;
  }
  return v;
}

// loads and stores through the global address space (the pointers live in a
// by-value argument struct, where the compiler cannot infer it)
typedef __attribute__((address_space(1))) const double gcdouble;
typedef __attribute__((address_space(1))) double gdouble;
// Element indices are below 2^28 (checked on the host), so masking makes the
// byte offset a provable 32-bit value: the loads then take a scalar base
// register plus one 32-bit vector offset instead of a 64-bit address per lane.
constexpr unsigned kIdxMask = 0x0fffffffu;
__device__ __forceinline__ double gld(const double *p, int i) {
  return ((gcdouble *)p)[(unsigned)i & kIdxMask];
}
__device__ __forceinline__ void gst(double *p, int i, double v) {
  ((gdouble *)p)[(unsigned)i & kIdxMask] = v;
}
// a store that is not kept in L2 (streamed to HBM): for the per-element
// outputs of the gather kernels, whose L2 lines are better spent on the
// neighbours' coefficients that the next elements read
__device__ __forceinline__ void gst_nt(double *p, int i, double v) {
  __builtin_nontemporal_store(v, &((gdouble *)p)[(unsigned)i & kIdxMask]);
}

// 16-byte loads and nontemporal stores of two doubles at an even index
typedef double dbl2 __attribute__((ext_vector_type(2)));
typedef __attribute__((address_space(1))) const dbl2 gcdbl2;
typedef __attribute__((address_space(1))) dbl2 gdbl2;
__device__ __forceinline__ dbl2 gld2(const double *p, int i) {
  return *(gcdbl2 *)((gcdouble *)p + ((unsigned)i & kIdxMask));
}
__device__ __forceinline__ void gst2_nt(double *p, int i, dbl2 v) {
  __builtin_nontemporal_store(v, (gdbl2 *)((gdouble *)p + ((unsigned)i & kIdxMask)));
}

// Materialise the accumulators here: the compiler would otherwise sink every
// flux evaluation to the end of the kernel and keep all traces live at once.
template <int N, int NP>
__device__ __forceinline__ void pin(double (&acc)[N][NP]) {
#pragma unroll
  for (int n = 0; n < N; n++)
#pragma unroll
    for (int i = 0; i < NP; i++) asm volatile("" : "+v"(acc[n][i]));
}
template <int NP>
__device__ __forceinline__ void pin(double (&acc)[NP]) {
#pragma unroll
  for (int i = 0; i < NP; i++) asm volatile("" : "+v"(acc[i]));
}

}  // namespace dgk
