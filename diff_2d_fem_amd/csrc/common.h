// common.h — error reporting and wavefront helpers shared by the library's translation units.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stdio.h>

#include "../../include/d2fem.h"
#include "d2f_math.h"

namespace d2f {

void set_error(const char *fmt, ...);

#define D2F_HIP(call)                                                              \
  do {                                                                             \
    hipError_t e_ = (call);                                                        \
    if (e_ != hipSuccess) {                                                        \
      ::d2f::set_error("%s failed: %s (%s:%d)", #call, hipGetErrorString(e_),       \
                       __FILE__, __LINE__);                                        \
      return D2F_EHIP;                                                             \
    }                                                                              \
  } while (0)

#define D2F_CHECK(cond, ...)            \
  do {                                  \
    if (!(cond)) {                      \
      ::d2f::set_error(__VA_ARGS__);    \
      return D2F_EINVAL;                \
    }                                   \
  } while (0)

constexpr int WAVE = 64;

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, WAVE);
  return v;
}
__device__ __forceinline__ double wave_min(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = fmin(v, __shfl_xor(v, o, WAVE));
  return v;
}
__device__ __forceinline__ double wave_max(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, WAVE));
  return v;
}

// Per-environment accumulation from a wavefront whose lanes mostly share one environment
// (items are laid out by environment): one atomic per wave when they all do, one per lane
// otherwise. Must be reached by every lane of the wave.
__device__ __forceinline__ void env_add(double *arr, int env, double v, bool valid) {
  const unsigned long long m = __ballot(valid);
  if (m == 0ull) return;
  const int lead = __ffsll((long long)m) - 1;
  const int e0 = __shfl(env, lead, WAVE);
  if (__all(!valid || env == e0)) {
    const double s = wave_sum(valid ? v : 0.0);
    if ((int)(threadIdx.x % WAVE) == lead) atomicAdd(arr + e0, s);
  } else if (valid) {
    atomicAdd(arr + env, v);
  }
}
__device__ __forceinline__ void env_min(double *arr, int env, double v, bool valid) {
  const unsigned long long m = __ballot(valid);
  if (m == 0ull) return;
  const int lead = __ffsll((long long)m) - 1;
  const int e0 = __shfl(env, lead, WAVE);
  if (__all(!valid || env == e0)) {
    const double s = wave_min(valid ? v : BIG * 10.0);
    if ((int)(threadIdx.x % WAVE) == lead) atomic_min_d(arr + e0, s);
  } else if (valid) {
    atomic_min_d(arr + env, v);
  }
}
__device__ __forceinline__ void env_max(double *arr, int env, double v, bool valid) {
  const unsigned long long m = __ballot(valid);
  if (m == 0ull) return;
  const int lead = __ffsll((long long)m) - 1;
  const int e0 = __shfl(env, lead, WAVE);
  if (__all(!valid || env == e0)) {
    const double s = wave_max(valid ? v : -BIG * 10.0);
    if ((int)(threadIdx.x % WAVE) == lead) atomic_max_d(arr + e0, s);
  } else if (valid) {
    atomic_max_d(arr + env, v);
  }
}

__host__ __device__ __forceinline__ Collider to_collider(const d2f_collider &c) {
  return Collider{c.kind, c.cx, c.cy, c.rx, c.ry, c.rref, c.r, c.mu};
}

}  // namespace d2f
