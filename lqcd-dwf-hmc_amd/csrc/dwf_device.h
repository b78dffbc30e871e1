// dwf_device.h -- device-side types shared by the HIP sources of libdwfhip.
//
// Storage traits: a field is stored as T (double, float, __half, bfloat16) and
// computed in C (double for T = double, float otherwise).
#pragma once

#include <hip/hip_runtime.h>
#include <hip/hip_fp16.h>
#include <hip/hip_bf16.h>
#include <stdint.h>

namespace dwf {

template <typename T> struct St;
template <> struct St<double> {
  using C = double;
  static __device__ __forceinline__ double ld(const double* p) { return *p; }
  static __device__ __forceinline__ void st(double* p, double v) { *p = v; }
};
template <> struct St<float> {
  using C = float;
  static __device__ __forceinline__ float ld(const float* p) { return *p; }
  static __device__ __forceinline__ void st(float* p, float v) { *p = v; }
};
template <> struct St<__half> {
  using C = float;
  static __device__ __forceinline__ float ld(const __half* p) { return __half2float(*p); }
  static __device__ __forceinline__ void st(__half* p, float v) { *p = __float2half(v); }
};
template <> struct St<__hip_bfloat16> {
  using C = float;
  static __device__ __forceinline__ float ld(const __hip_bfloat16* p) { return __bfloat162float(*p); }
  static __device__ __forceinline__ void st(__hip_bfloat16* p, float v) { *p = __float2bfloat16(v); }
};

// Kernel-side geometry of one even-odd half lattice.
struct Geo {
  int Vh;   // half volume (4D)
  int Ls;   // extent of the 5th dimension
};

}  // namespace dwf
