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
// ld/st: ordinary cached access.  ldn/stn: streaming (non-temporal) access for
// data touched once, so that it does not evict reusable lines from L2.
template <> struct St<double> {
  using C = double;
  static __device__ __forceinline__ double ld(const double* p) { return *p; }
  static __device__ __forceinline__ void st(double* p, double v) { *p = v; }
  static __device__ __forceinline__ double ldn(const double* p) { return __builtin_nontemporal_load(p); }
  static __device__ __forceinline__ void stn(double* p, double v) { __builtin_nontemporal_store(v, p); }
};
template <> struct St<float> {
  using C = float;
  static __device__ __forceinline__ float ld(const float* p) { return *p; }
  static __device__ __forceinline__ void st(float* p, float v) { *p = v; }
  static __device__ __forceinline__ float ldn(const float* p) { return __builtin_nontemporal_load(p); }
  static __device__ __forceinline__ void stn(float* p, float v) { __builtin_nontemporal_store(v, p); }
};
template <> struct St<__half> {
  using C = float;
  static __device__ __forceinline__ float ld(const __half* p) { return __half2float(*p); }
  static __device__ __forceinline__ void st(__half* p, float v) { *p = __float2half(v); }
  static __device__ __forceinline__ float ldn(const __half* p) {
    unsigned short b = __builtin_nontemporal_load(reinterpret_cast<const unsigned short*>(p));
    return __half2float(__ushort_as_half(b));
  }
  static __device__ __forceinline__ void stn(__half* p, float v) {
    __builtin_nontemporal_store(__half_as_ushort(__float2half(v)), reinterpret_cast<unsigned short*>(p));
  }
};
template <> struct St<__hip_bfloat16> {
  using C = float;
  static __device__ __forceinline__ float ld(const __hip_bfloat16* p) { return __bfloat162float(*p); }
  static __device__ __forceinline__ void st(__hip_bfloat16* p, float v) { *p = __float2bfloat16(v); }
  static __device__ __forceinline__ float ldn(const __hip_bfloat16* p) {
    unsigned short b = __builtin_nontemporal_load(reinterpret_cast<const unsigned short*>(p));
    return __uint_as_float((unsigned)b << 16);
  }
  static __device__ __forceinline__ void stn(__hip_bfloat16* p, float v) {
    __hip_bfloat16 h = __float2bfloat16(v);
    __builtin_nontemporal_store(*reinterpret_cast<const unsigned short*>(&h),
                                reinterpret_cast<unsigned short*>(p));
  }
};

// Kernel-side geometry of one even-odd half lattice.
struct Geo {
  int Vh;   // half volume (4D)
  int Ls;   // extent of the 5th dimension
};

}  // namespace dwf
