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

// Vx<T, V>: V consecutive elements (aligned to V) moved with one vector access.
// R is the raw register form (packed for 16-bit storage); get() unpacks one
// element to the compute type, ld/stn move all V.
typedef double dbl2_t __attribute__((ext_vector_type(2)));
typedef float flt2_t __attribute__((ext_vector_type(2)));
typedef float flt4_t __attribute__((ext_vector_type(4)));
typedef unsigned int u32x2_t __attribute__((ext_vector_type(2)));
typedef unsigned int u32x4_t __attribute__((ext_vector_type(4)));

template <typename T, int V> struct Vx;
template <> struct Vx<double, 1> {
  using R = double;
  static __device__ __forceinline__ R ld(const double* p) { return *p; }
  static __device__ __forceinline__ double get(const R& r, int) { return r; }
  static __device__ __forceinline__ void stn(double* p, const double* v) {
    __builtin_nontemporal_store(v[0], p);
  }
};
template <> struct Vx<double, 2> {
  using R = dbl2_t;
  static __device__ __forceinline__ R ld(const double* p) { return *reinterpret_cast<const R*>(p); }
  static __device__ __forceinline__ double get(const R& r, int v) { return r[v]; }
  static __device__ __forceinline__ void stn(double* p, const double* v) {
    R t = {v[0], v[1]};
    __builtin_nontemporal_store(t, reinterpret_cast<R*>(p));
  }
};
template <> struct Vx<float, 1> {
  using R = float;
  static __device__ __forceinline__ R ld(const float* p) { return *p; }
  static __device__ __forceinline__ float get(const R& r, int) { return r; }
  static __device__ __forceinline__ void stn(float* p, const float* v) {
    __builtin_nontemporal_store(v[0], p);
  }
};
template <> struct Vx<float, 2> {
  using R = flt2_t;
  static __device__ __forceinline__ R ld(const float* p) { return *reinterpret_cast<const R*>(p); }
  static __device__ __forceinline__ float get(const R& r, int v) { return r[v]; }
  static __device__ __forceinline__ void stn(float* p, const float* v) {
    R t = {v[0], v[1]};
    __builtin_nontemporal_store(t, reinterpret_cast<R*>(p));
  }
};
template <> struct Vx<float, 4> {
  using R = flt4_t;
  static __device__ __forceinline__ R ld(const float* p) { return *reinterpret_cast<const R*>(p); }
  static __device__ __forceinline__ float get(const R& r, int v) { return r[v]; }
  static __device__ __forceinline__ void stn(float* p, const float* v) {
    R t = {v[0], v[1], v[2], v[3]};
    __builtin_nontemporal_store(t, reinterpret_cast<R*>(p));
  }
};
// 16-bit storage: 2 elements per 32-bit word, element 0 in the low half
__device__ __forceinline__ float h_lo(unsigned w) { return __half2float(__ushort_as_half((unsigned short)(w & 0xffffu))); }
__device__ __forceinline__ float h_hi(unsigned w) { return __half2float(__ushort_as_half((unsigned short)(w >> 16))); }
__device__ __forceinline__ unsigned h_pack(float a, float b) {
  return (unsigned)__half_as_ushort(__float2half(a)) | ((unsigned)__half_as_ushort(__float2half(b)) << 16);
}
__device__ __forceinline__ float b_lo(unsigned w) { return __uint_as_float(w << 16); }
__device__ __forceinline__ float b_hi(unsigned w) { return __uint_as_float(w & 0xffff0000u); }
__device__ __forceinline__ unsigned b_pack(float a, float b) {
  const __hip_bfloat16 x = __float2bfloat16(a), y = __float2bfloat16(b);
  return (unsigned)*reinterpret_cast<const unsigned short*>(&x) |
         ((unsigned)*reinterpret_cast<const unsigned short*>(&y) << 16);
}
template <typename T> struct Pk16;
template <> struct Pk16<__half> {
  static __device__ __forceinline__ float lo(unsigned w) { return h_lo(w); }
  static __device__ __forceinline__ float hi(unsigned w) { return h_hi(w); }
  static __device__ __forceinline__ unsigned pack(float a, float b) { return h_pack(a, b); }
};
template <> struct Pk16<__hip_bfloat16> {
  static __device__ __forceinline__ float lo(unsigned w) { return b_lo(w); }
  static __device__ __forceinline__ float hi(unsigned w) { return b_hi(w); }
  static __device__ __forceinline__ unsigned pack(float a, float b) { return b_pack(a, b); }
};
// 16-bit outputs are written as 4- or 8-byte pieces of 32-byte rows: they go
// through L2 as ordinary stores, so that the pieces of a line are merged before
// the line is written back (non-temporal stores wrote 1.33x the bytes).
#ifndef DWF_NT_STORE16
#define DWF_NT_STORE16 0
#endif
template <typename R> __device__ __forceinline__ void st16(R* p, R v) {
#if DWF_NT_STORE16
  __builtin_nontemporal_store(v, p);
#else
  *p = v;
#endif
}
template <typename T> struct Vx16_2 {
  using R = unsigned;
  static __device__ __forceinline__ R ld(const T* p) { return *reinterpret_cast<const unsigned*>(p); }
  static __device__ __forceinline__ float get(const R& r, int v) {
    return v == 0 ? Pk16<T>::lo(r) : Pk16<T>::hi(r);
  }
  static __device__ __forceinline__ void stn(T* p, const float* v) {
    st16(reinterpret_cast<unsigned*>(p), Pk16<T>::pack(v[0], v[1]));
  }
};
template <typename T> struct Vx16_4 {
  using R = u32x2_t;
  static __device__ __forceinline__ R ld(const T* p) { return *reinterpret_cast<const R*>(p); }
  static __device__ __forceinline__ float get(const R& r, int v) {
    const unsigned w = r[v >> 1];
    return (v & 1) ? Pk16<T>::hi(w) : Pk16<T>::lo(w);
  }
  static __device__ __forceinline__ void stn(T* p, const float* v) {
    R t = {Pk16<T>::pack(v[0], v[1]), Pk16<T>::pack(v[2], v[3])};
    st16(reinterpret_cast<R*>(p), t);
  }
};
template <> struct Vx<__half, 2> : Vx16_2<__half> {};
template <> struct Vx<__half, 4> : Vx16_4<__half> {};
template <> struct Vx<__hip_bfloat16, 2> : Vx16_2<__hip_bfloat16> {};
template <> struct Vx<__hip_bfloat16, 4> : Vx16_4<__hip_bfloat16> {};
template <> struct Vx<__half, 1> {
  using R = float;
  static __device__ __forceinline__ R ld(const __half* p) { return St<__half>::ld(p); }
  static __device__ __forceinline__ float get(const R& r, int) { return r; }
  static __device__ __forceinline__ void stn(__half* p, const float* v) { St<__half>::stn(p, v[0]); }
};
template <> struct Vx<__hip_bfloat16, 1> {
  using R = float;
  static __device__ __forceinline__ R ld(const __hip_bfloat16* p) { return St<__hip_bfloat16>::ld(p); }
  static __device__ __forceinline__ float get(const R& r, int) { return r; }
  static __device__ __forceinline__ void stn(__hip_bfloat16* p, const float* v) {
    St<__hip_bfloat16>::stn(p, v[0]);
  }
};

// Kernel-side geometry of one even-odd half lattice.
struct Geo {
  int Vh;   // half volume (4D)
  int Ls;   // extent of the 5th dimension
};

}  // namespace dwf
