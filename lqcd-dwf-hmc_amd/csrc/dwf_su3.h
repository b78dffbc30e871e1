// dwf_su3.h -- 3x3 complex matrix helpers of the FP64 gauge kernels
// (smearing, its force, the HMC gauge action and update).  Matrices are the
// corelib Mat_SU_N layout: 18 doubles, element (a, b) at 2 (3 a + b) (re, im).
#pragma once

#include <hip/hip_runtime.h>

namespace dwf {

struct cplx {
  double r, i;
};
__device__ __forceinline__ cplx cadd(cplx a, cplx b) { return {a.r + b.r, a.i + b.i}; }
__device__ __forceinline__ cplx cmul(cplx a, cplx b) {
  return {a.r * b.r - a.i * b.i, a.r * b.i + a.i * b.r};
}
__device__ __forceinline__ cplx cscal(double s, cplx a) { return {s * a.r, s * a.i}; }

struct M3 {
  cplx a[3][3];
};

__device__ __forceinline__ M3 load3(const double* p) {
  M3 m;
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) m.a[i][j] = {p[2 * (3 * i + j)], p[2 * (3 * i + j) + 1]};
  return m;
}
__device__ __forceinline__ void store3(double* p, const M3& m) {
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      p[2 * (3 * i + j)] = m.a[i][j].r;
      p[2 * (3 * i + j) + 1] = m.a[i][j].i;
    }
}
__device__ __forceinline__ M3 mul_nn(const M3& x, const M3& y) {
  M3 r;
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      cplx s = cmul(x.a[i][0], y.a[0][j]);
      s = cadd(s, cmul(x.a[i][1], y.a[1][j]));
      s = cadd(s, cmul(x.a[i][2], y.a[2][j]));
      r.a[i][j] = s;
    }
  return r;
}
// x * y^+
__device__ __forceinline__ M3 mul_nd(const M3& x, const M3& y) {
  M3 r;
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      cplx s = {0, 0};
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        const cplx a = x.a[i][k], b = y.a[j][k];
        s.r += a.r * b.r + a.i * b.i;
        s.i += a.i * b.r - a.r * b.i;
      }
      r.a[i][j] = s;
    }
  return r;
}
// x^+ * y
__device__ __forceinline__ M3 mul_dn(const M3& x, const M3& y) {
  M3 r;
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      cplx s = {0, 0};
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        const cplx a = x.a[k][i], b = y.a[k][j];
        s.r += a.r * b.r + a.i * b.i;
        s.i += a.r * b.i - a.i * b.r;
      }
      r.a[i][j] = s;
    }
  return r;
}

__device__ __forceinline__ M3 dag3(const M3& x) {
  M3 r;
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) r.a[i][j] = {x.a[j][i].r, -x.a[j][i].i};
  return r;
}

// anti-hermitian traceless part: at(m) = (m - m^+)/2 - tr(m - m^+)/6
__device__ __forceinline__ M3 at3(const M3& m) {
  M3 r;
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j)
      r.a[i][j] = {0.5 * (m.a[i][j].r - m.a[j][i].r), 0.5 * (m.a[i][j].i + m.a[j][i].i)};
  const double tr = (r.a[0][0].i + r.a[1][1].i + r.a[2][2].i) / 3.0;
#pragma unroll
  for (int i = 0; i < 3; ++i) r.a[i][i].i -= tr;
  return r;
}

// lexical neighbour of site s in direction mu (dir = +1 or -1), periodic
__device__ __forceinline__ size_t shift_site(size_t s, int mu, int dir, int Nx, int Ny, int Nz,
                                             int Nt) {
  int x = (int)(s % Nx);
  size_t q = s / Nx;
  int y = (int)(q % Ny);
  q /= Ny;
  int z = (int)(q % Nz);
  int t = (int)(q / Nz);
  if (mu == 0) x = (x + dir + Nx) % Nx;
  else if (mu == 1) y = (y + dir + Ny) % Ny;
  else if (mu == 2) z = (z + dir + Nz) % Nz;
  else t = (t + dir + Nt) % Nt;
  return x + (size_t)Nx * (y + (size_t)Ny * (z + (size_t)Nz * t));
}

}  // namespace dwf
