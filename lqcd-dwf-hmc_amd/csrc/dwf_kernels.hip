// dwf_kernels.hip -- HIP kernels of libdwfhip for gfx950 (MI355X).
//
// The even-odd domain-wall operator is built from one fused kernel, k_dwf:
//   core(site, s) = hop(in)            4D Wilson hop, opposite-parity neighbours
//                 = g5 hop(g5 in)      its adjoint on the other parity
//                 = in(site, s)        site-local
//   out = xcoef * X + E(core)          E: per-chirality Ls x Ls matrix (fused)
//   out2 = F(out)                      optional second per-chirality matrix
// Every Ls-mixing part of the operator (the M5 part of D_eo, D_ee, its inverse,
// gamma_5 R and their products) is such a per-chirality real matrix, so the
// whole operator is a chain of k_dwf launches.
//
// Thread mapping: one thread per (4D half site, s); a block holds whole sites
// (spb sites x Ls threads), so the Ls-mixing epilogue runs through LDS inside
// the block.  The Ls threads of a site read contiguous spinor memory and share
// (by broadcast) the site's links.
#include <type_traits>
#include <vector>

#include "dwf_device.h"
#include "dwf_kernels.h"

// Hop loads of the next direction issued before the current one is computed.
#ifndef DWF_HOP_PIPE
#define DWF_HOP_PIPE 0
#endif

namespace dwf {

// Orders this wave's LDS stores before its later LDS loads in every lane
// (data handed between lanes of one wave through LDS).
__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  return v;
}

// Sum of n partials by one block of up to 8 waves, in a fixed order (thread t
// adds t, t + blockDim, ...; then the wave sums in wave order); with 256
// threads it is the order of k_red_finish.  sh holds 8 doubles.
__device__ __forceinline__ double block_sum(const double* __restrict__ p, unsigned n,
                                            double* sh) {
  double w = 0;
  for (unsigned i = threadIdx.x; i < n; i += blockDim.x) w += p[i];
  const double t = wave_sum(w);
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = t;
  __syncthreads();
  double r = 0;
  for (unsigned k = 0; k < (blockDim.x >> 6); ++k) r += sh[k];
  __syncthreads();
  return r;
}

size_t prec_bytes(int prec) {
  switch (prec) {
    case P64: return 8;
    case P32: return 4;
    default: return 2;
  }
}

// ---------------------------------------------------------------------------
// spin projection / reconstruction, Dirac basis.  dir = 2*mu + fb,
// fb = 0: (1 - gamma_mu) U_mu(x) psi(x+mu), fb = 1: (1 + gamma_mu) U^+ psi(x-mu).
// w[k], k = ri + 2*(c + 3*d).
// ---------------------------------------------------------------------------
template <int DIR, typename C>
__device__ __forceinline__ void project(const C* w, C* h1, C* h2) {
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    const int r = 2 * c, i = 2 * c + 1;
    const C* w0 = w;
    const C* w1 = w + 6;
    const C* w2 = w + 12;
    const C* w3 = w + 18;
    if (DIR == 0) {         // h1 = w0 + i w3, h2 = w1 + i w2
      h1[r] = w0[r] - w3[i]; h1[i] = w0[i] + w3[r];
      h2[r] = w1[r] - w2[i]; h2[i] = w1[i] + w2[r];
    } else if (DIR == 1) {  // h1 = w0 - i w3, h2 = w1 - i w2
      h1[r] = w0[r] + w3[i]; h1[i] = w0[i] - w3[r];
      h2[r] = w1[r] + w2[i]; h2[i] = w1[i] - w2[r];
    } else if (DIR == 2) {  // h1 = w0 + w3, h2 = w1 - w2
      h1[r] = w0[r] + w3[r]; h1[i] = w0[i] + w3[i];
      h2[r] = w1[r] - w2[r]; h2[i] = w1[i] - w2[i];
    } else if (DIR == 3) {  // h1 = w0 - w3, h2 = w1 + w2
      h1[r] = w0[r] - w3[r]; h1[i] = w0[i] - w3[i];
      h2[r] = w1[r] + w2[r]; h2[i] = w1[i] + w2[i];
    } else if (DIR == 4) {  // h1 = w0 + i w2, h2 = w1 - i w3
      h1[r] = w0[r] - w2[i]; h1[i] = w0[i] + w2[r];
      h2[r] = w1[r] + w3[i]; h2[i] = w1[i] - w3[r];
    } else if (DIR == 5) {  // h1 = w0 - i w2, h2 = w1 + i w3
      h1[r] = w0[r] + w2[i]; h1[i] = w0[i] - w2[r];
      h2[r] = w1[r] - w3[i]; h2[i] = w1[i] + w3[r];
    } else if (DIR == 6) {  // h1 = 2 w2, h2 = 2 w3
      h1[r] = w2[r] + w2[r]; h1[i] = w2[i] + w2[i];
      h2[r] = w3[r] + w3[r]; h2[i] = w3[i] + w3[i];
    } else {                // h1 = 2 w0, h2 = 2 w1
      h1[r] = w0[r] + w0[r]; h1[i] = w0[i] + w0[i];
      h2[r] = w1[r] + w1[r]; h2[i] = w1[i] + w1[i];
    }
  }
}

template <int DIR, typename C>
__device__ __forceinline__ void reconstruct(C* o, const C* a1, const C* a2) {
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    const int r = 2 * c, i = 2 * c + 1;
    C* o0 = o;
    C* o1 = o + 6;
    C* o2 = o + 12;
    C* o3 = o + 18;
    if (DIR != 6) {
      o0[r] += a1[r]; o0[i] += a1[i];
      o1[r] += a2[r]; o1[i] += a2[i];
    }
    if (DIR == 0) {         // o2 += -i a2, o3 += -i a1
      o2[r] += a2[i]; o2[i] -= a2[r]; o3[r] += a1[i]; o3[i] -= a1[r];
    } else if (DIR == 1) {  // o2 += i a2, o3 += i a1
      o2[r] -= a2[i]; o2[i] += a2[r]; o3[r] -= a1[i]; o3[i] += a1[r];
    } else if (DIR == 2) {  // o2 -= a2, o3 += a1
      o2[r] -= a2[r]; o2[i] -= a2[i]; o3[r] += a1[r]; o3[i] += a1[i];
    } else if (DIR == 3) {  // o2 += a2, o3 -= a1
      o2[r] += a2[r]; o2[i] += a2[i]; o3[r] -= a1[r]; o3[i] -= a1[i];
    } else if (DIR == 4) {  // o2 += -i a1, o3 += i a2
      o2[r] += a1[i]; o2[i] -= a1[r]; o3[r] -= a2[i]; o3[i] += a2[r];
    } else if (DIR == 5) {  // o2 += i a1, o3 += -i a2
      o2[r] -= a1[i]; o2[i] += a1[r]; o3[r] += a2[i]; o3[i] -= a2[r];
    } else if (DIR == 6) {  // lower components only
      o2[r] += a1[r]; o2[i] += a1[i]; o3[r] += a2[r]; o3[i] += a2[i];
    }
    // DIR == 7: upper components only (done above)
  }
}

// a1 = U h1, a2 = U h2 with U read from LDS element by element (row-major
// complex, entry ab = ri + 2*(b + 3*a) at up[ab * stride]): each entry is used
// for both half spinors and dies right after, so the link never occupies 18
// registers.
template <typename C>
__device__ __forceinline__ void su3_mul2_lds(const C* up, int stride, const C* h1, const C* h2,
                                             C* a1, C* a2) {
#pragma unroll
  for (int ia = 0; ia < 3; ++ia) {
    C r1 = 0, i1 = 0, r2 = 0, i2 = 0;
#pragma unroll
    for (int ib = 0; ib < 3; ++ib) {
      const C ur = up[(2 * (ib + 3 * ia)) * stride], ui = up[(2 * (ib + 3 * ia) + 1) * stride];
      r1 += ur * h1[2 * ib] - ui * h1[2 * ib + 1];
      i1 += ur * h1[2 * ib + 1] + ui * h1[2 * ib];
      r2 += ur * h2[2 * ib] - ui * h2[2 * ib + 1];
      i2 += ur * h2[2 * ib + 1] + ui * h2[2 * ib];
    }
    a1[2 * ia] = r1;
    a1[2 * ia + 1] = i1;
    a2[2 * ia] = r2;
    a2[2 * ia + 1] = i2;
  }
}

// a = U h for a 3x3 row-major complex matrix u[ab], ab = ri + 2*(b + 3*a)
template <typename C>
__device__ __forceinline__ void su3_mul(const C* u, const C* h, C* a) {
#pragma unroll
  for (int ia = 0; ia < 3; ++ia) {
    C re = 0, im = 0;
#pragma unroll
    for (int ib = 0; ib < 3; ++ib) {
      const C ur = u[2 * (ib + 3 * ia)], ui = u[2 * (ib + 3 * ia) + 1];
      re += ur * h[2 * ib] - ui * h[2 * ib + 1];
      im += ur * h[2 * ib + 1] + ui * h[2 * ib];
    }
    a[2 * ia] = re;
    a[2 * ia + 1] = im;
  }
}

// One hop direction for thread (site, s0 .. s0+VEC-1).  The block's links
// (already in the compute type) and neighbour indices sit in LDS:
// shL[(dir*18 + e)*spb + lsite], shN[dir*spb + lsite].  The VEC s-values of a
// spinor component are one vector load; the link is read once for all VEC.
template <int DIR, bool DAG, typename T, int VEC, typename C>
__device__ __forceinline__ void hop_dir(C (*acc)[24], const T* __restrict__ in, const C* shL,
                                        const int* shN, int lsite, int s0, int spb, int Ls) {
  using VL = Vx<T, VEC>;
  const int nb = shN[DIR * spb + lsite];
  const T* wp = in + (size_t)nb * 24 * Ls + s0;
  typename VL::R raw[24];
  if (!DAG) {
#pragma unroll
    for (int k = 0; k < 24; ++k) raw[k] = VL::ld(wp + k * Ls);
  } else {  // gamma_5 (Dirac basis) exchanges the upper and lower spin pairs
#pragma unroll
    for (int k = 0; k < 12; ++k) {
      raw[k] = VL::ld(wp + (k + 12) * Ls);
      raw[k + 12] = VL::ld(wp + k * Ls);
    }
  }
  const C* up = shL + DIR * 18 * spb + lsite;
#pragma unroll
  for (int v = 0; v < VEC; ++v) {
    C w[24];
#pragma unroll
    for (int k = 0; k < 24; ++k) w[k] = VL::get(raw[k], v);
    C h1[6], h2[6];
    project<DIR>(w, h1, h2);
    C a1[6], a2[6];
    su3_mul2_lds(up, spb, h1, h2, a1, a2);
    reconstruct<DIR>(acc[v], a1, a2);
  }
}

// The same split in two halves, so that the loads of the next direction can be
// in flight while the current one is computed (software pipelining).
template <int DIR, bool DAG, typename T, int VEC>
__device__ __forceinline__ void hop_load(typename Vx<T, VEC>::R (&raw)[24], const T* __restrict__ in,
                                         const int* shN, int lsite, int s0, int spb, int Ls) {
  using VL = Vx<T, VEC>;
  const int nb = shN[DIR * spb + lsite];
  const T* wp = in + (size_t)nb * 24 * Ls + s0;
  if (!DAG) {
#pragma unroll
    for (int k = 0; k < 24; ++k) raw[k] = VL::ld(wp + k * Ls);
  } else {
#pragma unroll
    for (int k = 0; k < 12; ++k) {
      raw[k] = VL::ld(wp + (k + 12) * Ls);
      raw[k + 12] = VL::ld(wp + k * Ls);
    }
  }
}
template <int DIR, typename T, int VEC, typename C>
__device__ __forceinline__ void hop_compute(C (*acc)[24], const typename Vx<T, VEC>::R (&raw)[24],
                                            const C* shL, int lsite, int spb) {
  using VL = Vx<T, VEC>;
  const C* up = shL + DIR * 18 * spb + lsite;
  C u[18];
#pragma unroll
  for (int e = 0; e < 18; ++e) u[e] = up[e * spb];
#pragma unroll
  for (int v = 0; v < VEC; ++v) {
    C w[24];
#pragma unroll
    for (int k = 0; k < 24; ++k) w[k] = VL::get(raw[k], v);
    C h1[6], h2[6];
    project<DIR>(w, h1, h2);
    C a1[6], a2[6];
    su3_mul(u, h1, a1);
    su3_mul(u, h2, a2);
    reconstruct<DIR>(acc[v], a1, a2);
  }
}
// FP16 storage, two s per thread: the hop on packed halves.  The spin
// projection runs on (s0, s1) pairs (v_pk_add_f16, one rounding to FP16 like
// the stored field itself), the pairs are regrouped to (re, im) per s, and the
// SU(3) product is a chain of v_dot2_f32_f16 with FP32 accumulation:
//   Re(u h) += dot2((ur, -ui), (hr, hi)),  Im(u h) += dot2((ui, ur), (hr, hi)).
// The block's links sit in LDS as those two words per entry, site-major with
// kPkStride words per site (8 directions x 20 words, 18 used, + 4 pad words so
// that the 16-byte reads of 8 sites of a wave cover all 32 banks once).
typedef _Float16 f16x2_t __attribute__((ext_vector_type(2)));
constexpr int kPkStride = 164;
__device__ __forceinline__ f16x2_t as_f16x2(unsigned w) { return __builtin_bit_cast(f16x2_t, w); }
__device__ __forceinline__ unsigned as_u32(f16x2_t h) { return __builtin_bit_cast(unsigned, h); }
// first product of a chain: the three-operand form with an inline zero (the
// builtin always becomes the accumulating two-operand form behind a register
// move of the zero)
__device__ __forceinline__ float dot2_first(f16x2_t a, f16x2_t b) {
  float r;
  asm("v_dot2_f32_f16 %0, %1, %2, 0" : "=v"(r) : "v"(a), "v"(b));
  return r;
}

template <int DIR, bool DAG, int LSC>
__device__ __forceinline__ void pk_load(f16x2_t (&w)[24], const __half* __restrict__ in,
                                        const int* shN, int lsite, int s0, int spb) {
  constexpr int Ls = LSC;
  const int nb = shN[DIR * spb + lsite];
  const __half* wp = in + (size_t)nb * 24 * Ls + s0;
  if (!DAG) {
#pragma unroll
    for (int k = 0; k < 24; ++k) w[k] = as_f16x2(*reinterpret_cast<const unsigned*>(wp + k * Ls));
  } else {  // gamma_5 (Dirac basis) exchanges the upper and lower spin pairs
#pragma unroll
    for (int k = 0; k < 12; ++k) {
      w[k] = as_f16x2(*reinterpret_cast<const unsigned*>(wp + (k + 12) * Ls));
      w[k + 12] = as_f16x2(*reinterpret_cast<const unsigned*>(wp + k * Ls));
    }
  }
}
template <int DIR>
__device__ __forceinline__ void pk_compute(float (*acc)[24], const f16x2_t (&w)[24],
                                           const unsigned* shP, int lsite) {
  f16x2_t h1[6], h2[6];
  project<DIR>(w, h1, h2);
  // (re, im) of colour c for s0 (p[0]) and s0 + 1 (p[1])
  unsigned p1[2][3], p2[2][3];
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    const unsigned r1 = as_u32(h1[2 * c]), i1 = as_u32(h1[2 * c + 1]);
    const unsigned r2 = as_u32(h2[2 * c]), i2 = as_u32(h2[2 * c + 1]);
    p1[0][c] = __builtin_amdgcn_perm(i1, r1, 0x05040100u);
    p1[1][c] = __builtin_amdgcn_perm(i1, r1, 0x07060302u);
    p2[0][c] = __builtin_amdgcn_perm(i2, r2, 0x05040100u);
    p2[1][c] = __builtin_amdgcn_perm(i2, r2, 0x07060302u);
  }
  const u32x4_t* up = reinterpret_cast<const u32x4_t*>(shP + lsite * kPkStride + DIR * 20);
  unsigned u[20];
#pragma unroll
  for (int q = 0; q < 5; ++q) {
    const u32x4_t t = up[q];
    u[4 * q] = t[0]; u[4 * q + 1] = t[1]; u[4 * q + 2] = t[2]; u[4 * q + 3] = t[3];
  }
#pragma unroll
  for (int v = 0; v < 2; ++v) {
    float a1[6], a2[6];
#pragma unroll
    for (int ia = 0; ia < 3; ++ia) {
      float r1, i1, r2, i2;
#pragma unroll
      for (int ib = 0; ib < 3; ++ib) {
        const f16x2_t X = as_f16x2(u[2 * (ib + 3 * ia)]), Y = as_f16x2(u[2 * (ib + 3 * ia) + 1]);
        const f16x2_t q1 = as_f16x2(p1[v][ib]), q2 = as_f16x2(p2[v][ib]);
        if (ib == 0) {
          r1 = dot2_first(X, q1);
          i1 = dot2_first(Y, q1);
          r2 = dot2_first(X, q2);
          i2 = dot2_first(Y, q2);
        } else {
          r1 = __builtin_amdgcn_fdot2(X, q1, r1, false);
          i1 = __builtin_amdgcn_fdot2(Y, q1, i1, false);
          r2 = __builtin_amdgcn_fdot2(X, q2, r2, false);
          i2 = __builtin_amdgcn_fdot2(Y, q2, i2, false);
        }
      }
      a1[2 * ia] = r1; a1[2 * ia + 1] = i1;
      a2[2 * ia] = r2; a2[2 * ia + 1] = i2;
    }
    reconstruct<DIR>(acc[v], a1, a2);
  }
}
template <bool DAG, int LSC>
__device__ __forceinline__ void hop_all_pk(float (*acc)[24], const __half* __restrict__ in,
                                           const unsigned* shP, const int* shN, int lsite,
                                           int s0, int spb) {
  f16x2_t wa[24];
  pk_load<0, DAG, LSC>(wa, in, shN, lsite, s0, spb); pk_compute<0>(acc, wa, shP, lsite);
  pk_load<1, DAG, LSC>(wa, in, shN, lsite, s0, spb); pk_compute<1>(acc, wa, shP, lsite);
  pk_load<2, DAG, LSC>(wa, in, shN, lsite, s0, spb); pk_compute<2>(acc, wa, shP, lsite);
  pk_load<3, DAG, LSC>(wa, in, shN, lsite, s0, spb); pk_compute<3>(acc, wa, shP, lsite);
  pk_load<4, DAG, LSC>(wa, in, shN, lsite, s0, spb); pk_compute<4>(acc, wa, shP, lsite);
  pk_load<5, DAG, LSC>(wa, in, shN, lsite, s0, spb); pk_compute<5>(acc, wa, shP, lsite);
  pk_load<6, DAG, LSC>(wa, in, shN, lsite, s0, spb); pk_compute<6>(acc, wa, shP, lsite);
  pk_load<7, DAG, LSC>(wa, in, shN, lsite, s0, spb); pk_compute<7>(acc, wa, shP, lsite);
}

template <bool DAG, typename T, int VEC, typename C>
__device__ __forceinline__ void hop_all(C (*acc)[24], const T* __restrict__ in, const C* shL,
                                        const int* shN, int lsite, int s0, int spb, int Ls) {
#if DWF_HOP_PIPE
  typename Vx<T, VEC>::R ra[24], rb[24];
  hop_load<0, DAG, T, VEC>(ra, in, shN, lsite, s0, spb, Ls);
  hop_load<1, DAG, T, VEC>(rb, in, shN, lsite, s0, spb, Ls);
  hop_compute<0, T, VEC>(acc, ra, shL, lsite, spb);
  hop_load<2, DAG, T, VEC>(ra, in, shN, lsite, s0, spb, Ls);
  hop_compute<1, T, VEC>(acc, rb, shL, lsite, spb);
  hop_load<3, DAG, T, VEC>(rb, in, shN, lsite, s0, spb, Ls);
  hop_compute<2, T, VEC>(acc, ra, shL, lsite, spb);
  hop_load<4, DAG, T, VEC>(ra, in, shN, lsite, s0, spb, Ls);
  hop_compute<3, T, VEC>(acc, rb, shL, lsite, spb);
  hop_load<5, DAG, T, VEC>(rb, in, shN, lsite, s0, spb, Ls);
  hop_compute<4, T, VEC>(acc, ra, shL, lsite, spb);
  hop_load<6, DAG, T, VEC>(ra, in, shN, lsite, s0, spb, Ls);
  hop_compute<5, T, VEC>(acc, rb, shL, lsite, spb);
  hop_load<7, DAG, T, VEC>(rb, in, shN, lsite, s0, spb, Ls);
  hop_compute<6, T, VEC>(acc, ra, shL, lsite, spb);
  hop_compute<7, T, VEC>(acc, rb, shL, lsite, spb);
#else
  hop_dir<0, DAG, T, VEC>(acc, in, shL, shN, lsite, s0, spb, Ls);
  hop_dir<1, DAG, T, VEC>(acc, in, shL, shN, lsite, s0, spb, Ls);
  hop_dir<2, DAG, T, VEC>(acc, in, shL, shN, lsite, s0, spb, Ls);
  hop_dir<3, DAG, T, VEC>(acc, in, shL, shN, lsite, s0, spb, Ls);
  hop_dir<4, DAG, T, VEC>(acc, in, shL, shN, lsite, s0, spb, Ls);
  hop_dir<5, DAG, T, VEC>(acc, in, shL, shN, lsite, s0, spb, Ls);
  hop_dir<6, DAG, T, VEC>(acc, in, shL, shN, lsite, s0, spb, Ls);
  hop_dir<7, DAG, T, VEC>(acc, in, shL, shN, lsite, s0, spb, Ls);
#endif
}

// Workgroups are dispatched round-robin over the 8 XCDs (each with its own
// L2).  Remap the hardware workgroup id so that every XCD walks a contiguous
// range of sites: neighbouring sites then share an L2.
__device__ __forceinline__ int xcd_block(int w, int G) {
  const int q = G >> 3, r = G & 7;
  const int x = w & 7, i = w >> 3;
  return x * q + (x < r ? x : r) + i;
}

// Matrix-core tile of the chiral epilogue: D(16x16) += A(16x4) B(4x16), lane l
// holding A[l % 16][l / 16] and B[l / 16][l % 16].  Register r of lane l holds
// D[4 (l / 16) + r][l % 16] for f32 and D[l / 16 + 4 r][l % 16] for f64 (both
// measured on gfx950).  row(i) names the output s of MFMA row i so that in
// both cases lane l ends up with s = 4 (l / 16) + r: identity for f32, the
// interleave 4 (i % 4) + i / 4 for f64.
template <typename C> struct Mma;
template <> struct Mma<float> {
  typedef float v4 __attribute__((ext_vector_type(4)));
  static __device__ __forceinline__ v4 run(float a, float b, v4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
  static __device__ __forceinline__ int row(int i) { return i; }
};
template <> struct Mma<double> {
  typedef double v4 __attribute__((ext_vector_type(4)));
  static __device__ __forceinline__ v4 run(double a, double b, v4 c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
  static __device__ __forceinline__ int row(int i) { return 4 * (i & 3) + (i >> 2); }
};
// four consecutive elements of a field (s_out = 4q .. 4q + 3 of one row)
template <typename T> struct Q4 {
  static __device__ __forceinline__ void ld(const T* p, float* o) {
    const typename Vx<T, 4>::R r = Vx<T, 4>::ld(p);
#pragma unroll
    for (int i = 0; i < 4; ++i) o[i] = Vx<T, 4>::get(r, i);
  }
  static __device__ __forceinline__ void stn(T* p, const float* v) { Vx<T, 4>::stn(p, v); }
};
template <> struct Q4<double> {
  static __device__ __forceinline__ void ld(const double* p, double* o) {
    const dbl2_t a = Vx<double, 2>::ld(p), b = Vx<double, 2>::ld(p + 2);
    o[0] = a[0]; o[1] = a[1]; o[2] = b[0]; o[3] = b[1];
  }
  static __device__ __forceinline__ void stn(double* p, const double* v) {
    Vx<double, 2>::stn(p, v);
    Vx<double, 2>::stn(p + 2, v + 2);
  }
};

// value of v after a round trip through the storage type
template <typename T>
__device__ __forceinline__ typename St<T>::C rounded(typename St<T>::C v) {
  T t;
  St<T>::st(&t, v);
  return St<T>::ld(&t);
}

// Occupancy target (waves per SIMD) of the fused kernel per storage type; the
// compiler caps registers to reach it.
#ifndef DWF_WAVES_F64
#define DWF_WAVES_F64 4
#endif
#ifndef DWF_WAVES_F32
#define DWF_WAVES_F32 4
#endif
#ifndef DWF_WAVES_F16
#define DWF_WAVES_F16 4
#endif
// 16-bit storage: the chiral epilogue on the FP16/BF16 matrix cores (operands
// rounded to the storage type, FP32 accumulation) instead of the f32 16x16x4 chain
#ifndef DWF_EPI_MFMA16
#define DWF_EPI_MFMA16 1
#endif
// FP16 storage: the packed hop (dot2 SU(3) products) instead of the FP32 one
#ifndef DWF_HOP_PK16
#define DWF_HOP_PK16 1
#endif
#ifndef DWF_HOP_THREADS
#define DWF_HOP_THREADS 256
#endif
#ifndef DWF_VEC_F32
#define DWF_VEC_F32 1
#endif
#ifndef DWF_VEC_F16
#define DWF_VEC_F16 2
#endif
// Per storage type: s-values per thread (VEC, one vector access per spinor
// component), threads per block and the occupancy target.
#ifndef DWF_HOP_THREADS16
#define DWF_HOP_THREADS16 256
#endif
template <typename T> struct KCfg {   // 16-bit storage
  static constexpr int VEC = DWF_VEC_F16;
  static constexpr int THREADS = DWF_HOP_THREADS16;
  static constexpr int WAVES = DWF_WAVES_F16;
};
#ifndef DWF_HOP_THREADS32
#define DWF_HOP_THREADS32 512
#endif
template <> struct KCfg<float> {
  static constexpr int VEC = DWF_VEC_F32;
  static constexpr int THREADS = DWF_HOP_THREADS32;
  static constexpr int WAVES = DWF_WAVES_F32;
};
#ifndef DWF_VEC_F64
#define DWF_VEC_F64 1
#endif
template <> struct KCfg<double> {
  static constexpr int VEC = DWF_VEC_F64;
  static constexpr int THREADS = DWF_HOP_THREADS;
  static constexpr int WAVES = DWF_WAVES_F64;
};

// Occupancy target of one launch: 16-bit site-local launches keep 3 waves/SIMD
// (their two-matrix epilogue needs the registers), hops use the storage type's.
template <typename T, int OP> constexpr int k_waves() {
  return (OP == K_SITE && sizeof(T) == 2) ? 3 : KCfg<T>::WAVES;
}

// Threads per site of a compiled-Ls kernel: Ls / VEC, except at Ls = 12, where
// the last threads of a site idle:
// * 16-bit storage: 6 threads per site padded to 8.  With 6, two of every three
//   4-lane groups of a gather span two sites, whose 24-byte rows then do not
//   coalesce: the hop made 4.6x the L1 accesses of Ls = 16 and ran at half its
//   bytes per second.
// * FP32: 12 padded to 16, so that a block of 32 sites is 8 waves: blocks of 6
//   waves fit only twice in a CU under the 4-waves-per-SIMD register budget
//   (D_eo at 16^3x32: 70.9 -> 57.0 us).
template <typename T> __host__ __device__ constexpr int tps_of(int Ls, int vec) {
  return (Ls == 12 && sizeof(T) == 2 && vec == 2)   ? 8
         : (Ls == 12 && sizeof(T) == 4 && vec == 1) ? 16
                                                    : Ls / vec;
}
// Sites per block of a compiled-Ls kernel: THREADS / tps, at Ls = 12 a multiple
// of 16 (whole waves of sites whose 12 spinor rows per chirality fill 16-column
// matrix-core tiles).  The launchers use the same rule.
template <typename T> __host__ __device__ constexpr int spb_of(int Ls, int vec) {
  return Ls == 12 ? (KCfg<T>::THREADS / tps_of<T>(12, vec)) / 16 * 16
                  : KCfg<T>::THREADS / tps_of<T>(Ls, vec);
}
// the chiral epilogue on the matrix cores (Ls x Ls in a 16 x 16 tile)
constexpr bool mfma_ls(int Ls) { return Ls == 16 || Ls == 12; }

// LSC > 0: Ls and the block shape (spb_of sites of tps_of threads) are
// compile-time constants, so every load and LDS access uses an immediate offset;
// LSC = 0: runtime Ls.  Thread (lsite, q) owns s = q*VEC .. q*VEC + VEC - 1 of
// its site (none when that is beyond Ls: padding threads).
//
// Several right-hand sides on one lattice (nr = 2^nrs of them, each a whole
// field, rstride reals after the previous one) run as one launch: the kernel
// walks "virtual sites" v = site * nr + rhs, so the nr spinors of a 4D site sit
// in the same block and share its staged links (read once for all of them).
// blk / nblk: this block's index among the launch's blocks (after the XCD
// remapping) and their number.
template <typename T, int OP, bool MAT, bool ADDX, bool OUT2, int LSC, int VEC>
__device__ __forceinline__ void
dwf_body(T* __restrict__ out, const T* __restrict__ in, const T* __restrict__ links,
         const int* __restrict__ nbr, int parity, int Vh, int Ls_rt, int spb_rt,
         const typename St<T>::C* __restrict__ E, typename St<T>::C scale,
         const T* __restrict__ X, typename St<T>::C xcoef, T* __restrict__ out2,
         const typename St<T>::C* __restrict__ F, const typename St<T>::C* __restrict__ G,
         const T* X2, double* __restrict__ nrm, const double* xnum, const double* xden,
         T* __restrict__ upd, const double* unum, const double* uden,
         const double* fin_partial, unsigned fin_n, double* fin_result, int blk, int nblk,
         int nrs, size_t rstride) {
  using S = St<T>;
  using C = typename S::C;
  using VL = Vx<T, VEC>;
  const int Ls = LSC > 0 ? LSC : Ls_rt;
  // coefficient of X: a kernel argument, or num / den read from device memory
  const C xc = xnum ? (C)(xnum[0] / xden[0]) : xcoef;
  // CG updates fused into the matrix-core epilogue (Ls = 16), ua = unum / uden:
  // site launch: upd += ua X;  hop launch: upd -= ua out, stored instead of out,
  // and the norm taken of the new upd
  const C ua = upd ? (C)(unum[0] / uden[0]) : C(0);
  const int Q = LSC > 0 ? tps_of<T>(LSC, VEC) : Ls / VEC;   // threads per site
  static_assert(KCfg<T>::THREADS % 64 == 0 && KCfg<T>::THREADS <= 512, "block of 1-8 waves");
  const int spb = LSC > 0 ? spb_of<T>(LSC, VEC) : spb_rt;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  // region A: hop staging (links, neighbours), later the chiral epilogue buffers
  // chiral rows of one site are 12*Ls + 1 apart, so the sites of a wave read
  // their rows from different LDS banks
  const int SU = 12 * Ls + 1;
  // packed FP16 hop (two s per thread, Ls = 16 or 12)
  constexpr bool PK = std::is_same<T, __half>::value && VEC == 2 && mfma_ls(LSC) && DWF_HOP_PK16;
  constexpr int LROW = PK ? kPkStride : 144;   // link words per site in LDS
  const int stageC = (OP == K_SITE) ? 0 : (LROW * spb + 8 * spb);
  const int epiC = MAT ? (mfma_ls(LSC) ? spb * 12 * 17 : 2 * spb * SU) : 0;
  const int regionA = stageC > epiC ? stageC : epiC;
  C* shA = reinterpret_cast<C*>(smem_raw);
  C* shL = shA;
  int* shN = reinterpret_cast<int*>(shA + LROW * spb);
  C* sh_u = shA;
  C* sh_l = shA + spb * SU;
  C* sh_E = shA + regionA;
  C* sh_F = sh_E + 2 * Ls * Ls;
  C* sh_G = sh_F + (OUT2 ? 2 * Ls * Ls : 0);
  // G != 0: the added field X enters through its own chiral matrix (uniform)
  const bool xmat = ADDX && MAT && G != nullptr;

  const int tid = threadIdx.x;
  // a reduction left by the previous launch, finished by this launch's last block
  if (fin_partial && blk == nblk - 1) {
    __shared__ double fsh[8];
    const double t = block_sum(fin_partial, fin_n, fsh);
    if (tid == 0) fin_result[0] = t;
  }
  const int lsite = tid / Q;
  const int s0 = (tid - lsite * Q) * VEC;
  // virtual sites (site * nr + rhs) of this block; the 4D sites they cover
  // are staged: rsite0 .. rsite0 + nsite - 1, this thread's one at row rl
  const int VhV = Vh << nrs;
  const int rmask = (1 << nrs) - 1;
  const int site0 = blk * spb;
  const int site = site0 + lsite;
  const bool valid = (lsite < spb) && (site < VhV) && (s0 < Ls);   // s0 >= Ls: padding thread
  const int rsite0 = site0 >> nrs;
  const int vend = (site0 + spb < VhV ? site0 + spb : VhV) - 1;
  const int nsite = (vend >> nrs) - rsite0 + 1;
  const int rl = (site >> nrs) - rsite0;
  // element offset of virtual site v (rhs field, 4D site)
  auto vaddr = [&](int v) -> size_t {
    return (size_t)(v & rmask) * rstride + (size_t)(v >> nrs) * 24 * Ls;
  };

  if (MAT) {
    const int LL = Ls * Ls;
    for (int i = tid; i < 2 * LL; i += blockDim.x) {
      const int chi = i / LL, r = i - chi * LL;
      const int so = r / Ls, si = r - so * Ls;
      sh_E[chi * LL + si * Ls + so] = E[i];
      if (OUT2) sh_F[chi * LL + si * Ls + so] = F[i];
      if (xmat) sh_G[chi * LL + si * Ls + so] = G[i];
    }
  }
  if (OP != K_SITE) {
    // coalesced cooperative load of the block's 8 x 18 link rows and 8 neighbour rows
    const T* L = links + (size_t)parity * 8 * 18 * Vh + rsite0;
    if constexpr (PK) {
      unsigned* shP = reinterpret_cast<unsigned*>(shA);
      const unsigned short* Lh = reinterpret_cast<const unsigned short*>(L);
      // eight sites per 16-byte load when the rows stay 16-byte aligned (Vh and
      // spb multiples of 8), else one site at a time
      if ((Vh & 7) == 0 && (spb & 7) == 0 && (rsite0 & 7) == 0) {
        const int h8 = spb >> 3;
        for (int i = tid; i < 72 * h8; i += blockDim.x) {
          const int row = i / h8, j = (i - row * h8) * 8;   // row = dir * 9 + entry
          const int dir = row / 9, ab = row - dir * 9;
          const unsigned short* pr = Lh + (size_t)(dir * 18 + 2 * ab) * Vh + j;
          const unsigned short* pi = pr + Vh;
          u32x4_t r4 = {0, 0, 0, 0}, i4 = {0, 0, 0, 0};
          if (j + 8 <= nsite) {
            r4 = __builtin_nontemporal_load(reinterpret_cast<const u32x4_t*>(pr));
            i4 = __builtin_nontemporal_load(reinterpret_cast<const u32x4_t*>(pi));
          } else {
            for (int m = 0; m < 8; ++m)
              if (j + m < nsite) {
                r4[m >> 1] |= (unsigned)__builtin_nontemporal_load(pr + m) << (16 * (m & 1));
                i4[m >> 1] |= (unsigned)__builtin_nontemporal_load(pi + m) << (16 * (m & 1));
              }
          }
          unsigned* d0 = shP + j * kPkStride + dir * 20 + 2 * ab;
#pragma unroll
          for (int m = 0; m < 8; ++m) {
            const unsigned ur = (r4[m >> 1] >> (16 * (m & 1))) & 0xffffu;
            const unsigned ui = (i4[m >> 1] >> (16 * (m & 1))) & 0xffffu;
            d0[m * kPkStride] = ur | ((ui ^ 0x8000u) << 16);   // (ur, -ui)
            d0[m * kPkStride + 1] = ui | (ur << 16);           // (ui, ur)
          }
        }
      } else {
        for (int i = tid; i < 72 * spb; i += blockDim.x) {
          const int row = i / spb, j = i - row * spb;
          const int dir = row / 9, ab = row - dir * 9;
          unsigned ur = 0, ui = 0;
          if (j < nsite) {
            ur = __builtin_nontemporal_load(Lh + (size_t)(dir * 18 + 2 * ab) * Vh + j);
            ui = __builtin_nontemporal_load(Lh + (size_t)(dir * 18 + 2 * ab + 1) * Vh + j);
          }
          shP[j * kPkStride + dir * 20 + 2 * ab] = ur | ((ui ^ 0x8000u) << 16);
          shP[j * kPkStride + dir * 20 + 2 * ab + 1] = ui | (ur << 16);
        }
      }
    } else if constexpr (sizeof(T) >= 4) {
      // 16-byte loads of W consecutive sites when the rows stay aligned
      constexpr int W = 16 / sizeof(T);
      typedef T tw_t __attribute__((ext_vector_type(W)));
      if ((Vh % W) == 0 && (spb % W) == 0 && (rsite0 % W) == 0) {
        const int sw = spb / W;
        for (int i = tid; i < 144 * sw; i += blockDim.x) {
          const int row = i / sw, j = (i - row * sw) * W;
          const T* src = L + (size_t)row * Vh + j;
          C* d = shL + row * spb + j;
          if (j + W <= nsite) {
            const tw_t v = __builtin_nontemporal_load(reinterpret_cast<const tw_t*>(src));
#pragma unroll
            for (int k = 0; k < W; ++k) d[k] = v[k];
          } else {
#pragma unroll
            for (int k = 0; k < W; ++k) d[k] = (j + k < nsite) ? S::ldn(src + k) : C(0);
          }
        }
      } else {
        for (int i = tid; i < 144 * spb; i += blockDim.x) {
          const int row = i / spb, j = i - row * spb;
          shL[i] = (j < nsite) ? S::ldn(L + (size_t)row * Vh + j) : C(0);
        }
      }
    } else {
      for (int i = tid; i < 144 * spb; i += blockDim.x) {
        const int row = i / spb, j = i - row * spb;
        shL[i] = (j < nsite) ? S::ldn(L + (size_t)row * Vh + j) : C(0);
      }
    }
    const int* N = nbr + (size_t)parity * 8 * Vh + rsite0;
    for (int i = tid; i < 8 * spb; i += blockDim.x) {
      const int row = i / spb, j = i - row * spb;
      shN[i] = (j < nsite) ? N[(size_t)row * Vh + j] : 0;
    }
    __syncthreads();
  }

  C acc[VEC][24];
#pragma unroll
  for (int v = 0; v < VEC; ++v)
#pragma unroll
    for (int k = 0; k < 24; ++k) acc[v][k] = 0;
  const size_t base = vaddr(site) + s0;
  const T* inr = in + (size_t)(site & rmask) * rstride;   // this rhs's input field
  if (valid) {
    if (OP == K_SITE) {
#pragma unroll
      for (int k = 0; k < 24; ++k) {
        const typename VL::R r = VL::ld(in + base + k * Ls);
#pragma unroll
        for (int v = 0; v < VEC; ++v) acc[v][k] = VL::get(r, v);
      }
    } else if (OP == K_HOP) {
      if constexpr (PK)
        hop_all_pk<false, LSC>(acc, reinterpret_cast<const __half*>(inr),
                          reinterpret_cast<const unsigned*>(shA), shN, rl, s0, spb);
      else
        hop_all<false, T, VEC>(acc, inr, shL, shN, rl, s0, spb, Ls);
    } else {
      if constexpr (PK)
        hop_all_pk<true, LSC>(acc, reinterpret_cast<const __half*>(inr),
                         reinterpret_cast<const unsigned*>(shA), shN, rl, s0, spb);
      else
        hop_all<true, T, VEC>(acc, inr, shL, shN, rl, s0, spb, Ls);
#pragma unroll
      for (int v = 0; v < VEC; ++v)
#pragma unroll
        for (int k = 0; k < 12; ++k) {
          C t = acc[v][k];
          acc[v][k] = acc[v][k + 12];
          acc[v][k + 12] = t;
        }
    }
  }

  if (!MAT) {
    if (valid) {
#pragma unroll
      for (int k = 0; k < 24; ++k) {
        C o[VEC];
#pragma unroll
        for (int v = 0; v < VEC; ++v) o[v] = scale * acc[v][k];
        if (ADDX) {
          const typename VL::R r = VL::ld(X + base + k * Ls);
#pragma unroll
          for (int v = 0; v < VEC; ++v) o[v] += xc * VL::get(r, v);
        }
        VL::stn(out + base + k * Ls, o);
      }
    }
    return;
  }

  if (OP != K_SITE) __syncthreads();  // staging area is reused below
  if constexpr (mfma_ls(LSC)) {
    // Matrix-core epilogue.  Per chirality the block's core is an Ls x (12 spb)
    // matrix U[s][col], col = lsite*12 + k; O = E U on 16-column tiles.  One
    // chirality at a time goes through one LDS buffer (column stride 17, bank
    // spread), which halves the LDS per block.  Ls = 12: the 12 x 12 matrices
    // in 16 x 16 tiles, A rows and columns beyond 12 zero, three k-steps of the
    // 16x16x4 chains, output rows s >= 12 (lanes 48-63) not stored.
    using MM = Mma<C>;
    using v4 = typename MM::v4;
    constexpr int LP = 17;
    constexpr int SPB = spb_of<T>(LSC, VEC);
    constexpr int NW = SPB * tps_of<T>(LSC, VEC) / 64;   // waves per block
    constexpr int NT = SPB * 12 / 16;              // 16-column tiles per block
    constexpr int NTW = (NT + NW - 1) / NW;        // tiles per wave
    constexpr int LL = LSC * LSC;                  // one chirality's matrix
    C* mb = shA;
    // 16-bit storage: the staged columns are kept in the storage type, which is
    // what the 16-bit matrix cores take (the same rounding as converting each
    // operand at the tile), LPH halves apart so that a lane's 4 rows are one
    // 8-byte read; rows beyond Ls are never staged and read as zero
    constexpr bool S16 = sizeof(T) == 2 && DWF_EPI_MFMA16;
    constexpr int LPH = 20;
    unsigned short* mh = reinterpret_cast<unsigned short*>(shA);
    auto h16 = [](C f) -> unsigned {
      if constexpr (std::is_same<T, __half>::value) {
        return __builtin_bit_cast(unsigned short, (_Float16)f);
      } else {
        const __hip_bfloat16 h = __float2bfloat16(f);
        return *reinterpret_cast<const unsigned short*>(&h);
      }
    };
    const int lane = tid & 63, wave = tid >> 6, nwaves = blockDim.x >> 6;
    const int li = lane & 15, lk = lane >> 4;
    const int ntiles = (spb * 12) / 16;
    const int arow = MM::row(li);   // output s of this lane's A row
    v4 cu[NTW], cl[NTW];
    // D[j] (+)= M_chi B on the staged chirality
    // one 16-column tile: init + M_chi B, B = the tile's columns of the staged
    // chirality (the columns of a tile are read only by the wave that owns it)
    auto mma_tile = [&](const C* M, int col, v4 init) -> v4 {
      if constexpr (std::is_same<T, __half>::value && DWF_EPI_MFMA16) {
        // FP16 storage: one v_mfma_f32_16x16x16f16, operands rounded to FP16
        // (lane l holds A[l % 16][4 (l / 16) + i] and B[4 (l / 16) + i][l % 16]);
        // the output layout is that of the f32 16x16x4 instruction
        typedef _Float16 h4_t __attribute__((ext_vector_type(4)));
        h4_t av;
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          const int sin = 4 * lk + i;
          av[i] = (_Float16)(sin < LSC && arow < LSC ? M[sin * LSC + arow] : C(0));
        }
        const u32x2_t bw = 4 * lk < LSC ? *reinterpret_cast<const u32x2_t*>(mh + col * LPH + 4 * lk)
                                        : u32x2_t{0u, 0u};
        return __builtin_amdgcn_mfma_f32_16x16x16f16(av, __builtin_bit_cast(h4_t, bw), init, 0, 0,
                                                     0);
      } else if constexpr (std::is_same<T, __hip_bfloat16>::value && DWF_EPI_MFMA16) {
        // BF16 storage: the same with v_mfma_f32_16x16x16_bf16 (round to nearest even)
        typedef short s4_t __attribute__((ext_vector_type(4)));
        auto bf = [](float f) -> short {
          const __hip_bfloat16 h = __float2bfloat16(f);
          return *reinterpret_cast<const short*>(&h);
        };
        s4_t av;
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          const int sin = 4 * lk + i;
          av[i] = bf(sin < LSC && arow < LSC ? M[sin * LSC + arow] : C(0));
        }
        const u32x2_t bw = 4 * lk < LSC ? *reinterpret_cast<const u32x2_t*>(mh + col * LPH + 4 * lk)
                                        : u32x2_t{0u, 0u};
        return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(av, __builtin_bit_cast(s4_t, bw), init, 0,
                                                         0, 0);
      } else {
        v4 a = init;
#pragma unroll
        for (int kc = 0; kc < LSC / 4; ++kc) {
          const int sin = kc * 4 + lk;
          a = MM::run(arow < LSC ? M[sin * LSC + arow] : C(0), mb[col * LP + sin], a);
        }
        return a;
      }
    };
    // D[j] (+)= M_chi B on the staged chirality, every tile of this wave
    auto mma_one = [&](const C* M, v4 (&D)[NTW], bool add) {
#pragma unroll
      for (int j = 0; j < NTW; ++j) {
        const int t = wave + j * nwaves;
        if (t < ntiles) D[j] = mma_tile(M, t * 16 + li, add ? D[j] : v4{0, 0, 0, 0});
      }
    };
    // core: u = acc_k + acc_{k+12} (P+), then l = acc_k - acc_{k+12} (P-)
    auto stage_acc = [&](C sign) {
      if (valid) {
#pragma unroll
        for (int k = 0; k < 12; ++k) {
          if constexpr (S16) {   // VEC = 2: s0 and s0 + 1 in one word
            *reinterpret_cast<unsigned*>(mh + (lsite * 12 + k) * LPH + s0) =
                h16(acc[0][k] + sign * acc[0][k + 12]) |
                (h16(acc[1][k] + sign * acc[1][k + 12]) << 16);
          } else {
#pragma unroll
            for (int v = 0; v < VEC; ++v)
              mb[(lsite * 12 + k) * LP + s0 + v] = acc[v][k] + sign * acc[v][k + 12];
          }
        }
      }
    };
    stage_acc(C(1));
    __syncthreads();
    mma_one(sh_E, cu, false);
    __syncthreads();
    stage_acc(C(-1));
    __syncthreads();
    mma_one(sh_E + LL, cl, false);
    if (xmat) {   // + G X, X staged the same way
      auto stage_x = [&](C sign) {
        if (valid) {
#pragma unroll
          for (int k = 0; k < 12; ++k) {
            const typename VL::R ra = VL::ld(X + base + k * Ls);
            const typename VL::R rb = VL::ld(X + base + (k + 12) * Ls);
            if constexpr (S16) {
              *reinterpret_cast<unsigned*>(mh + (lsite * 12 + k) * LPH + s0) =
                  h16(VL::get(ra, 0) + sign * VL::get(rb, 0)) |
                  (h16(VL::get(ra, 1) + sign * VL::get(rb, 1)) << 16);
            } else {
#pragma unroll
              for (int v = 0; v < VEC; ++v)
                mb[(lsite * 12 + k) * LP + s0 + v] = VL::get(ra, v) + sign * VL::get(rb, v);
            }
          }
        }
      };
      __syncthreads();
      stage_x(C(1));
      __syncthreads();
      mma_one(sh_G, cu, true);
      __syncthreads();
      stage_x(C(-1));
      __syncthreads();
      mma_one(sh_G + LL, cl, true);
    }
    // combine, add X, store; lane: column col, rows s_out = 4 lk + r.  With
    // OUT2 the tile's chiral combinations are staged in its own LDS columns and
    // F is applied right away (after the last E product every wave reads only
    // the columns of its own tiles, so no block barrier is needed).
    double nacc = 0;
#pragma unroll
    for (int j = 0; j < NTW; ++j) {
      const int t = wave + j * nwaves;
      if (t < ntiles) {
        const int col = t * 16 + li;
        const int cs = col / 12, ck = col - cs * 12;
        const bool cvalid = site0 + cs < VhV && 4 * lk < LSC;   // rows s >= Ls: padding
        const size_t ob = vaddr(site0 + cs) + (size_t)ck * Ls + 4 * lk;
        C o1[4], o2[4];   // rows k and k + 12
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          o1[r] = cu[j][r] + cl[j][r];
          o2[r] = cu[j][r] - cl[j][r];
        }
        if (cvalid) {
          if (ADDX && !xmat) {
            C xa[4], xb[4];
            Q4<T>::ld(X + ob, xa);
            Q4<T>::ld(X + ob + 12 * Ls, xb);
#pragma unroll
            for (int r = 0; r < 4; ++r) {
              o1[r] += xc * xa[r];
              o2[r] += xc * xb[r];
            }
            if (OP == K_SITE && upd) {
              C ya[4], yb[4];
              Q4<T>::ld(upd + ob, ya);
              Q4<T>::ld(upd + ob + 12 * Ls, yb);
#pragma unroll
              for (int r = 0; r < 4; ++r) {
                ya[r] += ua * xa[r];
                yb[r] += ua * xb[r];
              }
              Q4<T>::stn(upd + ob, ya);
              Q4<T>::stn(upd + ob + 12 * Ls, yb);
            }
          }
          if (OP != K_SITE && upd) {
            C ya[4], yb[4];
            Q4<T>::ld(upd + ob, ya);
            Q4<T>::ld(upd + ob + 12 * Ls, yb);
#pragma unroll
            for (int r = 0; r < 4; ++r) {
              ya[r] = rounded<T>(ya[r] - ua * o1[r]);
              yb[r] = rounded<T>(yb[r] - ua * o2[r]);
            }
            if (nrm) {
#pragma unroll
              for (int r = 0; r < 4; ++r)
                nacc += (double)ya[r] * (double)ya[r] + (double)yb[r] * (double)yb[r];
            }
            Q4<T>::stn(upd + ob, ya);
            Q4<T>::stn(upd + ob + 12 * Ls, yb);
          } else {
            if (out) {   // out == nullptr: only out2 and the norm are wanted
              Q4<T>::stn(out + ob, o1);
              Q4<T>::stn(out + ob + 12 * Ls, o2);
            }
            if (nrm) {   // (uniform: the norm is wanted by few launches)
#pragma unroll
              for (int r = 0; r < 4; ++r)
                nacc += (double)o1[r] * (double)o1[r] + (double)o2[r] * (double)o2[r];
            }
          }
        }
        if (OUT2) {   // out2 = F out, tile by tile in the wave's own LDS columns
          // The FP32/FP64 chain reads rows written by other lanes of the wave,
          // and the second staging overwrites rows other lanes have just read:
          // each hand-over is a wave-scope release/acquire around a wave barrier.
          auto stage_col = [&](C sign) {   // this lane's 4 rows of column col
            if constexpr (S16) {
              *reinterpret_cast<u32x2_t*>(mh + col * LPH + 4 * lk) =
                  u32x2_t{h16(o1[0] + sign * o2[0]) | (h16(o1[1] + sign * o2[1]) << 16),
                          h16(o1[2] + sign * o2[2]) | (h16(o1[3] + sign * o2[3]) << 16)};
            } else {
#pragma unroll
              for (int r = 0; r < 4; ++r) mb[col * LP + 4 * lk + r] = o1[r] + sign * o2[r];
            }
          };
          stage_col(C(1));
          wave_lds_sync();
          const v4 fu = mma_tile(sh_F, col, v4{0, 0, 0, 0});
          wave_lds_sync();
          stage_col(C(-1));
          wave_lds_sync();
          const v4 fl = mma_tile(sh_F + LL, col, v4{0, 0, 0, 0});
          if (cvalid) {
            C p1[4], p2[4];
#pragma unroll
            for (int r = 0; r < 4; ++r) {
              p1[r] = fu[r] + fl[r];
              p2[r] = fu[r] - fl[r];
            }
            if (X2) {
              C xa[4], xb[4];
              Q4<T>::ld(X2 + ob, xa);
              Q4<T>::ld(X2 + ob + 12 * Ls, xb);
#pragma unroll
              for (int r = 0; r < 4; ++r) {
                p1[r] += xa[r];
                p2[r] += xb[r];
              }
            }
            Q4<T>::stn(out2 + ob, p1);
            Q4<T>::stn(out2 + ob + 12 * Ls, p2);
          }
        }
      }
    }
    if (nrm) {   // |out|^2: one partial per block (uniform branch)
      __shared__ double wsum2[8];
      const double t = wave_sum(nacc);
      if (lane == 0) wsum2[wave] = t;
      __syncthreads();
      if (tid == 0) {
        double a = 0;
        for (int w = 0; w < nwaves; ++w) a += wsum2[w];
        nrm[blk] = a;
      }
    }
    return;
  }
  // chiral components of the core, per (site, k < 12, s): P+ ~ u, P- ~ l
  if (valid) {
    C* u = sh_u + lsite * SU + s0;
    C* l = sh_l + lsite * SU + s0;
#pragma unroll
    for (int k = 0; k < 12; ++k)
#pragma unroll
      for (int v = 0; v < VEC; ++v) {
        u[k * Ls + v] = acc[v][k] + acc[v][k + 12];
        l[k * Ls + v] = acc[v][k] - acc[v][k + 12];
      }
  }
  __syncthreads();
  const int LL = Ls * Ls;
  // ou[v][k] = sum_s E+[s0+v][s] u[k][s] (and E- on l); E sits in LDS as [chi][s][s_out]
  C ou[VEC][12], ol[VEC][12];
#pragma unroll
  for (int v = 0; v < VEC; ++v)
#pragma unroll
    for (int k = 0; k < 12; ++k) { ou[v][k] = 0; ol[v][k] = 0; }
  if (valid) {
    const C* u = sh_u + lsite * SU;
    const C* l = sh_l + lsite * SU;
    for (int s = 0; s < Ls; ++s) {
      C ep[VEC], em[VEC];
#pragma unroll
      for (int v = 0; v < VEC; ++v) {
        ep[v] = sh_E[s * Ls + s0 + v];
        em[v] = sh_E[LL + s * Ls + s0 + v];
      }
#pragma unroll
      for (int k = 0; k < 12; ++k) {
        const C us = u[k * Ls + s], ls_ = l[k * Ls + s];
#pragma unroll
        for (int v = 0; v < VEC; ++v) {
          ou[v][k] += ep[v] * us;
          ol[v][k] += em[v] * ls_;
        }
      }
    }
  }
  if (xmat) {
    // + G X: stage the chiral components of X in the same LDS rows
    __syncthreads();
    if (valid) {
      C* u = sh_u + lsite * SU + s0;
      C* l = sh_l + lsite * SU + s0;
#pragma unroll
      for (int k = 0; k < 12; ++k) {
        const typename VL::R ra = VL::ld(X + base + k * Ls);
        const typename VL::R rb = VL::ld(X + base + (k + 12) * Ls);
#pragma unroll
        for (int v = 0; v < VEC; ++v) {
          const C a = VL::get(ra, v), b = VL::get(rb, v);
          u[k * Ls + v] = a + b;
          l[k * Ls + v] = a - b;
        }
      }
    }
    __syncthreads();
    if (valid) {
      const C* u = sh_u + lsite * SU;
      const C* l = sh_l + lsite * SU;
      for (int s = 0; s < Ls; ++s) {
        C gp[VEC], gm[VEC];
#pragma unroll
        for (int v = 0; v < VEC; ++v) {
          gp[v] = sh_G[s * Ls + s0 + v];
          gm[v] = sh_G[LL + s * Ls + s0 + v];
        }
#pragma unroll
        for (int k = 0; k < 12; ++k) {
          const C us = u[k * Ls + s], ls_ = l[k * Ls + s];
#pragma unroll
          for (int v = 0; v < VEC; ++v) {
            ou[v][k] += gp[v] * us;
            ol[v][k] += gm[v] * ls_;
          }
        }
      }
    }
  }
  C o[VEC][24];
  if (valid) {
#pragma unroll
    for (int v = 0; v < VEC; ++v)
#pragma unroll
      for (int k = 0; k < 12; ++k) {
        o[v][k] = ou[v][k] + ol[v][k];
        o[v][k + 12] = ou[v][k] - ol[v][k];
      }
    if (ADDX && !xmat) {
#pragma unroll
      for (int k = 0; k < 24; ++k) {
        const typename VL::R r = VL::ld(X + base + k * Ls);
#pragma unroll
        for (int v = 0; v < VEC; ++v) o[v][k] += xc * VL::get(r, v);
      }
    }
#pragma unroll
    for (int k = 0; k < 24; ++k) {
      C t[VEC];
#pragma unroll
      for (int v = 0; v < VEC; ++v) t[v] = o[v][k];
      VL::stn(out + base + k * Ls, t);
    }
  }
  if (nrm) {   // |out|^2: one partial per block (uniform branch)
    double t = 0;
    if (valid) {
#pragma unroll
      for (int v = 0; v < VEC; ++v)
#pragma unroll
        for (int k = 0; k < 24; ++k) t += (double)o[v][k] * (double)o[v][k];
    }
    __shared__ double wsum[8];
    t = wave_sum(t);
    if ((tid & 63) == 0) wsum[tid >> 6] = t;
    __syncthreads();
    if (tid == 0) {
      double a = 0;
      for (int w = 0; w < (int)((blockDim.x + 63) >> 6); ++w) a += wsum[w];
      nrm[blk] = a;
    }
  }
  if (OUT2) {
    __syncthreads();
    if (valid) {
      C* u = sh_u + lsite * SU + s0;
      C* l = sh_l + lsite * SU + s0;
#pragma unroll
      for (int k = 0; k < 12; ++k)
#pragma unroll
        for (int v = 0; v < VEC; ++v) {
          u[k * Ls + v] = o[v][k] + o[v][k + 12];
          l[k * Ls + v] = o[v][k] - o[v][k + 12];
        }
    }
    __syncthreads();
    if (valid) {
#pragma unroll
      for (int v = 0; v < VEC; ++v)
#pragma unroll
        for (int k = 0; k < 12; ++k) { ou[v][k] = 0; ol[v][k] = 0; }
      const C* u = sh_u + lsite * SU;
      const C* l = sh_l + lsite * SU;
      for (int s = 0; s < Ls; ++s) {
        C fp[VEC], fm[VEC];
#pragma unroll
        for (int v = 0; v < VEC; ++v) {
          fp[v] = sh_F[s * Ls + s0 + v];
          fm[v] = sh_F[LL + s * Ls + s0 + v];
        }
#pragma unroll
        for (int k = 0; k < 12; ++k) {
          const C us = u[k * Ls + s], ls_ = l[k * Ls + s];
#pragma unroll
          for (int v = 0; v < VEC; ++v) {
            ou[v][k] += fp[v] * us;
            ol[v][k] += fm[v] * ls_;
          }
        }
      }
      // X2 != 0: out2 = F o + X2 (X2 may be out2 itself: each element is read
      // and written by this thread only)
      if (X2) {
#pragma unroll
        for (int k = 0; k < 12; ++k) {
          const typename VL::R ra = VL::ld(X2 + base + k * Ls);
          const typename VL::R rb = VL::ld(X2 + base + (k + 12) * Ls);
#pragma unroll
          for (int v = 0; v < VEC; ++v) {
            const C a = VL::get(ra, v), b = VL::get(rb, v);
            ou[v][k] += (C)0.5 * (a + b);
            ol[v][k] += (C)0.5 * (a - b);
          }
        }
      }
#pragma unroll
      for (int k = 0; k < 12; ++k) {
        C t1[VEC], t2[VEC];
#pragma unroll
        for (int v = 0; v < VEC; ++v) {
          t1[v] = ou[v][k] + ol[v][k];
          t2[v] = ou[v][k] - ol[v][k];
        }
        VL::stn(out2 + base + k * Ls, t1);
        VL::stn(out2 + base + (k + 12) * Ls, t2);
      }
    }
  }
}

template <typename T, int OP, bool MAT, bool ADDX, bool OUT2, int LSC, int VEC>
__global__ void __launch_bounds__(KCfg<T>::THREADS, (k_waves<T, OP>()))
k_dwf(T* __restrict__ out, const T* __restrict__ in, const T* __restrict__ links,
      const int* __restrict__ nbr, int parity, int Vh, int Ls_rt, int spb_rt,
      const typename St<T>::C* __restrict__ E, typename St<T>::C scale,
      const T* __restrict__ X, typename St<T>::C xcoef, T* __restrict__ out2,
      const typename St<T>::C* __restrict__ F, const typename St<T>::C* __restrict__ G,
      const T* X2, double* __restrict__ nrm, const double* xnum, const double* xden,
      T* __restrict__ upd, const double* unum, const double* uden,
      const double* fin_partial, unsigned fin_n, double* fin_result, int nrs, size_t rstride) {
  dwf_body<T, OP, MAT, ADDX, OUT2, LSC, VEC>(
      out, in, links, nbr, parity, Vh, Ls_rt, spb_rt, E, scale, X, xcoef, out2, F, G, X2, nrm,
      xnum, xden, upd, unum, uden, fin_partial, fin_n, fin_result,
      xcd_block(blockIdx.x, gridDim.x), gridDim.x, nrs, rstride);
}

// Arguments of one member of a grouped launch (one lattice, or several
// right-hand sides of one lattice), read by every block from device memory.
template <typename T> struct DwfArgs {
  using C = typename St<T>::C;
  T* out; const T* in; const T* links; const int* nbr;
  const C* E; const T* X; T* out2; const C* F; const C* G; const T* X2;
  double* nrm; const double* xnum; const double* xden;
  T* upd; const double* unum; const double* uden;
  const double* fin_partial; double* fin_result;
  size_t rstride;
  C scale, xcoef;
  int parity, Vh, nrs, blk_begin, nblk;
  unsigned fin_n;
};

// Lattices of different volume (same Ls, precision and epilogue form) in one
// launch: block w (after the XCD remapping) belongs to the last member whose
// blk_begin <= w; members are at most kGroupMax.
template <typename T, int OP, bool MAT, bool ADDX, bool OUT2, int LSC, int VEC>
__global__ void __launch_bounds__(KCfg<T>::THREADS, (k_waves<T, OP>()))
k_dwf_group(const DwfArgs<T>* __restrict__ tab, int n) {
  const int w = xcd_block(blockIdx.x, gridDim.x);
  int g = 0;
#pragma unroll
  for (int i = 1; i < kGroupMax; ++i)
    if (i < n && w >= tab[i].blk_begin) g = i;
  const DwfArgs<T>& a = tab[g];
  dwf_body<T, OP, MAT, ADDX, OUT2, LSC, VEC>(
      a.out, a.in, a.links, a.nbr, a.parity, a.Vh, LSC, 0, a.E, a.scale, a.X, a.xcoef, a.out2,
      a.F, a.G, a.X2, a.nrm, a.xnum, a.xden, a.upd, a.unum, a.uden, a.fin_partial, a.fin_n,
      a.fin_result, w - a.blk_begin, a.nblk, a.nrs, a.rstride);
}

template <typename T, int OP, bool MAT, bool ADDX, bool OUT2, int LSC>
static void launch_k(dim3 grid, dim3 block, size_t lds, hipStream_t st, void* out,
                     const void* in, const void* links, const int* nbr, int parity, int Vh,
                     int Ls, int spb, const Epi& e) {
  using C = typename St<T>::C;
  // compile-time Ls (all multiples of 4): vector accesses; runtime Ls: scalar
  constexpr int VEC = LSC > 0 ? KCfg<T>::VEC : 1;
  hipLaunchKernelGGL((k_dwf<T, OP, MAT, ADDX, OUT2, LSC, VEC>), grid, block, lds, st, (T*)out,
                     (const T*)in, (const T*)links, nbr, parity, Vh, Ls, spb, (const C*)e.E,
                     (C)e.scale, (const T*)e.X, (C)e.xcoef, (T*)e.out2, (const C*)e.F,
                     (const C*)e.G, (const T*)e.X2, e.nrm_partial, e.xnum, e.xden, (T*)e.upd,
                     e.unum, e.uden, e.fin_partial, e.fin_n, e.fin_result, e.nrs, e.rstride);
}

// the (op, epilogue) combinations the operator chains use
template <typename T, int LSC>
static hipError_t launch_combo(int op, bool mat, bool addx, bool out2, dim3 g, dim3 b,
                               size_t lds, hipStream_t st, void* out, const void* in,
                               const void* links, const int* nbr, int parity, int Vh, int Ls,
                               int spb, const Epi& e) {
#define LK(OPV, M, A, O2) \
  launch_k<T, OPV, M, A, O2, LSC>(g, b, lds, st, out, in, links, nbr, parity, Vh, Ls, spb, e)
  if (op == K_SITE) {
    if (!mat) return hipErrorInvalidValue;
    if (addx) {
      if (!out2) return hipErrorInvalidValue;
      LK(K_SITE, true, true, true);
    } else if (out2) {
      LK(K_SITE, true, false, true);
    } else {
      LK(K_SITE, true, false, false);
    }
  } else if (op == K_HOP) {
    if (!mat) {
      if (addx || out2) return hipErrorInvalidValue;
      LK(K_HOP, false, false, false);
    } else if (!addx) {
      if (out2) LK(K_HOP, true, false, true);
      else LK(K_HOP, true, false, false);
    } else if (out2) {
      LK(K_HOP, true, true, true);
    } else {
      LK(K_HOP, true, true, false);
    }
  } else {
    if (!mat) return hipErrorInvalidValue;
    if (addx && out2) LK(K_HOPDAG, true, true, true);
    else if (addx) LK(K_HOPDAG, true, true, false);
    else if (out2) return hipErrorInvalidValue;
    else LK(K_HOPDAG, true, false, false);
  }
#undef LK
  return hipGetLastError();
}

static void finish_norm(const double* partial, unsigned nblocks, double* result, hipStream_t st,
                        double* host);

template <typename T>
static hipError_t launch_dwf_t(int op, void* out, const void* in, const void* links,
                               const int* nbr, int parity, int Vh, int Ls, const Epi& e,
                               hipStream_t st, unsigned* nblocks) {
  using C = typename St<T>::C;
  const bool mat = e.E != nullptr;
  const bool addx = e.X != nullptr;
  const bool out2 = e.out2 != nullptr;
  if (out2 && !mat) return hipErrorInvalidValue;
  if (e.nrm_partial && !mat) return hipErrorInvalidValue;
  const int vec = KCfg<T>::VEC;
  const bool lsc = (Ls == 4 || Ls == 8 || Ls == 12 || Ls == 16) && Ls % vec == 0;
  int spb = lsc ? spb_of<T>(Ls, vec) : KCfg<T>::THREADS / Ls;
  if (spb < 1) return hipErrorInvalidValue;
  const size_t cs = sizeof(C);
  if ((e.G && (!mat || !addx)) || (e.X2 && !out2)) return hipErrorInvalidValue;
  const size_t matb =
      mat ? 2 * (size_t)Ls * Ls * cs * (1 + (out2 ? 1 : 0) + (e.G ? 1 : 0)) : 0;
  const size_t lds_cap = 64 * 1024;
  auto lds_for = [&](int sp) {
    const bool pk = std::is_same<T, __half>::value && KCfg<T>::VEC == 2 && mfma_ls(Ls) &&
                    DWF_HOP_PK16;
    const size_t stage = (op == K_SITE) ? 0 : (size_t)((pk ? kPkStride : 144) + 8) * sp * cs;
    const size_t epi = mat ? (size_t)sp * (mfma_ls(Ls) && lsc ? 12 * 17 : 2 * (12 * Ls + 1)) * cs : 0;
    return (stage > epi ? stage : epi) + matb;
  };
  // The compiled-Ls kernels fix their block at THREADS / (Ls / VEC) sites; when
  // that block's link staging exceeds the LDS cap (FP64 and FP32 at Ls = 4) the
  // runtime-Ls kernel below runs instead, with fewer sites per block.
  // the fused CG updates, and an output that is not stored, exist in the
  // matrix-core epilogue only
  const bool mfma_path = mat && mfma_ls(Ls) && lsc && lds_for(spb) <= lds_cap;
  if (e.upd && !(mfma_path && (op == K_SITE ? addx : !out2))) return hipErrorInvalidValue;
  if (!out && !e.upd && !(mfma_path && out2)) return hipErrorInvalidValue;
  const int VhV = Vh << e.nrs;   // virtual sites (site, rhs)
  if (lsc && lds_for(spb) <= lds_cap) {
    const size_t lds = lds_for(spb);
    const dim3 block(spb * tps_of<T>(Ls, vec)), grid((VhV + spb - 1) / spb);
    if (e.fin_partial && block.x % 64) return hipErrorInvalidValue;   // block_sum: whole waves
    *nblocks = grid.x;
    switch (Ls) {
      case 4: return launch_combo<T, 4>(op, mat, addx, out2, grid, block, lds, st, out, in, links, nbr, parity, Vh, Ls, spb, e);
      case 8: return launch_combo<T, 8>(op, mat, addx, out2, grid, block, lds, st, out, in, links, nbr, parity, Vh, Ls, spb, e);
      case 12: return launch_combo<T, 12>(op, mat, addx, out2, grid, block, lds, st, out, in, links, nbr, parity, Vh, Ls, spb, e);
      case 16: return launch_combo<T, 16>(op, mat, addx, out2, grid, block, lds, st, out, in, links, nbr, parity, Vh, Ls, spb, e);
      default: break;
    }
  }
  spb = KCfg<T>::THREADS / Ls;
  while (spb > 1 && lds_for(spb) > lds_cap) --spb;
  if (lds_for(spb) > lds_cap) return hipErrorInvalidValue;
  const size_t lds = lds_for(spb);
  const dim3 block(spb * Ls), grid((VhV + spb - 1) / spb);
  if (e.fin_partial && block.x % 64) return hipErrorInvalidValue;
  *nblocks = grid.x;
  return launch_combo<T, 0>(op, mat, addx, out2, grid, block, lds, st, out, in, links, nbr,
                            parity, Vh, Ls, spb, e);
}

hipError_t launch_dwf(int prec, int op, void* out, const void* in, const void* links,
                      const int* nbr, int parity, int Vh, int Ls, const Epi& e,
                      hipStream_t st) {
  hipError_t err = hipErrorInvalidValue;
  unsigned nb = 0;
  switch (prec) {
    case P64: err = launch_dwf_t<double>(op, out, in, links, nbr, parity, Vh, Ls, e, st, &nb); break;
    case P32: err = launch_dwf_t<float>(op, out, in, links, nbr, parity, Vh, Ls, e, st, &nb); break;
    case P16: err = launch_dwf_t<__half>(op, out, in, links, nbr, parity, Vh, Ls, e, st, &nb); break;
    case PBF16:
      err = launch_dwf_t<__hip_bfloat16>(op, out, in, links, nbr, parity, Vh, Ls, e, st, &nb);
      break;
  }
  if (e.nblocks_out) *e.nblocks_out = nb;
  if (err == hipSuccess && e.nrm_partial && !e.nrm_defer) {
    finish_norm(e.nrm_partial, nb, e.nrm_result, st, e.nrm_host);
    err = hipGetLastError();
  }
  return err;
}

// ---------------------------------------------------------------------------
// grouped launches (several lattices in one launch)
// ---------------------------------------------------------------------------
template <typename T, int OP, bool MAT, bool ADDX, bool OUT2, int LSC>
static void launch_g(const GroupLaunch& gl, hipStream_t st) {
  hipLaunchKernelGGL((k_dwf_group<T, OP, MAT, ADDX, OUT2, LSC, KCfg<T>::VEC>), dim3(gl.grid),
                     dim3(gl.block), gl.lds, st, (const DwfArgs<T>*)gl.table, gl.n);
}
// the combinations the D^+D chain uses
template <typename T, int LSC>
static hipError_t launch_group_t(const GroupLaunch& gl, hipStream_t st) {
  if (!gl.mat) return hipErrorInvalidValue;
  if (gl.op == K_SITE && !gl.addx && !gl.out2) launch_g<T, K_SITE, true, false, false, LSC>(gl, st);
  else if (gl.op == K_HOP && !gl.addx && !gl.out2) launch_g<T, K_HOP, true, false, false, LSC>(gl, st);
  else if (gl.op == K_HOP && gl.addx && gl.out2) launch_g<T, K_HOP, true, true, true, LSC>(gl, st);
  else if (gl.op == K_HOPDAG && !gl.addx && !gl.out2) launch_g<T, K_HOPDAG, true, false, false, LSC>(gl, st);
  else if (gl.op == K_HOPDAG && gl.addx && !gl.out2) launch_g<T, K_HOPDAG, true, true, false, LSC>(gl, st);
  else return hipErrorInvalidValue;
  return hipGetLastError();
}

template <typename T>
static hipError_t prepare_group_t(int op, int Ls, int n, const GroupMember* m, GroupLaunch* gl) {
  using C = typename St<T>::C;
  if (KCfg<T>::VEC != 1 && Ls % KCfg<T>::VEC) return hipErrorInvalidValue;
  const int spb = spb_of<T>(Ls, KCfg<T>::VEC);
  const Epi& e0 = m[0].e;
  const bool mat = e0.E != nullptr, addx = e0.X != nullptr, out2 = e0.out2 != nullptr;
  std::vector<DwfArgs<T>> tab(n);
  int begin = 0;
  for (int i = 0; i < n; ++i) {
    const Epi& e = m[i].e;
    if ((e.E != nullptr) != mat || (e.X != nullptr) != addx || (e.out2 != nullptr) != out2 ||
        (e.G != nullptr) != (e0.G != nullptr) || (e.X2 != nullptr) != (e0.X2 != nullptr) ||
        e.upd || e.xnum || e.nrm_partial || e.fin_partial)
      return hipErrorInvalidValue;   // one epilogue form per launch; no fused CG updates
    DwfArgs<T>& a = tab[i];
    a.out = (T*)m[i].out;
    a.in = (const T*)m[i].in;
    a.links = (const T*)m[i].links;
    a.nbr = m[i].nbr;
    a.E = (const C*)e.E;
    a.X = (const T*)e.X;
    a.out2 = (T*)e.out2;
    a.F = (const C*)e.F;
    a.G = (const C*)e.G;
    a.X2 = (const T*)e.X2;
    a.nrm = nullptr;
    a.xnum = a.xden = nullptr;
    a.upd = nullptr;
    a.unum = a.uden = nullptr;
    a.fin_partial = nullptr;
    a.fin_result = nullptr;
    a.fin_n = 0;
    a.rstride = e.rstride;
    a.scale = (C)e.scale;
    a.xcoef = (C)e.xcoef;
    a.parity = m[i].parity;
    a.Vh = m[i].Vh;
    a.nrs = e.nrs;
    a.nblk = ((m[i].Vh << e.nrs) + spb - 1) / spb;
    a.blk_begin = begin;
    begin += a.nblk;
  }
  const size_t cs = sizeof(C);
  const size_t matb = 2 * (size_t)Ls * Ls * cs * (1 + (out2 ? 1 : 0) + (e0.G ? 1 : 0));
  const size_t stage = (op == K_SITE) ? 0 : (size_t)(144 + 8) * spb * cs;
  const size_t epi = (size_t)spb * (mfma_ls(Ls) ? 12 * 17 : 2 * (12 * Ls + 1)) * cs;
  gl->lds = (stage > epi ? stage : epi) + matb;
  if (gl->lds > 64 * 1024) return hipErrorInvalidValue;
  gl->op = op;
  gl->Ls = Ls;
  gl->n = n;
  gl->mat = mat;
  gl->addx = addx;
  gl->out2 = out2;
  gl->grid = begin;
  gl->block = spb * tps_of<T>(Ls, KCfg<T>::VEC);
  hipError_t err = hipMalloc(&gl->table, sizeof(DwfArgs<T>) * n);
  if (err != hipSuccess) return err;
  return hipMemcpy(gl->table, tab.data(), sizeof(DwfArgs<T>) * n, hipMemcpyHostToDevice);
}

hipError_t prepare_dwf_group(int prec, int op, int Ls, int n, const GroupMember* m,
                             GroupLaunch* gl) {
  if (n < 1 || n > kGroupMax || (Ls != 12 && Ls != 16)) return hipErrorInvalidValue;
  gl->prec = prec;
  switch (prec) {
    case P64: return prepare_group_t<double>(op, Ls, n, m, gl);
    case P32: return prepare_group_t<float>(op, Ls, n, m, gl);
    default: return hipErrorInvalidValue;
  }
}

hipError_t launch_dwf_group(const GroupLaunch& gl, hipStream_t st) {
  if (gl.prec == P64) return gl.Ls == 12 ? launch_group_t<double, 12>(gl, st)
                                         : launch_group_t<double, 16>(gl, st);
  if (gl.prec == P32) return gl.Ls == 12 ? launch_group_t<float, 12>(gl, st)
                                         : launch_group_t<float, 16>(gl, st);
  return hipErrorInvalidValue;
}

void free_dwf_group(GroupLaunch* gl) {
  if (gl->table) (void)hipFree(gl->table);
  gl->table = nullptr;
}

// ---------------------------------------------------------------------------
// layout conversion: corelib Field(24, Vh, Ls) <-> device (site, k, s)
// ---------------------------------------------------------------------------
template <typename T>
__global__ void k_to_dev(T* __restrict__ out, const double* __restrict__ in, int Vh, int Ls) {
  const size_t n = (size_t)Vh * 24 * Ls;
  for (size_t idx = blockIdx.x * (size_t)blockDim.x + threadIdx.x; idx < n;
       idx += (size_t)gridDim.x * blockDim.x) {
    const int s = (int)(idx % Ls);
    const size_t r = idx / Ls;
    const int k = (int)(r % 24);
    const size_t site = r / 24;
    St<T>::st(out + idx, (typename St<T>::C)in[k + 24 * (site + (size_t)Vh * s)]);
  }
}
template <typename T>
__global__ void k_to_corelib(double* __restrict__ out, const T* __restrict__ in, int Vh, int Ls) {
  const size_t n = (size_t)Vh * 24 * Ls;
  for (size_t idx = blockIdx.x * (size_t)blockDim.x + threadIdx.x; idx < n;
       idx += (size_t)gridDim.x * blockDim.x) {
    const int k = (int)(idx % 24);
    const size_t r = idx / 24;
    const size_t site = r % Vh;
    const int s = (int)(r / Vh);
    out[idx] = (double)St<T>::ld(in + (site * 24 + k) * Ls + s);
  }
}

static dim3 grid_for(size_t n, int block = 256, size_t maxg = 16384) {
  size_t g = (n + block - 1) / block;
  if (g > maxg) g = maxg;
  if (g < 1) g = 1;
  return dim3((unsigned)g);
}

#define DWF_DISPATCH(prec, F, ...)                          \
  switch (prec) {                                          \
    case P64: F<double> __VA_ARGS__; break;                \
    case P32: F<float> __VA_ARGS__; break;                 \
    case P16: F<__half> __VA_ARGS__; break;                \
    case PBF16: F<__hip_bfloat16> __VA_ARGS__; break;      \
    default: return hipErrorInvalidValue;                  \
  }

template <typename T>
static void to_dev_t(void* out, const double* in, int Vh, int Ls, hipStream_t st) {
  size_t n = (size_t)Vh * 24 * Ls;
  hipLaunchKernelGGL(k_to_dev<T>, grid_for(n), dim3(256), 0, st, (T*)out, in, Vh, Ls);
}
template <typename T>
static void to_core_t(double* out, const void* in, int Vh, int Ls, hipStream_t st) {
  size_t n = (size_t)Vh * 24 * Ls;
  hipLaunchKernelGGL(k_to_corelib<T>, grid_for(n), dim3(256), 0, st, out, (const T*)in, Vh, Ls);
}

hipError_t launch_to_device_layout(int prec, void* out, const double* in, int Vh, int Ls,
                                   hipStream_t st) {
  DWF_DISPATCH(prec, to_dev_t, (out, in, Vh, Ls, st));
  return hipGetLastError();
}
hipError_t launch_to_corelib_layout(int prec, double* out, const void* in, int Vh, int Ls,
                                    hipStream_t st) {
  DWF_DISPATCH(prec, to_core_t, (out, in, Vh, Ls, st));
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// link packing: lexical Field_G -> per-parity, per-direction, doubly stored
// ---------------------------------------------------------------------------
template <typename T>
__global__ void k_pack_links(T* __restrict__ links, const double* __restrict__ U, int Nx,
                             int Ny, int Nz, int Nt, int bcx, int bcy, int bcz, int bct) {
  const int Nx2 = Nx / 2;
  const int Vh = Nx2 * Ny * Nz * Nt;
  const size_t V = 2 * (size_t)Vh;
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= 2 * Vh) return;
  const int p = idx / Vh;
  const int s2 = idx - p * Vh;
  int r = s2 / Nx2;
  const int x2 = s2 - r * Nx2;
  const int y = r % Ny;
  r /= Ny;
  const int z = r % Nz;
  const int t = r / Nz;
  const int x = 2 * x2 + ((y + z + t + p) & 1);
  const int N[4] = {Nx, Ny, Nz, Nt};
  const int bc[4] = {bcx, bcy, bcz, bct};
  const int xc[4] = {x, y, z, t};
  for (int mu = 0; mu < 4; ++mu) {
    // forward: U_mu(x), boundary factor if x + mu wraps
    {
      const size_t lex = x + (size_t)Nx * (y + (size_t)Ny * (z + (size_t)Nz * t));
      const double f = (xc[mu] == N[mu] - 1) ? (double)bc[mu] : 1.0;
      const double* u = U + 18 * (lex + V * mu);
      T* dst = links + ((size_t)(p * 8 + 2 * mu) * 18) * Vh + s2;
      for (int e = 0; e < 18; ++e) St<T>::st(dst + (size_t)e * Vh, (typename St<T>::C)(f * u[e]));
    }
    // backward: U_mu(x - mu)^dagger, boundary factor if x - mu wraps
    {
      int xb[4] = {x, y, z, t};
      xb[mu] = (xb[mu] - 1 + N[mu]) % N[mu];
      const size_t lex = xb[0] + (size_t)Nx * (xb[1] + (size_t)Ny * (xb[2] + (size_t)Nz * xb[3]));
      const double f = (xc[mu] == 0) ? (double)bc[mu] : 1.0;
      const double* u = U + 18 * (lex + V * mu);
      T* dst = links + ((size_t)(p * 8 + 2 * mu + 1) * 18) * Vh + s2;
      for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b) {
          // (U^+)[a][b] = conj(U[b][a])
          const double re = u[2 * (a + 3 * b)], im = -u[2 * (a + 3 * b) + 1];
          St<T>::st(dst + (size_t)(2 * (b + 3 * a)) * Vh, (typename St<T>::C)(f * re));
          St<T>::st(dst + (size_t)(2 * (b + 3 * a) + 1) * Vh, (typename St<T>::C)(f * im));
        }
    }
  }
}

template <typename T>
static void pack_t(void* links, const double* U, int Nx, int Ny, int Nz, int Nt, int bcx,
                   int bcy, int bcz, int bct, hipStream_t st) {
  const int Vh = Nx / 2 * Ny * Nz * Nt;
  hipLaunchKernelGGL(k_pack_links<T>, dim3((2 * Vh + 255) / 256), dim3(256), 0, st, (T*)links,
                     U, Nx, Ny, Nz, Nt, bcx, bcy, bcz, bct);
}

hipError_t launch_pack_links(int prec, void* links, const double* U, int Nx, int Ny, int Nz,
                             int Nt, int bcx, int bcy, int bcz, int bct, hipStream_t st) {
  DWF_DISPATCH(prec, pack_t, (links, U, Nx, Ny, Nz, Nt, bcx, bcy, bcz, bct, st));
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// BLAS-1.  Fields are processed in 16-byte packs (2 doubles, 4 floats or 8
// 16-bit values per load) so that every precision streams at full width.
// ---------------------------------------------------------------------------
template <typename T> struct Vec;
template <> struct Vec<double> {
  static constexpr int N = 2;
  static __device__ __forceinline__ void ld(const double* p, double* o) {
    const double2 v = *reinterpret_cast<const double2*>(p);
    o[0] = v.x; o[1] = v.y;
  }
  static __device__ __forceinline__ void st(double* p, const double* o) {
    *reinterpret_cast<double2*>(p) = make_double2(o[0], o[1]);
  }
};
template <> struct Vec<float> {
  static constexpr int N = 4;
  static __device__ __forceinline__ void ld(const float* p, float* o) {
    const float4 v = *reinterpret_cast<const float4*>(p);
    o[0] = v.x; o[1] = v.y; o[2] = v.z; o[3] = v.w;
  }
  static __device__ __forceinline__ void st(float* p, const float* o) {
    *reinterpret_cast<float4*>(p) = make_float4(o[0], o[1], o[2], o[3]);
  }
};
template <> struct Vec<__half> {
  static constexpr int N = 8;
  static __device__ __forceinline__ void ld(const __half* p, float* o) {
    const uint4 v = *reinterpret_cast<const uint4*>(p);
    const unsigned w[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      __half2 h = *reinterpret_cast<const __half2*>(&w[i]);
      float2 f = __half22float2(h);
      o[2 * i] = f.x; o[2 * i + 1] = f.y;
    }
  }
  static __device__ __forceinline__ void st(__half* p, const float* o) {
    unsigned w[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      __half2 h = __floats2half2_rn(o[2 * i], o[2 * i + 1]);
      w[i] = *reinterpret_cast<const unsigned*>(&h);
    }
    *reinterpret_cast<uint4*>(p) = make_uint4(w[0], w[1], w[2], w[3]);
  }
};
template <> struct Vec<__hip_bfloat16> {
  static constexpr int N = 8;
  static __device__ __forceinline__ void ld(const __hip_bfloat16* p, float* o) {
    const uint4 v = *reinterpret_cast<const uint4*>(p);
    const unsigned w[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      o[2 * i] = __uint_as_float(w[i] << 16);
      o[2 * i + 1] = __uint_as_float(w[i] & 0xffff0000u);
    }
  }
  static __device__ __forceinline__ void st(__hip_bfloat16* p, const float* o) {
    unsigned w[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      __hip_bfloat16 a = __float2bfloat16(o[2 * i]), b = __float2bfloat16(o[2 * i + 1]);
      w[i] = (unsigned)(*reinterpret_cast<const unsigned short*>(&a)) |
             ((unsigned)(*reinterpret_cast<const unsigned short*>(&b)) << 16);
    }
    *reinterpret_cast<uint4*>(p) = make_uint4(w[0], w[1], w[2], w[3]);
  }
};

// grid-stride over packs; the scalar tail (n % N) is done by thread 0 of block 0
#define PACK_LOOP(T, n)                                                            \
  constexpr int NV = Vec<T>::N;                                                    \
  const size_t npk = (n) / NV;                                                     \
  for (size_t ip = blockIdx.x * (size_t)blockDim.x + threadIdx.x; ip < npk;        \
       ip += (size_t)gridDim.x * blockDim.x)
#define TAIL_LOOP(n)                                                               \
  if (blockIdx.x == 0 && threadIdx.x == 0)                                         \
    for (size_t i = npk * NV; i < (n); ++i)

template <typename T>
__global__ void __launch_bounds__(256) k_set_zero(T* __restrict__ x, size_t n) {
  using C = typename St<T>::C;
  C z[Vec<T>::N] = {};
  PACK_LOOP(T, n) Vec<T>::st(x + ip * NV, z);
  TAIL_LOOP(n) St<T>::st(x + i, 0);
}
template <typename T>
__global__ void __launch_bounds__(256) k_set_value(T* __restrict__ x, typename St<T>::C a,
                                                   size_t n) {
  using C = typename St<T>::C;
  C z[Vec<T>::N];
#pragma unroll
  for (int j = 0; j < Vec<T>::N; ++j) z[j] = a;
  PACK_LOOP(T, n) Vec<T>::st(x + ip * NV, z);
  TAIL_LOOP(n) St<T>::st(x + i, a);
}
template <typename T>
__global__ void __launch_bounds__(256) k_copy(T* __restrict__ y, const T* __restrict__ x, size_t n) {
  const size_t nb = n * sizeof(T) / 16;
  const uint4* xs = reinterpret_cast<const uint4*>(x);
  uint4* ys = reinterpret_cast<uint4*>(y);
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < nb;
       i += (size_t)gridDim.x * blockDim.x)
    ys[i] = xs[i];
  if (blockIdx.x == 0 && threadIdx.x == 0)
    for (size_t i = nb * 16 / sizeof(T); i < n; ++i) y[i] = x[i];
}
template <typename T>
__global__ void __launch_bounds__(256)
k_axpy(T* __restrict__ y, typename St<T>::C a, const T* __restrict__ x, size_t n) {
  using C = typename St<T>::C;
  PACK_LOOP(T, n) {
    C xv[NV], yv[NV];
    Vec<T>::ld(x + ip * NV, xv);
    Vec<T>::ld(y + ip * NV, yv);
#pragma unroll
    for (int j = 0; j < NV; ++j) yv[j] += a * xv[j];
    Vec<T>::st(y + ip * NV, yv);
  }
  TAIL_LOOP(n) St<T>::st(y + i, St<T>::ld(y + i) + a * St<T>::ld(x + i));
}
template <typename T>
__global__ void __launch_bounds__(256)
k_aypx(typename St<T>::C a, T* __restrict__ y, const T* __restrict__ x, size_t n) {
  using C = typename St<T>::C;
  PACK_LOOP(T, n) {
    C xv[NV], yv[NV];
    Vec<T>::ld(x + ip * NV, xv);
    Vec<T>::ld(y + ip * NV, yv);
#pragma unroll
    for (int j = 0; j < NV; ++j) yv[j] = a * yv[j] + xv[j];
    Vec<T>::st(y + ip * NV, yv);
  }
  TAIL_LOOP(n) St<T>::st(y + i, a * St<T>::ld(y + i) + St<T>::ld(x + i));
}
template <typename T>
__global__ void __launch_bounds__(256) k_scal(T* __restrict__ x, typename St<T>::C a, size_t n) {
  using C = typename St<T>::C;
  PACK_LOOP(T, n) {
    C xv[NV];
    Vec<T>::ld(x + ip * NV, xv);
#pragma unroll
    for (int j = 0; j < NV; ++j) xv[j] *= a;
    Vec<T>::st(x + ip * NV, xv);
  }
  TAIL_LOOP(n) St<T>::st(x + i, a * St<T>::ld(x + i));
}
template <typename TO, typename TI>
__global__ void __launch_bounds__(256) k_convert(TO* __restrict__ y, const TI* __restrict__ x, size_t n) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x)
    St<TO>::st(y + i, (typename St<TO>::C)St<TI>::ld(x + i));
}

template <typename T>
static void setz_t(void* x, size_t n, hipStream_t st) {
  hipLaunchKernelGGL(k_set_zero<T>, grid_for(n), dim3(256), 0, st, (T*)x, n);
}
template <typename T>
static void setv_t(void* x, double a, size_t n, hipStream_t st) {
  hipLaunchKernelGGL(k_set_value<T>, grid_for(n), dim3(256), 0, st, (T*)x,
                     (typename St<T>::C)a, n);
}
template <typename T>
static void copy_t(void* y, const void* x, size_t n, hipStream_t st) {
  hipLaunchKernelGGL(k_copy<T>, grid_for(n), dim3(256), 0, st, (T*)y, (const T*)x, n);
}
template <typename T>
static void axpy_t(void* y, double a, const void* x, size_t n, hipStream_t st) {
  hipLaunchKernelGGL(k_axpy<T>, grid_for(n), dim3(256), 0, st, (T*)y,
                     (typename St<T>::C)a, (const T*)x, n);
}
template <typename T>
static void aypx_t(double a, void* y, const void* x, size_t n, hipStream_t st) {
  hipLaunchKernelGGL(k_aypx<T>, grid_for(n), dim3(256), 0, st, (typename St<T>::C)a, (T*)y,
                     (const T*)x, n);
}
template <typename T>
static void scal_t(void* x, double a, size_t n, hipStream_t st) {
  hipLaunchKernelGGL(k_scal<T>, grid_for(n), dim3(256), 0, st, (T*)x, (typename St<T>::C)a, n);
}

hipError_t launch_set_zero(int prec, void* x, size_t n, hipStream_t st) {
  DWF_DISPATCH(prec, setz_t, (x, n, st));
  return hipGetLastError();
}
hipError_t launch_set_value(int prec, void* x, double a, size_t n, hipStream_t st) {
  DWF_DISPATCH(prec, setv_t, (x, a, n, st));
  return hipGetLastError();
}
hipError_t launch_copy(int prec, void* y, const void* x, size_t n, hipStream_t st) {
  DWF_DISPATCH(prec, copy_t, (y, x, n, st));
  return hipGetLastError();
}
hipError_t launch_axpy(int prec, void* y, double a, const void* x, size_t n, hipStream_t st) {
  DWF_DISPATCH(prec, axpy_t, (y, a, x, n, st));
  return hipGetLastError();
}
hipError_t launch_aypx(int prec, double a, void* y, const void* x, size_t n, hipStream_t st) {
  DWF_DISPATCH(prec, aypx_t, (a, y, x, n, st));
  return hipGetLastError();
}
hipError_t launch_scal(int prec, void* x, double a, size_t n, hipStream_t st) {
  DWF_DISPATCH(prec, scal_t, (x, a, n, st));
  return hipGetLastError();
}

template <typename TO>
static hipError_t convert_from(void* y, int prec_in, const void* x, size_t n, hipStream_t st) {
  dim3 g = grid_for(n), b(256);
  switch (prec_in) {
    case P64: hipLaunchKernelGGL((k_convert<TO, double>), g, b, 0, st, (TO*)y, (const double*)x, n); break;
    case P32: hipLaunchKernelGGL((k_convert<TO, float>), g, b, 0, st, (TO*)y, (const float*)x, n); break;
    case P16: hipLaunchKernelGGL((k_convert<TO, __half>), g, b, 0, st, (TO*)y, (const __half*)x, n); break;
    case PBF16: hipLaunchKernelGGL((k_convert<TO, __hip_bfloat16>), g, b, 0, st, (TO*)y, (const __hip_bfloat16*)x, n); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}
// y = a x across storage types (the product in double, one rounding to y's type)
template <typename TO, typename TI>
__global__ void __launch_bounds__(256) k_convert_scaled(TO* __restrict__ y, const TI* __restrict__ x,
                                                        double a, size_t n) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x)
    St<TO>::st(y + i, (typename St<TO>::C)(a * (double)St<TI>::ld(x + i)));
}
// y += a x with x in another storage type (double y)
template <typename TI>
__global__ void __launch_bounds__(256) k_axpy_from(double* __restrict__ y, double a,
                                                   const TI* __restrict__ x, size_t n) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x)
    y[i] += a * (double)St<TI>::ld(x + i);
}
hipError_t launch_convert_scaled(int prec_out, void* y, int prec_in, const void* x, double a,
                                 size_t n, hipStream_t st) {
  if (prec_in != P64) return hipErrorInvalidValue;
  const dim3 g = grid_for(n), b(256);
  const double* xd = (const double*)x;
  switch (prec_out) {
    case P64: hipLaunchKernelGGL((k_convert_scaled<double, double>), g, b, 0, st, (double*)y, xd, a, n); break;
    case P32: hipLaunchKernelGGL((k_convert_scaled<float, double>), g, b, 0, st, (float*)y, xd, a, n); break;
    case P16: hipLaunchKernelGGL((k_convert_scaled<__half, double>), g, b, 0, st, (__half*)y, xd, a, n); break;
    case PBF16: hipLaunchKernelGGL((k_convert_scaled<__hip_bfloat16, double>), g, b, 0, st, (__hip_bfloat16*)y, xd, a, n); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}
hipError_t launch_axpy_from(int prec_y, void* y, double a, int prec_x, const void* x, size_t n,
                            hipStream_t st) {
  if (prec_y != P64) return hipErrorInvalidValue;
  const dim3 g = grid_for(n), b(256);
  switch (prec_x) {
    case P64: hipLaunchKernelGGL((k_axpy_from<double>), g, b, 0, st, (double*)y, a, (const double*)x, n); break;
    case P32: hipLaunchKernelGGL((k_axpy_from<float>), g, b, 0, st, (double*)y, a, (const float*)x, n); break;
    case P16: hipLaunchKernelGGL((k_axpy_from<__half>), g, b, 0, st, (double*)y, a, (const __half*)x, n); break;
    case PBF16: hipLaunchKernelGGL((k_axpy_from<__hip_bfloat16>), g, b, 0, st, (double*)y, a, (const __hip_bfloat16*)x, n); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t launch_convert(int prec_out, void* y, int prec_in, const void* x, size_t n,
                          hipStream_t st) {
  switch (prec_out) {
    case P64: return convert_from<double>(y, prec_in, x, n, st);
    case P32: return convert_from<float>(y, prec_in, x, n, st);
    case P16: return convert_from<__half>(y, prec_in, x, n, st);
    case PBF16: return convert_from<__hip_bfloat16>(y, prec_in, x, n, st);
  }
  return hipErrorInvalidValue;
}

// ---------------------------------------------------------------------------
// reductions: block partials, then the last block to finish sums the partials
// in a fixed order (deterministic) and resets the counter.
// ---------------------------------------------------------------------------

template <int NR>
__device__ void block_finish(double (&v)[NR], double* partial, unsigned* counter,
                             double* result) {
  __shared__ double sh[NR][4];
  __shared__ bool amlast;
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  const int nw = (blockDim.x + 63) >> 6;
#pragma unroll
  for (int r = 0; r < NR; ++r) {
    double t = wave_sum(v[r]);
    if (lane == 0) sh[r][wid] = t;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
#pragma unroll
    for (int r = 0; r < NR; ++r) {
      double t = 0;
      for (int w = 0; w < nw; ++w) t += sh[r][w];
      partial[(size_t)blockIdx.x * NR + r] = t;
    }
    if (!kRedTwoKernels) {
      __threadfence();
      unsigned prev = atomicAdd(counter, 1u);
      amlast = (prev == gridDim.x - 1);
    }
  }
  if (kRedTwoKernels) return;   // k_red_finish adds the block partials
  __syncthreads();
  if (!amlast) return;
  __threadfence();
  double w[NR];
#pragma unroll
  for (int r = 0; r < NR; ++r) w[r] = 0;
  for (unsigned i = threadIdx.x; i < gridDim.x; i += blockDim.x) {
#pragma unroll
    for (int r = 0; r < NR; ++r)
      w[r] += __hip_atomic_load(partial + (size_t)i * NR + r, __ATOMIC_RELAXED,
                                __HIP_MEMORY_SCOPE_AGENT);
  }
  __syncthreads();
#pragma unroll
  for (int r = 0; r < NR; ++r) {
    double t = wave_sum(w[r]);
    if (lane == 0) sh[r][wid] = t;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
#pragma unroll
    for (int r = 0; r < NR; ++r) {
      double t = 0;
      for (int ww = 0; ww < nw; ++ww) t += sh[r][ww];
      result[r] = t;
    }
    __hip_atomic_store(counter, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
}

// Second pass of a reduction: block y adds the nblocks partials (NR values
// each) at partial + y*pstride in a fixed order into result + y*rstride.
template <int NR>
__global__ void __launch_bounds__(256)
k_red_finish(const double* __restrict__ partial, unsigned nblocks, size_t pstride,
             double* __restrict__ result, int rstride, double* host) {
  const double* p = partial + (size_t)blockIdx.y * pstride;
  double w[NR];
#pragma unroll
  for (int r = 0; r < NR; ++r) w[r] = 0;
  for (unsigned i = threadIdx.x; i < nblocks; i += blockDim.x)
#pragma unroll
    for (int r = 0; r < NR; ++r) w[r] += p[(size_t)i * NR + r];
  __shared__ double sh[NR][4];
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
  for (int r = 0; r < NR; ++r) {
    const double t = wave_sum(w[r]);
    if (lane == 0) sh[r][wid] = t;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
#pragma unroll
    for (int r = 0; r < NR; ++r) {
      const double t = sh[r][0] + sh[r][1] + sh[r][2] + sh[r][3];
      result[(size_t)blockIdx.y * rstride + r] = t;
      if (host) host[(size_t)blockIdx.y * rstride + r] = t;   // coherent host memory
    }
    if (host) __threadfence_system();
  }
}

template <int NR>
static void red_finish(const double* partial, unsigned nblocks, double* result, hipStream_t st,
                       unsigned ny = 1, size_t pstride = 0, int rstride = 0) {
  if (!kRedTwoKernels) return;
  hipLaunchKernelGGL(k_red_finish<NR>, dim3(1, ny), dim3(256), 0, st, partial, nblocks, pstride,
                     result, rstride, (double*)nullptr);
}

hipError_t launch_finish_partials(int nr, const double* partial, unsigned nblocks,
                                  double* result, hipStream_t st) {
  if (nr == 1)
    hipLaunchKernelGGL(k_red_finish<1>, dim3(1, 1), dim3(256), 0, st, partial, nblocks, (size_t)0,
                       result, 0, (double*)nullptr);
  else if (nr == 2)
    hipLaunchKernelGGL(k_red_finish<2>, dim3(1, 1), dim3(256), 0, st, partial, nblocks, (size_t)0,
                       result, 0, (double*)nullptr);
  else
    return hipErrorInvalidValue;
  return hipGetLastError();
}

static void finish_norm(const double* partial, unsigned nblocks, double* result, hipStream_t st,
                        double* host) {
  hipLaunchKernelGGL(k_red_finish<1>, dim3(1, 1), dim3(256), 0, st, partial, nblocks, (size_t)0,
                     result, 0, host);
}

static dim3 red_grid(size_t n) {
  size_t g = (n + 255) / 256;
  if (g > (size_t)kRedBlocks) g = kRedBlocks;
  if (g < 1) g = 1;
  return dim3((unsigned)g);
}


template <typename T>
__global__ void __launch_bounds__(256)
k_norm2(const T* __restrict__ x, size_t n, double* partial, unsigned* counter, double* result) {
  using C = typename St<T>::C;
  double v[1] = {0};
  PACK_LOOP(T, n) {
    C a[NV];
    Vec<T>::ld(x + ip * NV, a);
#pragma unroll
    for (int j = 0; j < NV; ++j) v[0] += (double)a[j] * (double)a[j];
  }
  TAIL_LOOP(n) {
    double a = (double)St<T>::ld(x + i);
    v[0] += a * a;
  }
  block_finish<1>(v, partial, counter, result);
}
template <typename T>
__global__ void __launch_bounds__(256)
k_dot(const T* __restrict__ x, const T* __restrict__ y, size_t n, double* partial,
      unsigned* counter, double* result) {
  using C = typename St<T>::C;
  double v[1] = {0};
  PACK_LOOP(T, n) {
    C a[NV], b[NV];
    Vec<T>::ld(x + ip * NV, a);
    Vec<T>::ld(y + ip * NV, b);
#pragma unroll
    for (int j = 0; j < NV; ++j) v[0] += (double)a[j] * (double)b[j];
  }
  TAIL_LOOP(n) v[0] += (double)St<T>::ld(x + i) * (double)St<T>::ld(y + i);
  block_finish<1>(v, partial, counter, result);
}
// complex <x, y> with the device layout: real and imaginary parts are the
// even/odd k of the same (site, s)
template <typename T>
__global__ void __launch_bounds__(256)
k_dotc(const T* __restrict__ x, const T* __restrict__ y, size_t npairs, int Ls, double* partial,
       unsigned* counter, double* result) {
  double v[2] = {0, 0};
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < npairs;
       i += (size_t)gridDim.x * blockDim.x) {
    const int s = (int)(i % Ls);
    const size_t r = i / Ls;            // site * 12 + c2
    const size_t idx = (r * 2) * Ls + s;  // k = 2 * c2 (real), k + 1 imaginary
    double xr = St<T>::ld(x + idx), xi = St<T>::ld(x + idx + Ls);
    double yr = St<T>::ld(y + idx), yi = St<T>::ld(y + idx + Ls);
    v[0] += xr * yr + xi * yi;
    v[1] += xr * yi - xi * yr;
  }
  block_finish<2>(v, partial, counter, result);
}
// x += a p;  r -= a s;  |r|^2 of the stored r
template <typename T>
__global__ void __launch_bounds__(256)
k_cg_xr(T* __restrict__ x, T* __restrict__ r, const T* __restrict__ p, const T* __restrict__ s,
        typename St<T>::C a, size_t n, double* partial, unsigned* counter, double* result) {
  using C = typename St<T>::C;
  double v[1] = {0};
  PACK_LOOP(T, n) {
    C xv[NV], rv[NV], pv[NV], sv[NV];
    Vec<T>::ld(x + ip * NV, xv);
    Vec<T>::ld(p + ip * NV, pv);
    Vec<T>::ld(r + ip * NV, rv);
    Vec<T>::ld(s + ip * NV, sv);
#pragma unroll
    for (int j = 0; j < NV; ++j) {
      xv[j] += a * pv[j];
      rv[j] = rounded<T>(rv[j] - a * sv[j]);
      v[0] += (double)rv[j] * (double)rv[j];
    }
    Vec<T>::st(x + ip * NV, xv);
    Vec<T>::st(r + ip * NV, rv);
  }
  TAIL_LOOP(n) {
    St<T>::st(x + i, St<T>::ld(x + i) + a * St<T>::ld(p + i));
    C rn = rounded<T>(St<T>::ld(r + i) - a * St<T>::ld(s + i));
    St<T>::st(r + i, rn);
    v[0] += (double)rn * (double)rn;
  }
  block_finish<1>(v, partial, counter, result);
}
// x += a p;  r_out = r_in - a (s + sig p) (r_out another buffer);  |r_out|^2
template <typename T>
__global__ void __launch_bounds__(256)
k_cg_xr_to(T* __restrict__ x, T* __restrict__ rout, const T* __restrict__ rin,
           const T* __restrict__ p, const T* __restrict__ s, typename St<T>::C a,
           typename St<T>::C sig, size_t n, double* partial, unsigned* counter,
           double* result) {
  using C = typename St<T>::C;
  double v[1] = {0};
  PACK_LOOP(T, n) {
    C xv[NV], rv[NV], pv[NV], sv[NV];
    Vec<T>::ld(x + ip * NV, xv);
    Vec<T>::ld(p + ip * NV, pv);
    Vec<T>::ld(rin + ip * NV, rv);
    Vec<T>::ld(s + ip * NV, sv);
#pragma unroll
    for (int j = 0; j < NV; ++j) {
      xv[j] += a * pv[j];
      rv[j] = rounded<T>(rv[j] - a * (sv[j] + sig * pv[j]));
      v[0] += (double)rv[j] * (double)rv[j];
    }
    Vec<T>::st(x + ip * NV, xv);
    Vec<T>::st(rout + ip * NV, rv);
  }
  TAIL_LOOP(n) {
    const C pv = St<T>::ld(p + i);
    St<T>::st(x + i, St<T>::ld(x + i) + a * pv);
    C rn = rounded<T>(St<T>::ld(rin + i) - a * (St<T>::ld(s + i) + sig * pv));
    St<T>::st(rout + i, rn);
    v[0] += (double)rn * (double)rn;
  }
  block_finish<1>(v, partial, counter, result);
}
// y = a y + x;  |y|^2 of the stored y
template <typename T>
__global__ void __launch_bounds__(256)
k_aypx_norm(typename St<T>::C a, T* __restrict__ y, const T* __restrict__ x, size_t n,
            double* partial, unsigned* counter, double* result) {
  using C = typename St<T>::C;
  double v[1] = {0};
  PACK_LOOP(T, n) {
    C xv[NV], yv[NV];
    Vec<T>::ld(x + ip * NV, xv);
    Vec<T>::ld(y + ip * NV, yv);
#pragma unroll
    for (int j = 0; j < NV; ++j) {
      yv[j] = rounded<T>(a * yv[j] + xv[j]);
      v[0] += (double)yv[j] * (double)yv[j];
    }
    Vec<T>::st(y + ip * NV, yv);
  }
  TAIL_LOOP(n) {
    const C yv = rounded<T>(a * St<T>::ld(y + i) + St<T>::ld(x + i));
    St<T>::st(y + i, yv);
    v[0] += (double)yv * (double)yv;
  }
  block_finish<1>(v, partial, counter, result);
}
// y += a x;  Re<y, x> of the stored y
template <typename T>
__global__ void __launch_bounds__(256)
k_axpy_dot(T* __restrict__ y, typename St<T>::C a, const T* __restrict__ x, size_t n,
           double* partial, unsigned* counter, double* result) {
  using C = typename St<T>::C;
  double v[1] = {0};
  PACK_LOOP(T, n) {
    C xv[NV], yv[NV];
    Vec<T>::ld(x + ip * NV, xv);
    Vec<T>::ld(y + ip * NV, yv);
#pragma unroll
    for (int j = 0; j < NV; ++j) {
      yv[j] = rounded<T>(yv[j] + a * xv[j]);
      v[0] += (double)yv[j] * (double)xv[j];
    }
    Vec<T>::st(y + ip * NV, yv);
  }
  TAIL_LOOP(n) {
    const C xv = St<T>::ld(x + i);
    const C yv = rounded<T>(St<T>::ld(y + i) + a * xv);
    St<T>::st(y + i, yv);
    v[0] += (double)yv * (double)xv;
  }
  block_finish<1>(v, partial, counter, result);
}
// r = ax - b;  |r|^2 of the stored r
template <typename T>
__global__ void __launch_bounds__(256)
k_sub_norm2(T* __restrict__ r, const T* __restrict__ ax, const T* __restrict__ b, size_t n,
            double* partial, unsigned* counter, double* result) {
  using C = typename St<T>::C;
  double v[1] = {0};
  PACK_LOOP(T, n) {
    C av[NV], bv[NV];
    Vec<T>::ld(ax + ip * NV, av);
    Vec<T>::ld(b + ip * NV, bv);
#pragma unroll
    for (int j = 0; j < NV; ++j) {
      av[j] = rounded<T>(av[j] - bv[j]);
      v[0] += (double)av[j] * (double)av[j];
    }
    Vec<T>::st(r + ip * NV, av);
  }
  TAIL_LOOP(n) {
    const C t = rounded<T>(St<T>::ld(ax + i) - St<T>::ld(b + i));
    St<T>::st(r + i, t);
    v[0] += (double)t * (double)t;
  }
  block_finish<1>(v, partial, counter, result);
}

template <typename T>
static void norm2_t(const void* x, size_t n, const RedScratch& r, int slot, hipStream_t st) {
  hipLaunchKernelGGL(k_norm2<T>, red_grid(n), dim3(256), 0, st, (const T*)x, n, r.partial,
                     r.counter, r.result + slot);
  red_finish<1>(r.partial, red_grid(n).x, r.result + slot, st);
}
template <typename T>
static void dot_t(const void* x, const void* y, size_t n, const RedScratch& r, int slot,
                  hipStream_t st) {
  hipLaunchKernelGGL(k_dot<T>, red_grid(n), dim3(256), 0, st, (const T*)x, (const T*)y, n,
                     r.partial, r.counter, r.result + slot);
  red_finish<1>(r.partial, red_grid(n).x, r.result + slot, st);
}
template <typename T>
static void cg_xr_t(void* x, void* rr, const void* p, const void* s, double a, size_t n,
                    const RedScratch& r, int slot, hipStream_t st) {
  hipLaunchKernelGGL(k_cg_xr<T>, red_grid(n), dim3(256), 0, st, (T*)x, (T*)rr, (const T*)p,
                     (const T*)s, (typename St<T>::C)a, n, r.partial, r.counter, r.result + slot);
  red_finish<1>(r.partial, red_grid(n).x, r.result + slot, st);
}
template <typename T>
static void cg_xr_to_t(void* x, void* rout, const void* rin, const void* p, const void* s,
                       double a, double sig, size_t n, const RedScratch& r, int slot,
                       hipStream_t st) {
  hipLaunchKernelGGL(k_cg_xr_to<T>, red_grid(n), dim3(256), 0, st, (T*)x, (T*)rout,
                     (const T*)rin, (const T*)p, (const T*)s, (typename St<T>::C)a,
                     (typename St<T>::C)sig, n, r.partial, r.counter, r.result + slot);
  red_finish<1>(r.partial, red_grid(n).x, r.result + slot, st);
}
template <typename T>
static void aypx_norm_t(double a, void* y, const void* x, size_t n, const RedScratch& r, int slot,
                        hipStream_t st) {
  hipLaunchKernelGGL(k_aypx_norm<T>, red_grid(n), dim3(256), 0, st, (typename St<T>::C)a, (T*)y,
                     (const T*)x, n, r.partial, r.counter, r.result + slot);
  red_finish<1>(r.partial, red_grid(n).x, r.result + slot, st);
}
template <typename T>
static void axpy_dot_t(void* y, double a, const void* x, size_t n, const RedScratch& r, int slot,
                       hipStream_t st) {
  hipLaunchKernelGGL(k_axpy_dot<T>, red_grid(n), dim3(256), 0, st, (T*)y, (typename St<T>::C)a,
                     (const T*)x, n, r.partial, r.counter, r.result + slot);
  red_finish<1>(r.partial, red_grid(n).x, r.result + slot, st);
}
template <typename T>
static void sub_norm2_t(void* rr, const void* ax, const void* b, size_t n, const RedScratch& r,
                        int slot, hipStream_t st) {
  hipLaunchKernelGGL(k_sub_norm2<T>, red_grid(n), dim3(256), 0, st, (T*)rr, (const T*)ax,
                     (const T*)b, n, r.partial, r.counter, r.result + slot);
  red_finish<1>(r.partial, red_grid(n).x, r.result + slot, st);
}

hipError_t launch_norm2(int prec, const void* x, size_t n, const RedScratch& r, int slot,
                        hipStream_t st) {
  DWF_DISPATCH(prec, norm2_t, (x, n, r, slot, st));
  return hipGetLastError();
}
hipError_t launch_dot(int prec, const void* x, const void* y, size_t n, const RedScratch& r,
                      int slot, hipStream_t st) {
  DWF_DISPATCH(prec, dot_t, (x, y, n, r, slot, st));
  return hipGetLastError();
}
hipError_t launch_dotc_ls(int prec, const void* x, const void* y, size_t n, int Ls,
                          const RedScratch& r, int slot, hipStream_t st) {
  const size_t np = n / 2;
  switch (prec) {
    case P64: hipLaunchKernelGGL(k_dotc<double>, red_grid(np), dim3(256), 0, st, (const double*)x, (const double*)y, np, Ls, r.partial, r.counter, r.result + slot); break;
    case P32: hipLaunchKernelGGL(k_dotc<float>, red_grid(np), dim3(256), 0, st, (const float*)x, (const float*)y, np, Ls, r.partial, r.counter, r.result + slot); break;
    case P16: hipLaunchKernelGGL(k_dotc<__half>, red_grid(np), dim3(256), 0, st, (const __half*)x, (const __half*)y, np, Ls, r.partial, r.counter, r.result + slot); break;
    case PBF16: hipLaunchKernelGGL(k_dotc<__hip_bfloat16>, red_grid(np), dim3(256), 0, st, (const __hip_bfloat16*)x, (const __hip_bfloat16*)y, np, Ls, r.partial, r.counter, r.result + slot); break;
    default: return hipErrorInvalidValue;
  }
  red_finish<2>(r.partial, red_grid(np).x, r.result + slot, st);
  return hipGetLastError();
}
hipError_t launch_cg_xr(int prec, void* x, void* r, const void* p, const void* s, double a,
                        size_t n, const RedScratch& red, int slot, hipStream_t st) {
  DWF_DISPATCH(prec, cg_xr_t, (x, r, p, s, a, n, red, slot, st));
  return hipGetLastError();
}

hipError_t launch_cg_xr_to(int prec, void* x, void* rout, const void* rin, const void* p,
                           const void* s, double a, double sig, size_t n, const RedScratch& red,
                           int slot, hipStream_t st) {
  DWF_DISPATCH(prec, cg_xr_to_t, (x, rout, rin, p, s, a, sig, n, red, slot, st));
  return hipGetLastError();
}
hipError_t launch_aypx_norm(int prec, double a, void* y, const void* x, size_t n,
                            const RedScratch& red, int slot, hipStream_t st) {
  DWF_DISPATCH(prec, aypx_norm_t, (a, y, x, n, red, slot, st));
  return hipGetLastError();
}

// CG updates with the coefficients in device memory (no host round trip):
//   x += a p, r -= a s, |r|^2 -> result[slot], a = num[0] / den[0]
//   p = (num[0] / den[0]) p + r
template <typename T>
__global__ void __launch_bounds__(256)
k_cg_xr_dev(T* __restrict__ x, T* __restrict__ r, const T* __restrict__ p,
            const T* __restrict__ s, const double* num, const double* den, size_t n,
            double* partial, unsigned* counter, double* result) {
  using C = typename St<T>::C;
  const C a = (C)(num[0] / den[0]);
  double v[1] = {0};
  PACK_LOOP(T, n) {
    C xv[NV], rv[NV], pv[NV], sv[NV];
    Vec<T>::ld(x + ip * NV, xv);
    Vec<T>::ld(p + ip * NV, pv);
    Vec<T>::ld(r + ip * NV, rv);
    Vec<T>::ld(s + ip * NV, sv);
#pragma unroll
    for (int j = 0; j < NV; ++j) {
      xv[j] += a * pv[j];
      rv[j] = rounded<T>(rv[j] - a * sv[j]);
      v[0] += (double)rv[j] * (double)rv[j];
    }
    Vec<T>::st(x + ip * NV, xv);
    Vec<T>::st(r + ip * NV, rv);
  }
  TAIL_LOOP(n) {
    St<T>::st(x + i, St<T>::ld(x + i) + a * St<T>::ld(p + i));
    C rn = rounded<T>(St<T>::ld(r + i) - a * St<T>::ld(s + i));
    St<T>::st(r + i, rn);
    v[0] += (double)rn * (double)rn;
  }
  block_finish<1>(v, partial, counter, result);
}
template <typename T>
__global__ void __launch_bounds__(256)
k_aypx_dev(T* __restrict__ y, const T* __restrict__ x, const double* num, const double* den,
           size_t n) {
  using C = typename St<T>::C;
  const C a = (C)(num[0] / den[0]);
  PACK_LOOP(T, n) {
    C yv[NV], xv[NV];
    Vec<T>::ld(y + ip * NV, yv);
    Vec<T>::ld(x + ip * NV, xv);
#pragma unroll
    for (int j = 0; j < NV; ++j) yv[j] = a * yv[j] + xv[j];
    Vec<T>::st(y + ip * NV, yv);
  }
  TAIL_LOOP(n) St<T>::st(y + i, a * St<T>::ld(y + i) + St<T>::ld(x + i));
}
template <typename T>
static void cg_xr_dev_t(void* x, void* rr, const void* p, const void* s, const double* num,
                        const double* den, size_t n, const RedScratch& r, int slot,
                        hipStream_t st) {
  hipLaunchKernelGGL(k_cg_xr_dev<T>, red_grid(n), dim3(256), 0, st, (T*)x, (T*)rr, (const T*)p,
                     (const T*)s, num, den, n, r.partial, r.counter, r.result + slot);
  red_finish<1>(r.partial, red_grid(n).x, r.result + slot, st);
}
template <typename T>
static void aypx_dev_t(void* y, const void* x, const double* num, const double* den, size_t n,
                       hipStream_t st) {
  hipLaunchKernelGGL(k_aypx_dev<T>, grid_for(n), dim3(256), 0, st, (T*)y, (const T*)x, num, den,
                     n);
}
hipError_t launch_cg_xr_dev(int prec, void* x, void* r, const void* p, const void* s,
                            const double* num, const double* den, size_t n,
                            const RedScratch& red, int slot, hipStream_t st) {
  DWF_DISPATCH(prec, cg_xr_dev_t, (x, r, p, s, num, den, n, red, slot, st));
  return hipGetLastError();
}
hipError_t launch_aypx_dev(int prec, void* y, const void* x, const double* num, const double* den,
                           size_t n, hipStream_t st) {
  DWF_DISPATCH(prec, aypx_dev_t, (y, x, num, den, n, st));
  return hipGetLastError();
}

// The rest of a CG step once r (and |r|^2) are updated:
//   x += (rr[0] / pap[0]) p,  p = r + (rr_new[0] / rr[0]) p
template <typename T>
__global__ void __launch_bounds__(256)
k_cg_px_dev(T* __restrict__ x, T* __restrict__ p, const T* __restrict__ r, const double* rr,
            const double* pap, const double* rr_new, size_t n) {
  using C = typename St<T>::C;
  const C a = (C)(rr[0] / pap[0]);
  const C b = (C)(rr_new[0] / rr[0]);
  PACK_LOOP(T, n) {
    C xv[NV], pv[NV], rv[NV];
    Vec<T>::ld(x + ip * NV, xv);
    Vec<T>::ld(p + ip * NV, pv);
    Vec<T>::ld(r + ip * NV, rv);
#pragma unroll
    for (int j = 0; j < NV; ++j) {
      xv[j] += a * pv[j];
      pv[j] = rv[j] + b * pv[j];
    }
    Vec<T>::st(x + ip * NV, xv);
    Vec<T>::st(p + ip * NV, pv);
  }
  TAIL_LOOP(n) {
    const C pv = St<T>::ld(p + i);
    St<T>::st(x + i, St<T>::ld(x + i) + a * pv);
    St<T>::st(p + i, St<T>::ld(r + i) + b * pv);
  }
}
template <typename T>
static void cg_px_dev_t(void* x, void* p, const void* r, const double* rr, const double* pap,
                        const double* rr_new, size_t n, hipStream_t st) {
  hipLaunchKernelGGL(k_cg_px_dev<T>, grid_for(n), dim3(256), 0, st, (T*)x, (T*)p, (const T*)r,
                     rr, pap, rr_new, n);
}
hipError_t launch_cg_px_dev(int prec, void* x, void* p, const void* r, const double* rr,
                            const double* pap, const double* rr_new, size_t n, hipStream_t st) {
  DWF_DISPATCH(prec, cg_px_dev_t, (x, p, r, rr, pap, rr_new, n, st));
  return hipGetLastError();
}
hipError_t launch_axpy_dot(int prec, void* y, double a, const void* x, size_t n,
                           const RedScratch& red, int slot, hipStream_t st) {
  DWF_DISPATCH(prec, axpy_dot_t, (y, a, x, n, red, slot, st));
  return hipGetLastError();
}
hipError_t launch_sub_norm2(int prec, void* r, const void* ax, const void* b, size_t n,
                            const RedScratch& red, int slot, hipStream_t st) {
  DWF_DISPATCH(prec, sub_norm2_t, (r, ax, b, n, red, slot, st));
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// multishift update of all active shifts j = 1..m-1 in one launch
// (ashiftsolver_CG-tmpl.h:243-266): x_j -= beta_s p_j; p_j = alpha_s p_j + zeta r;
// pp[j] = |p_j|^2.  blockIdx.y = j - 1; double precision fields.
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256)
k_shift_update(double* const* __restrict__ x, double* const* __restrict__ p,
               const double* __restrict__ r, const double* __restrict__ coef, size_t n,
               double* partial, unsigned* counters, double* pp) {
  const int j = blockIdx.y + 1;
  double* xj = x[j];
  double* pj = p[j];
  const double c0 = coef[3 * j], c1 = coef[3 * j + 1], c2 = coef[3 * j + 2];
  double v[1] = {0};
  PACK_LOOP(double, n) {
    double xv[2], pv[2], rv[2];
    Vec<double>::ld(xj + ip * 2, xv);
    Vec<double>::ld(pj + ip * 2, pv);
    Vec<double>::ld(r + ip * 2, rv);
#pragma unroll
    for (int k = 0; k < 2; ++k) {
      xv[k] += c0 * pv[k];
      pv[k] = c1 * pv[k] + c2 * rv[k];
      v[0] += pv[k] * pv[k];
    }
    Vec<double>::st(xj + ip * 2, xv);
    Vec<double>::st(pj + ip * 2, pv);
  }
  TAIL_LOOP(n) {
    const double pv = pj[i];
    xj[i] += c0 * pv;
    const double np = c1 * pv + c2 * r[i];
    pj[i] = np;
    v[0] += np * np;
  }
  block_finish<1>(v, partial + (size_t)blockIdx.y * kRedBlocks, counters + blockIdx.y,
                  pp + j);
}

// The same update with the shifts looped inside each thread (m <= kShiftMax):
// r is read once per element instead of once per shift (4 passes per shift
// + 1 instead of 5 per shift).  The partial sums of |p_j|^2 are laid out as in
// k_shift_update (region j - 1 of kRedBlocks per shift).  n must be even.
constexpr int kShiftMax = 32;
__global__ void __launch_bounds__(256)
k_shift_update_all(double* const* __restrict__ x, double* const* __restrict__ p,
                   const double* __restrict__ r, const double* __restrict__ coef, int m,
                   size_t n, double* partial) {
  double v[kShiftMax];
#pragma unroll
  for (int j = 0; j < kShiftMax; ++j) v[j] = 0;
  PACK_LOOP(double, n) {
    double rv[2];
    Vec<double>::ld(r + ip * 2, rv);
#pragma unroll
    for (int j = 1; j < kShiftMax; ++j) {
      if (j < m) {
        double* xj = x[j];
        double* pj = p[j];
        const double c0 = coef[3 * j], c1 = coef[3 * j + 1], c2 = coef[3 * j + 2];
        double xv[2], pv[2];
        Vec<double>::ld(xj + ip * 2, xv);
        Vec<double>::ld(pj + ip * 2, pv);
#pragma unroll
        for (int k = 0; k < 2; ++k) {
          xv[k] += c0 * pv[k];
          pv[k] = c1 * pv[k] + c2 * rv[k];
          v[j] += pv[k] * pv[k];
        }
        Vec<double>::st(xj + ip * 2, xv);
        Vec<double>::st(pj + ip * 2, pv);
      }
    }
  }
  // (n is even: no tail; 256-thread blocks: 4 waves)
  __shared__ double sh[4];
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
  for (int j = 1; j < kShiftMax; ++j) {
    if (j < m) {
      const double t = wave_sum(v[j]);
      if (lane == 0) sh[wid] = t;
      __syncthreads();
      if (threadIdx.x == 0)
        partial[(size_t)(j - 1) * kRedBlocks + blockIdx.x] = sh[0] + sh[1] + sh[2] + sh[3];
      __syncthreads();
    }
  }
}

hipError_t launch_shift_update(void* const* x, void* const* p, const void* r,
                               const double* coef, int m, size_t n, double* partial,
                               unsigned* counters, double* pp, hipStream_t st) {
  if (m <= 1) return hipSuccess;
  dim3 g = red_grid(n);
  if (m <= kShiftMax && n % 2 == 0 && kRedTwoKernels) {
    hipLaunchKernelGGL(k_shift_update_all, g, dim3(256), 0, st, (double* const*)x,
                       (double* const*)p, (const double*)r, coef, m, n, partial);
    red_finish<1>(partial, g.x, pp + 1, st, m - 1, kRedBlocks, 1);
    return hipGetLastError();
  }
  g.y = m - 1;
  hipLaunchKernelGGL(k_shift_update, g, dim3(256), 0, st, (double* const*)x, (double* const*)p,
                     (const double*)r, coef, n, partial, counters, pp);
  red_finish<1>(partial, g.x, pp + 1, st, g.y, kRedBlocks, 1);
  return hipGetLastError();
}

// Krylov-block fold of the multishift CG.  After m <= kFoldMax iterations
// whose unshifted residuals are R_1 .. R_m, shift j's vectors are linear
// combinations of their values at the start of the block and the R_l:
//   x_j <- x_j + b_j p_j + sum_l d_jl R_l,   p_j <- a_j p_j + sum_l c_jl R_l,
// coef per shift j (1 <= j < nsh): [a, b, c_1 .. c_K, d_1 .. d_K], K = kFoldMax,
// zero beyond m.  One pass reads each R_l once for all shifts (instead of r
// and x_j, p_j of every shift every iteration); |p_j|^2 per shift as block
// partials (region j - 1 of kRedBlocks), as k_shift_update_all.
__global__ void __launch_bounds__(256)
k_ms_fold(double* const* __restrict__ x, double* const* __restrict__ p,
          const double* const* __restrict__ R, const double* __restrict__ coef, int nsh, int m,
          size_t n, double* partial) {
  double v[kShiftMax];
#pragma unroll
  for (int j = 0; j < kShiftMax; ++j) v[j] = 0;
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x) {
    double rv[kFoldMax];
#pragma unroll
    for (int l = 0; l < kFoldMax; ++l) rv[l] = l < m ? R[l][i] : 0.0;
#pragma unroll
    for (int j = 1; j < kShiftMax; ++j) {
      if (j < nsh) {
        const double* c = coef + j * kFoldCoef;
        double* pj = p[j];
        double* xj = x[j];
        const double pv = pj[i];
        double pn = c[0] * pv, xn = xj[i] + c[1] * pv;
#pragma unroll
        for (int l = 0; l < kFoldMax; ++l) {
          pn += c[2 + l] * rv[l];
          xn += c[2 + kFoldMax + l] * rv[l];
        }
        xj[i] = xn;
        pj[i] = pn;
        v[j] += pn * pn;
      }
    }
  }
  __shared__ double sh[4];
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
  for (int j = 1; j < kShiftMax; ++j) {
    if (j < nsh) {
      const double t = wave_sum(v[j]);
      if (lane == 0) sh[wid] = t;
      __syncthreads();
      if (threadIdx.x == 0)
        partial[(size_t)(j - 1) * kRedBlocks + blockIdx.x] = sh[0] + sh[1] + sh[2] + sh[3];
      __syncthreads();
    }
  }
}

hipError_t launch_ms_fold(void* const* x, void* const* p, const void* const* R,
                          const double* coef, int nsh, int m, size_t n, double* partial,
                          double* pp, hipStream_t st) {
  if (nsh <= 1) return hipSuccess;
  if (nsh > kShiftMax || m < 1 || m > kFoldMax) return hipErrorInvalidValue;
  const dim3 g = red_grid(n);
  hipLaunchKernelGGL(k_ms_fold, g, dim3(256), 0, st, (double* const*)x, (double* const*)p,
                     (const double* const*)R, coef, nsh, m, n, partial);
  red_finish<1>(partial, g.x, pp + 1, st, nsh - 1, kRedBlocks, 1);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// rational combination (afopr_Rational-tmpl.h:211-215):
//   v = a0 b + sum_j c[j] x[j],  j < m;  one read of each field, one write of v
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256)
k_lincomb(double* __restrict__ v, const double* __restrict__ b, double a0,
          const double* const* __restrict__ x, const double* __restrict__ c, int m, size_t n) {
  PACK_LOOP(double, n) {
    double acc[2], t[2];
    Vec<double>::ld(b + ip * 2, acc);
    acc[0] *= a0;
    acc[1] *= a0;
    for (int j = 0; j < m; ++j) {
      Vec<double>::ld(x[j] + ip * 2, t);
      const double cj = c[j];
      acc[0] += cj * t[0];
      acc[1] += cj * t[1];
    }
    Vec<double>::st(v + ip * 2, acc);
  }
  TAIL_LOOP(n) {
    double acc = a0 * b[i];
    for (int j = 0; j < m; ++j) acc += c[j] * x[j][i];
    v[i] = acc;
  }
}

hipError_t launch_lincomb(void* v, const void* b, double a0, void* const* x, const double* c,
                          int m, size_t n, hipStream_t st) {
  hipLaunchKernelGGL(k_lincomb, red_grid(n), dim3(256), 0, st, (double*)v, (const double*)b, a0,
                     (const double* const*)x, c, m, n);
  return hipGetLastError();
}

}  // namespace dwf
