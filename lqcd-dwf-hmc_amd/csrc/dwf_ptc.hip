// dwf_ptc.hip -- gauge-covariant parallel-transport convolution (PTC) network
// kernels: the learned preconditioner of the flexible CG.
//
// A layer maps C_in channels to C_out channels of 5D spinor fields on the full
// lattice (both parities):
//   out[a](x, s) = sum_{q < 11} sum_b W[q][a][b] T_q(in[b])(x, s)
// with W[q][a][b] a complex 4x4 matrix on spin (identity on colour) and the
// transports
//   q = 0          identity
//   q = 1 + dir    U(x, dir) in(x + dir), dir = 2 mu + fb the 8 hops of the
//                  operator (the packed links of the even-odd operator, boundary
//                  phases included: forward U_mu(x), backward U_mu(x - mu)^+)
//   q = 9, 10      in(x, s + 1), in(x, s - 1)   (zero beyond the walls)
// Every T_q commutes with gauge transformations of in and out, so the layer
// is gauge covariant.  The adjoint layer (the backward pass for the input
// gradient) is the same kernel with the reversed transport (dir -> dir ^ 1,
// 9 <-> 10) and W^+.  Weights are stored in bfloat16 and applied in FP32.
//
// Activation layout: channel-major, then parity, then the device spinor
// layout of the operator, element ((site * 24 + k) * Ls + s), k = ri + 2 (c + 3 d).
#include <cstdlib>
#include <cstring>

#include "dwf_device.h"
#include "dwf_kernels.h"

namespace dwf {

constexpr int kPtcPaths = 11;

// (re, im) += (ar + i ai)(br + i bi), two fused multiply-adds per part
__device__ __forceinline__ void cmac(float& re, float& im, float ar, float ai, float br,
                                     float bi) {
  re = fmaf(-ai, bi, fmaf(ar, br, re));
  im = fmaf(ai, br, fmaf(ar, bi, im));
}

__device__ __forceinline__ int ptc_rev(int q) {
  if (q == 0) return 0;
  if (q <= 8) return 1 + ((q - 1) ^ 1);
  return q == 9 ? 10 : 9;
}

// t = T_q(in)(p, site, s) for one channel field `in` (both parities, Vh each);
// returns false if the path leaves the fifth dimension (t = 0)
__device__ __forceinline__ bool ptc_transport(float (&t)[24], const float* __restrict__ in,
                                              const float* __restrict__ links,
                                              const int* __restrict__ nbr, int q, int p,
                                              int site, int s, int Vh, int Ls) {
  const size_t fsz = (size_t)Vh * 24 * Ls;
  if (q == 0 || q >= 9) {
    const int ss = q == 0 ? s : (q == 9 ? s + 1 : s - 1);
    if (ss < 0 || ss >= Ls) return false;
    const float* src = in + (size_t)p * fsz + (size_t)site * 24 * Ls + ss;
#pragma unroll
    for (int k = 0; k < 24; ++k) t[k] = src[k * Ls];
    return true;
  }
  const int dir = q - 1;
  const int nb = nbr[((size_t)p * 8 + dir) * Vh + site];
  const float* src = in + (size_t)(1 - p) * fsz + (size_t)nb * 24 * Ls + s;
  float w[24];
#pragma unroll
  for (int k = 0; k < 24; ++k) w[k] = src[k * Ls];
  const float* L = links + ((size_t)p * 8 + dir) * 18 * Vh + site;
  float u[18];
#pragma unroll
  for (int e = 0; e < 18; ++e) u[e] = L[(size_t)e * Vh];
#pragma unroll
  for (int d = 0; d < 4; ++d)
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      float re = 0, im = 0;
#pragma unroll
      for (int b = 0; b < 3; ++b) {
        cmac(re, im, u[2 * (b + 3 * a)], u[2 * (b + 3 * a) + 1], w[2 * (b + 3 * d)],
             w[2 * (b + 3 * d) + 1]);
      }
      t[2 * (a + 3 * d)] = re;
      t[2 * (a + 3 * d) + 1] = im;
    }
  return true;
}

// One layer (ADJ = false) or its adjoint (ADJ = true).  Thread per (parity,
// site, s), s fastest; block row blockIdx.y owns the AG output channels
// a0 .. a0 + AG - 1, so each transport T_q(in[b]) is formed once per thread and
// applied to AG output channels, and the weight addresses are uniform over the
// block (scalar loads).  W: bf16 [q][a][b][alpha][beta][ri] of the forward layer
// (C_out x C_in channels), read as one 32-bit word per complex entry (re low,
// im high); for ADJ the kernel's input channels are the forward outputs and
// out[b] = sum W^+ T_rev(q)(in[a]).  Each output channel's sum runs over q, then
// b, as in the per-channel form.
template <bool ADJ, int AG>
__global__ void __launch_bounds__(256)
k_ptc_layer(float* __restrict__ out, const float* __restrict__ in, int cin,
            const unsigned* __restrict__ W, int wcin, int wcout,
            const float* __restrict__ links, const int* __restrict__ nbr, int Vh, int Ls) {
  const size_t fsz = (size_t)Vh * 24 * Ls;
  const size_t total = (size_t)2 * Vh * Ls;
  const int a0 = blockIdx.y * AG;
  for (size_t idx = blockIdx.x * (size_t)blockDim.x + threadIdx.x; idx < total;
       idx += (size_t)gridDim.x * blockDim.x) {
    const int s = (int)(idx % Ls);
    const size_t r = idx / Ls;
    const int site = (int)(r % Vh);
    const int p = (int)(r / Vh);
    float acc[AG][24];
#pragma unroll
    for (int j = 0; j < AG; ++j)
#pragma unroll
      for (int k = 0; k < 24; ++k) acc[j][k] = 0;
    for (int q = 0; q < kPtcPaths; ++q) {
      const int qw = ADJ ? ptc_rev(q) : q;   // weight path
      for (int b = 0; b < cin; ++b) {
        float t[24];
        if (!ptc_transport(t, in + (size_t)b * 2 * fsz, links, nbr, q, p, site, s, Vh, Ls))
          continue;
#pragma unroll
        for (int j = 0; j < AG; ++j) {
          // forward: W[qw][a][b];  adjoint: (W[qw][b][a])^+
          const int a = a0 + j;
          const int wa = ADJ ? b : a, wb = ADJ ? a : b;
          const unsigned* w = W + (((size_t)qw * wcout + wa) * wcin + wb) * 16;
#pragma unroll
          for (int al = 0; al < 4; ++al)
#pragma unroll
            for (int be = 0; be < 4; ++be) {
              // element (al, be) of the applied spin matrix
              const unsigned e = w[ADJ ? (be * 4 + al) : (al * 4 + be)];
              const float wr = __uint_as_float(e << 16);
              const float wi = ADJ ? -__uint_as_float(e & 0xffff0000u)
                                   : __uint_as_float(e & 0xffff0000u);
#pragma unroll
              for (int c = 0; c < 3; ++c) {
                cmac(acc[j][2 * (c + 3 * al)], acc[j][2 * (c + 3 * al) + 1], wr, wi,
                     t[2 * (c + 3 * be)], t[2 * (c + 3 * be) + 1]);
              }
            }
        }
      }
    }
#pragma unroll
    for (int j = 0; j < AG; ++j) {
      float* o = out + (size_t)(a0 + j) * 2 * fsz + (size_t)p * fsz + (size_t)site * 24 * Ls + s;
#pragma unroll
      for (int k = 0; k < 24; ++k) o[k * Ls] = acc[j][k];
    }
  }
}

// Raw buffer resource over nbytes at base (gfx9 descriptor word 3): loads at
// byte offsets >= nbytes return 0.
__device__ __forceinline__ __amdgpu_buffer_rsrc_t ptc_rsrc(const void* base, unsigned nbytes) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(base), 0, (int)nbytes, 0x00020000);
}
// the b32 load returns the 32 loaded bits as an integer: reinterpret, never convert
__device__ __forceinline__ unsigned ptc_ldu(__amdgpu_buffer_rsrc_t r, unsigned voff,
                                            unsigned soff) {
  const auto v = __builtin_amdgcn_raw_buffer_load_b32(r, (int)voff, (int)soff, 0);
  static_assert(sizeof(v) == 4, "32-bit buffer load");
  return __builtin_bit_cast(unsigned, v);
}
__device__ __forceinline__ float ptc_ld(__amdgpu_buffer_rsrc_t r, unsigned voff, unsigned soff) {
  return __uint_as_float(ptc_ldu(r, voff, soff));
}

// Matrix-core form of the same layer (any number of output channels up to 16;
// an odd count leaves the second channel of the last tile with zero weights and
// does not store it).  Per path q and input channel b the contraction is the
// real product
//   [out_re; out_im](a, alpha) += [Wr -Wi; Wi Wr](a, alpha; beta) [t_re; t_im](beta)
// for every colour, computed as v_mfma_f32_16x16x4f32 tiles D(16x16) +=
// A(16x4) B(4x16) with
//   D rows  m = 8 a' + 2 alpha + re/im  (NMT tiles of two output channels),
//   K       beta (4 spins), one product for t_re and one for t_im,
//   columns 16 points (site, s) of one parity, consecutive in site * Ls + s.
// Lane l holds A[l % 16][l / 16] and B[l / 16][l % 16], so lane l computes
// the transport of spin beta = l / 16 at point l % 16 itself (3 colours, 36
// multiply-adds for a hop path) and holds the weight of row l % 16, column
// beta; register r of its result holds D[4 (l / 16) + r][l % 16].  One wave
// per 16 points, all output channels; each output element is a sum over
// (q, b) of 4-spin matrix-core products in FP32.
// three waves per SIMD (at most 168 registers) fit without spills up to NMT = 5
template <bool ADJ, int NMT>
__global__ void __launch_bounds__(256, NMT <= 5 ? 3 : 2)
k_ptc_layer_mfma(float* __restrict__ out, const float* __restrict__ in, int cin, int cout,
                 const unsigned* __restrict__ W, int wcin, int wcout,
                 const float* __restrict__ links, const int* __restrict__ nbr, int Vh, int Ls) {
  // loads go through buffer resources with 32-bit byte offsets: per lane the
  // point / spin part, per wave (SGPR) the channel and element part; a lane
  // whose source is outside the lattice or the fifth dimension reads at
  // kOut and gets 0.  The launcher checks that every offset fits.
  constexpr unsigned kOut = 0x80000000u;
  const int lane = threadIdx.x & 63;
  const int n = lane & 15, g = lane >> 4;
  const size_t fsz = (size_t)Vh * 24 * Ls;
  const int npts = Vh * Ls;            // points of one parity
  const int tpp = (npts + 15) / 16;    // tiles of one parity
  const int wpb = blockDim.x >> 6;
  const unsigned lsb = 4u * (unsigned)Ls, vhb = 4u * (unsigned)Vh;
  const __amdgpu_buffer_rsrc_t rl = ptc_rsrc(links, (unsigned)(2 * 8 * 18) * vhb);
  const __amdgpu_buffer_rsrc_t rw =
      ptc_rsrc(W, 4u * (unsigned)(kPtcPaths * wcin * wcout * 16));
  // the A row of this lane: channel 2 mt + am, spin row al, re (0) / im (1);
  // its weight word within the (q, b, mt) block
  const int am = n >> 3, al = (n >> 1) & 3, ro = n & 1;
  const unsigned wlane =
      4u * (ADJ ? (unsigned)(am * 16 + g * 4 + al) : (unsigned)(am * wcin * 16 + al * 4 + g));
  for (int tile = blockIdx.x * wpb + (threadIdx.x >> 6); tile < 2 * tpp;
       tile += gridDim.x * wpb) {
    const int p = tile / tpp;
    const int P = (tile - p * tpp) * 16 + n;
    const bool valid = P < npts;
    const int Pc = valid ? P : npts - 1;
    const int site = Pc / Ls, s = Pc - (Pc / Ls) * Ls;
    flt4_t acc[NMT][3];
#pragma unroll
    for (int mt = 0; mt < NMT; ++mt)
#pragma unroll
      for (int c = 0; c < 3; ++c) acc[mt][c] = flt4_t{0.f, 0.f, 0.f, 0.f};
    for (int q = 0; q < kPtcPaths; ++q) {
      const int qw = ADJ ? ptc_rev(q) : q;   // weight path
      const bool hop = q >= 1 && q <= 8;
      float u[18];
      unsigned woff;   // byte offset of spin g of the source point in a channel
      if (hop) {
        const int dir = q - 1;
        const int nb = nbr[((size_t)p * 8 + dir) * Vh + site];
        const unsigned loff = 4u * (unsigned)(((p * 8 + dir) * 18) * Vh + site);
#pragma unroll
        for (int e = 0; e < 18; ++e) u[e] = ptc_ld(rl, loff, (unsigned)e * vhb);
        woff = valid ? 4u * (unsigned)((1 - p) * fsz + ((size_t)nb * 24 + 6 * g) * Ls + s) : kOut;
      } else {
        const int ss = q == 0 ? s : (q == 9 ? s + 1 : s - 1);
        woff = valid && ss >= 0 && ss < Ls
                   ? 4u * (unsigned)(p * fsz + ((size_t)site * 24 + 6 * g) * Ls + ss)
                   : kOut;
      }
      // the spinor of input channel b at the source point and this lane's
      // weight words; loaded one channel ahead of the matrix-core products
      auto load = [&](int b, float (&w)[6], unsigned (&e)[NMT]) {
        const __amdgpu_buffer_rsrc_t ri = ptc_rsrc(in + (size_t)b * 2 * fsz, 8u * (unsigned)fsz);
#pragma unroll
        for (int k = 0; k < 6; ++k) w[k] = ptc_ld(ri, woff, (unsigned)k * lsb);
        // element (al, beta = g) of the applied spin matrix of channel
        // a = 2 mt + am: forward W[qw][a][b], adjoint (W[qw][b][a])^+
#pragma unroll
        for (int mt = 0; mt < NMT; ++mt) {
          const unsigned wblk = ADJ ? (unsigned)(((qw * wcout + b) * wcin + 2 * mt) * 16)
                                    : (unsigned)(((qw * wcout + 2 * mt) * wcin + b) * 16);
          // an odd cout pads the last tile with a zero channel
          e[mt] = ptc_ldu(rw, 2 * mt + am < cout ? wlane : kOut, 4u * wblk);
        }
      };
      float wn[6];
      unsigned en[NMT];
      load(0, wn, en);
      for (int b = 0; b < cin; ++b) {
        float w[6];
        unsigned e[NMT];
#pragma unroll
        for (int k = 0; k < 6; ++k) w[k] = wn[k];
#pragma unroll
        for (int mt = 0; mt < NMT; ++mt) e[mt] = en[mt];
        if (b + 1 < cin) load(b + 1, wn, en);
        float ar[NMT], ai[NMT];
#pragma unroll
        for (int mt = 0; mt < NMT; ++mt) {
          const float wr = __uint_as_float(e[mt] << 16);
          const float wi = ADJ ? -__uint_as_float(e[mt] & 0xffff0000u)
                               : __uint_as_float(e[mt] & 0xffff0000u);
          ar[mt] = ro ? wi : wr;    // multiplies t_re
          ai[mt] = ro ? wr : -wi;   // multiplies t_im
        }
        // B operand: the transported colour vector of spin g at this point
        float t[6];
        if (hop) {
#pragma unroll
          for (int ca = 0; ca < 3; ++ca) {
            float re = 0.f, im = 0.f;
#pragma unroll
            for (int cb = 0; cb < 3; ++cb)
              cmac(re, im, u[2 * (cb + 3 * ca)], u[2 * (cb + 3 * ca) + 1], w[2 * cb],
                   w[2 * cb + 1]);
            t[2 * ca] = re;
            t[2 * ca + 1] = im;
          }
        } else {
#pragma unroll
          for (int k = 0; k < 6; ++k) t[k] = w[k];
        }
#pragma unroll
        for (int c = 0; c < 3; ++c)
#pragma unroll
          for (int mt = 0; mt < NMT; ++mt)
            acc[mt][c] = __builtin_amdgcn_mfma_f32_16x16x4f32(ar[mt], t[2 * c], acc[mt][c], 0, 0, 0);
#pragma unroll
        for (int c = 0; c < 3; ++c)
#pragma unroll
          for (int mt = 0; mt < NMT; ++mt)
            acc[mt][c] =
                __builtin_amdgcn_mfma_f32_16x16x4f32(ai[mt], t[2 * c + 1], acc[mt][c], 0, 0, 0);
      }
    }
    if (valid) {
      // register r: row 4 g + r = channel 2 mt + (g >> 1), spin 2 (g & 1) + (r >> 1),
      // re / im r & 1
#pragma unroll
      for (int mt = 0; mt < NMT; ++mt) {
        if (2 * mt + (g >> 1) >= cout) continue;
        float* o = out + (size_t)(2 * mt + (g >> 1)) * 2 * fsz + (size_t)p * fsz +
                   (size_t)site * 24 * Ls + s;
#pragma unroll
        for (int c = 0; c < 3; ++c)
#pragma unroll
          for (int r = 0; r < 4; ++r) {
            const int alo = 2 * (g & 1) + (r >> 1);
            o[(size_t)((r & 1) + 2 * (c + 3 * alo)) * Ls] = acc[mt][c][r];
          }
      }
    }
  }
}

// Weight gradient of one layer: block (x, y = q * cin + b, z) adds, over its
// share of the lattice, G[q][a][b][al][be] = sum_{x, s, c}
// g[a](al, c) conj(T_q(in[b])(be, c)) for the AG output channels
// a = z AG .. z AG + AG - 1 (the transport formed once for all of them);
// block partials [y][a][block][32].
template <int AG>
__global__ void __launch_bounds__(256)
k_ptc_wgrad(double* __restrict__ partial, const float* __restrict__ g,
            const float* __restrict__ in, int cin, int cout, const float* __restrict__ links,
            const int* __restrict__ nbr, int Vh, int Ls) {
  const int q = blockIdx.y / cin, b = blockIdx.y - (blockIdx.y / cin) * cin;
  const int a0 = blockIdx.z * AG;
  const size_t fsz = (size_t)Vh * 24 * Ls;
  const size_t total = (size_t)2 * Vh * Ls;
  float acc[AG][32];
#pragma unroll
  for (int j = 0; j < AG; ++j)
#pragma unroll
    for (int k = 0; k < 32; ++k) acc[j][k] = 0;
  for (size_t idx = blockIdx.x * (size_t)blockDim.x + threadIdx.x; idx < total;
       idx += (size_t)gridDim.x * blockDim.x) {
    const int s = (int)(idx % Ls);
    const size_t r = idx / Ls;
    const int site = (int)(r % Vh);
    const int p = (int)(r / Vh);
    float t[24];
    if (!ptc_transport(t, in + (size_t)b * 2 * fsz, links, nbr, q, p, site, s, Vh, Ls)) continue;
#pragma unroll
    for (int j = 0; j < AG; ++j) {
      const float* gp =
          g + (size_t)(a0 + j) * 2 * fsz + (size_t)p * fsz + (size_t)site * 24 * Ls + s;
      float gv[24];
#pragma unroll
      for (int k = 0; k < 24; ++k) gv[k] = gp[k * Ls];
#pragma unroll
      for (int al = 0; al < 4; ++al)
#pragma unroll
        for (int be = 0; be < 4; ++be)
#pragma unroll
          for (int c = 0; c < 3; ++c)
            // g conj(t)
            cmac(acc[j][2 * (al * 4 + be)], acc[j][2 * (al * 4 + be) + 1], gv[2 * (c + 3 * al)],
                 gv[2 * (c + 3 * al) + 1], t[2 * (c + 3 * be)], -t[2 * (c + 3 * be) + 1]);
    }
  }
  __shared__ double sh[4][AG * 32];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll
  for (int j = 0; j < AG; ++j)
#pragma unroll
    for (int k = 0; k < 32; ++k) {
      double v = acc[j][k];
#pragma unroll
      for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
      if (lane == 0) sh[wv][j * 32 + k] = v;
    }
  __syncthreads();
  if (threadIdx.x < AG * 32) {
    const int j = threadIdx.x >> 5, k = threadIdx.x & 31;
    const double v = sh[0][threadIdx.x] + sh[1][threadIdx.x] + sh[2][threadIdx.x] +
                     sh[3][threadIdx.x];
    partial[(((size_t)blockIdx.y * cout + a0 + j) * gridDim.x + blockIdx.x) * 32 + k] = v;
  }
}

// out[(y z) 32 + k] = sum over the nblk block partials, fixed order
__global__ void __launch_bounds__(64)
k_ptc_wgrad_finish(double* __restrict__ out, const double* __restrict__ partial, int nblk,
                   int nyz) {
  const int yz = blockIdx.x;
  const int k = threadIdx.x;
  if (yz >= nyz || k >= 32) return;
  double v = 0;
  for (int i = 0; i < nblk; ++i) v += partial[((size_t)yz * nblk + i) * 32 + k];
  out[(size_t)yz * 32 + k] = v;
}

// even part of channel 0 <- x (FP32 eo field), odd part <- 0
__global__ void __launch_bounds__(256)
k_ptc_embed(float* __restrict__ act, const float* __restrict__ x, size_t n) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x) {
    act[i] = x[i];
    act[n + i] = 0.f;
  }
}

// g = 2 (y - t) on the even part of channel 0 (y = act), 0 on the odd part;
// block partials of |y - t|^2
__global__ void __launch_bounds__(256)
k_ptc_residual(float* __restrict__ g, const float* __restrict__ act, const float* __restrict__ t,
               size_t n, double* __restrict__ partial) {
  double acc = 0;
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x) {
    const float d = act[i] - t[i];
    acc += (double)d * (double)d;
    g[i] = 2.f * d;
    g[n + i] = 0.f;
  }
  __shared__ double sh[4];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) acc += __shfl_down(acc, off, 64);
  if (lane == 0) sh[wv] = acc;
  __syncthreads();
  if (threadIdx.x == 0) partial[blockIdx.x] = sh[0] + sh[1] + sh[2] + sh[3];
}

static unsigned ptc_grid(size_t n) {
  size_t g = (n + 255) / 256;
  if (g > 4096) g = 4096;
  return (unsigned)(g < 1 ? 1 : g);
}

template <bool ADJ, int AG>
static void ptc_layer_launch(float* out, const float* in, int cin, int cout, const unsigned* W,
                             int wcin, int wcout, const float* links, const int* nbr, int Vh,
                             int Ls, hipStream_t st) {
  const dim3 grid(ptc_grid((size_t)2 * Vh * Ls), cout / AG);
  hipLaunchKernelGGL((k_ptc_layer<ADJ, AG>), grid, dim3(256), 0, st, out, in, cin, W, wcin, wcout,
                     links, nbr, Vh, Ls);
}

template <bool ADJ, int NMT>
static void ptc_layer_mfma_launch(float* out, const float* in, int cin, int cout,
                                  const unsigned* W, int wcin, int wcout, const float* links,
                                  const int* nbr, int Vh, int Ls, hipStream_t st) {
  const size_t tiles = 2 * (((size_t)Vh * Ls + 15) / 16);
  size_t blocks = (tiles + 3) / 4;
  if (blocks > 8192) blocks = 8192;
  hipLaunchKernelGGL((k_ptc_layer_mfma<ADJ, NMT>), dim3((unsigned)blocks), dim3(256), 0, st, out,
                     in, cin, cout, W, wcin, wcout, links, nbr, Vh, Ls);
}

template <bool ADJ>
static void ptc_layer_dispatch(float* out, const float* in, int cin, int cout, const unsigned* W,
                               int wcin, int wcout, const float* links, const int* nbr, int Vh,
                               int Ls, hipStream_t st) {
  // the matrix-core form (all output channels of 16 points per wave, an odd
  // count padded with a zero channel) wherever its 32-bit buffer offsets fit:
  // one channel's two parities, the links and the weights each below 2^31 bytes
  // DWF_PTC_KERNEL=valu selects the VALU form everywhere (its tests use it)
  const size_t fsz = (size_t)Vh * 24 * Ls;
  const bool fits = 8 * fsz < (1ull << 31) && (size_t)2 * 8 * 18 * 4 * Vh < (1ull << 31);
  const char* force = std::getenv("DWF_PTC_KERNEL");
  if (fits && !(force && std::strcmp(force, "valu") == 0)) {
    switch ((cout + 1) / 2) {
#define PTC_MFMA(n)                                                                          \
  case n:                                                                                    \
    return ptc_layer_mfma_launch<ADJ, n>(out, in, cin, cout, W, wcin, wcout, links, nbr, Vh, \
                                         Ls, st);
      PTC_MFMA(1) PTC_MFMA(2) PTC_MFMA(3) PTC_MFMA(4)
      PTC_MFMA(5) PTC_MFMA(6) PTC_MFMA(7) PTC_MFMA(8)
#undef PTC_MFMA
      default: break;
    }
  }
  // otherwise the VALU form: four output channels per thread (227 registers, two
  // waves per SIMD); eight need 324 registers, one wave per SIMD, and are slower;
  // a thread per point and group, with 64-bit addressing
  if (cout % 4 == 0)
    ptc_layer_launch<ADJ, 4>(out, in, cin, cout, W, wcin, wcout, links, nbr, Vh, Ls, st);
  else if (cout % 2 == 0)
    ptc_layer_launch<ADJ, 2>(out, in, cin, cout, W, wcin, wcout, links, nbr, Vh, Ls, st);
  else
    ptc_layer_launch<ADJ, 1>(out, in, cin, cout, W, wcin, wcout, links, nbr, Vh, Ls, st);
}

hipError_t launch_ptc_layer(bool adj, float* out, const float* in, int cin, int cout,
                            const unsigned short* W, int wcin, int wcout, const float* links,
                            const int* nbr, int Vh, int Ls, hipStream_t st) {
  // each complex weight is one aligned 32-bit word (re, im)
  if (reinterpret_cast<uintptr_t>(W) % 4 != 0 || cout < 1) return hipErrorInvalidValue;
  const unsigned* w = reinterpret_cast<const unsigned*>(W);
  if (adj)
    ptc_layer_dispatch<true>(out, in, cin, cout, w, wcin, wcout, links, nbr, Vh, Ls, st);
  else
    ptc_layer_dispatch<false>(out, in, cin, cout, w, wcin, wcout, links, nbr, Vh, Ls, st);
  return hipGetLastError();
}

hipError_t launch_ptc_wgrad(double* G, double* partial, int nblk, const float* g, const float* in,
                            int cin, int cout, const float* links, const int* nbr, int Vh, int Ls,
                            hipStream_t st) {
  if (cout < 1) return hipErrorInvalidValue;
  // two output channels per thread: four need 128 accumulators (266 registers,
  // one wave per SIMD), which is slower
  const int ag = cout % 2 == 0 ? 2 : 1;
  const dim3 grid(nblk, kPtcPaths * cin, cout / ag);
  if (ag == 2)
    hipLaunchKernelGGL(k_ptc_wgrad<2>, grid, dim3(256), 0, st, partial, g, in, cin, cout, links,
                       nbr, Vh, Ls);
  else
    hipLaunchKernelGGL(k_ptc_wgrad<1>, grid, dim3(256), 0, st, partial, g, in, cin, cout, links,
                       nbr, Vh, Ls);
  const int nyz = kPtcPaths * cin * cout;
  hipLaunchKernelGGL(k_ptc_wgrad_finish, dim3(nyz), dim3(64), 0, st, G, partial, nblk, nyz);
  return hipGetLastError();
}

hipError_t launch_ptc_embed(float* act, const float* x, size_t n, hipStream_t st) {
  hipLaunchKernelGGL(k_ptc_embed, dim3(ptc_grid(n)), dim3(256), 0, st, act, x, n);
  return hipGetLastError();
}

hipError_t launch_ptc_residual(float* g, const float* act, const float* t, size_t n,
                               double* partial, unsigned* nblk, hipStream_t st) {
  unsigned gr = ptc_grid(n);
  if (gr > (unsigned)kRedBlocks) gr = kRedBlocks;
  *nblk = gr;
  hipLaunchKernelGGL(k_ptc_residual, dim3(gr), dim3(256), 0, st, g, act, t, n, partial);
  return hipGetLastError();
}

}  // namespace dwf
