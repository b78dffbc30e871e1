// dwf_lpc.hip -- kernels of the learned polynomial preconditioner.
//
// Training fits per-chirality Ls x Ls weights by least squares, which needs
// the chiral Gram matrices of a few 5D fields f_0..f_{F-1}:
//   G_ij[chi][s][t] = sum_{site, k < 12} p_chi(f_i)(site, k, s) p_chi(f_j)(site, k, t),
//   p_+(f)_k = f_k + f_{k+12},  p_-(f)_k = f_k - f_{k+12}   (Dirac basis, gamma_5
// exchanges the upper and lower spin pairs).  Over the (site, k) samples this is
// a tall-skinny GEMM (Ls x 12 Vh) (12 Vh x Ls): for FP32 fields with Ls <= 16
// it runs on the matrix cores (v_mfma_f32_16x16x4f32, one 16 x 16 tile per
// chirality and field pair); otherwise on the vector ALUs.  Every field is
// read once for all pairs.  Partial sums go per block to a buffer and are
// added in a fixed order (deterministic).
#include "dwf_device.h"
#include "dwf_kernels.h"

namespace dwf {

// ---- uniform random fill: x_i in [-1, 1) from a counter hash (splitmix64) ----
__device__ __forceinline__ double hash_uniform(unsigned long long seed, unsigned long long i) {
  unsigned long long z = seed + (i + 1ull) * 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  z ^= z >> 31;
  return (double)(z >> 11) * (1.0 / 4503599627370496.0) - 1.0;   // 2^-52
}

template <typename T>
__global__ void k_fill_uniform(T* __restrict__ x, size_t n, unsigned long long seed) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x)
    St<T>::st(x + i, (typename St<T>::C)hash_uniform(seed, i));
}

hipError_t launch_fill_uniform(int prec, void* x, size_t n, unsigned long long seed,
                               hipStream_t st) {
  size_t g = (n + 255) / 256;
  if (g > 4096) g = 4096;
  const dim3 grid((unsigned)g);
  switch (prec) {
    case P64: hipLaunchKernelGGL(k_fill_uniform<double>, grid, dim3(256), 0, st, (double*)x, n, seed); break;
    case P32: hipLaunchKernelGGL(k_fill_uniform<float>, grid, dim3(256), 0, st, (float*)x, n, seed); break;
    case P16: hipLaunchKernelGGL(k_fill_uniform<__half>, grid, dim3(256), 0, st, (__half*)x, n, seed); break;
    case PBF16:
      hipLaunchKernelGGL(k_fill_uniform<__hip_bfloat16>, grid, dim3(256), 0, st,
                         (__hip_bfloat16*)x, n, seed);
      break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

// ---- chiral Gram matrices --------------------------------------------------
constexpr int kGramMaxF = 6;
constexpr int kGramMaxPairs = kGramMaxF * (kGramMaxF + 1) / 2;
constexpr int kGramBlocks = 512;

struct GramFields {
  const void* f[kGramMaxF];
};

__host__ __device__ constexpr int gram_pairs(int F) { return F * (F + 1) / 2; }

// Matrix-core path: FP32 fields, Ls <= 16.  One wave takes 4 (site, k)
// samples per step; lane l holds s = l % 16 of sample l / 16, which is the
// A (16 x 4, i = s) and B (4 x 16, j = t) operand layout of
// v_mfma_f32_16x16x4f32; the result D[i][j] sits in lane j + 16 (i / 4),
// register i % 4.
typedef float v4f __attribute__((ext_vector_type(4)));

template <int F>
__global__ void __launch_bounds__(256)
k_gram_mfma(GramFields fs, int Vh, int Ls, double* __restrict__ partial) {
  constexpr int NP = F * (F + 1) / 2;
  __shared__ float shw[NP * 2 * 256];   // the block's tiles, summed wave by wave
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int s = lane & 15, kk = lane >> 4;
  v4f acc[NP][2];
#pragma unroll
  for (int p = 0; p < NP; ++p) {
    acc[p][0] = v4f{0.f, 0.f, 0.f, 0.f};
    acc[p][1] = v4f{0.f, 0.f, 0.f, 0.f};
  }
  const size_t groups = (size_t)Vh * 3;   // 12 Vh samples, 4 per step
  const size_t nw = (size_t)gridDim.x * 4;
  for (size_t g = (size_t)blockIdx.x * 4 + wave; g < groups; g += nw) {
    const size_t n = 4 * g + kk;
    const size_t site = n / 12;
    const int k = (int)(n - site * 12);
    const size_t o1 = (site * 24 + k) * Ls + s, o2 = o1 + 12 * (size_t)Ls;
    float u[F], l[F];
#pragma unroll
    for (int i = 0; i < F; ++i) {
      const float* fp = static_cast<const float*>(fs.f[i]);
      const float a = (s < Ls) ? fp[o1] : 0.f;
      const float b = (s < Ls) ? fp[o2] : 0.f;
      u[i] = a + b;
      l[i] = a - b;
    }
    int p = 0;
#pragma unroll
    for (int i = 0; i < F; ++i)
#pragma unroll
      for (int j = i; j < F; ++j, ++p) {
        acc[p][0] = __builtin_amdgcn_mfma_f32_16x16x4f32(u[i], u[j], acc[p][0], 0, 0, 0);
        acc[p][1] = __builtin_amdgcn_mfma_f32_16x16x4f32(l[i], l[j], acc[p][1], 0, 0, 0);
      }
  }
  // tiles to LDS, one wave after the other: element (p, chi, i, j) with
  // i = 4 (lane / 16) + r, j = lane % 16
  for (int w = 0; w < 4; ++w) {
    if (wave == w) {
#pragma unroll
      for (int p = 0; p < NP; ++p)
#pragma unroll
        for (int c = 0; c < 2; ++c)
#pragma unroll
          for (int r = 0; r < 4; ++r) {
            const int i = 4 * (lane >> 4) + r, j = lane & 15;
            float* d = shw + (p * 2 + c) * 256 + i * 16 + j;
            *d = (w == 0) ? acc[p][c][r] : *d + acc[p][c][r];
          }
    }
    __syncthreads();
  }
  for (int e = threadIdx.x; e < NP * 2 * 256; e += blockDim.x)
    partial[(size_t)blockIdx.x * NP * 2 * 256 + e] = (double)shw[e];
}

// Vector path, any storage type and Ls <= 16: thread (s, t) of a 16 x 16 block
// accumulates all pairs and both chiralities over the block's samples, which
// are staged through LDS in chunks.
template <typename T, int F>
__global__ void __launch_bounds__(256)
k_gram_valu(GramFields fs, int Vh, int Ls, double* __restrict__ partial) {
  constexpr int NP = F * (F + 1) / 2;
  constexpr int CH = 16;   // samples per chunk
  using C = typename St<T>::C;
  __shared__ double shu[F][CH][16], shl[F][CH][16];
  const int t = threadIdx.x & 15, s = threadIdx.x >> 4;
  double au[NP], al[NP];
#pragma unroll
  for (int p = 0; p < NP; ++p) { au[p] = 0; al[p] = 0; }
  const size_t S = (size_t)Vh * 12;
  const size_t nchunks = (S + CH - 1) / CH;
  for (size_t ch = blockIdx.x; ch < nchunks; ch += gridDim.x) {
    __syncthreads();
    // stage: thread (s, t) loads sample t of the chunk at index s (all fields)
    {
      const size_t n = ch * CH + (size_t)(threadIdx.x >> 4);
      const int si = threadIdx.x & 15;
#pragma unroll
      for (int i = 0; i < F; ++i) {
        double a = 0, b = 0;
        if (n < S && si < Ls) {
          const size_t site = n / 12;
          const int k = (int)(n - site * 12);
          const T* fp = static_cast<const T*>(fs.f[i]);
          a = (double)St<T>::ld(fp + (site * 24 + k) * Ls + si);
          b = (double)St<T>::ld(fp + (site * 24 + k + 12) * Ls + si);
        }
        shu[i][threadIdx.x >> 4][si] = a + b;
        shl[i][threadIdx.x >> 4][si] = a - b;
      }
    }
    __syncthreads();
#pragma unroll
    for (int m = 0; m < CH; ++m) {
      int p = 0;
#pragma unroll
      for (int i = 0; i < F; ++i)
#pragma unroll
        for (int j = i; j < F; ++j, ++p) {
          au[p] += shu[i][m][s] * shu[j][m][t];
          al[p] += shl[i][m][s] * shl[j][m][t];
        }
    }
  }
  (void)sizeof(C);
#pragma unroll
  for (int p = 0; p < NP; ++p) {
    partial[(size_t)blockIdx.x * NP * 2 * 256 + (p * 2 + 0) * 256 + s * 16 + t] = au[p];
    partial[(size_t)blockIdx.x * NP * 2 * 256 + (p * 2 + 1) * 256 + s * 16 + t] = al[p];
  }
}

// out[p][chi][s][t] (Ls x Ls, double) = sum over blocks of the 16 x 16 tiles
__global__ void k_gram_finish(const double* __restrict__ partial, int nblocks, int NP, int Ls,
                              double* __restrict__ out) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  const int per = 2 * Ls * Ls;
  if (e >= NP * per) return;
  const int p = e / per, r = e - p * per;
  const int c = r / (Ls * Ls), q = r - c * Ls * Ls;
  const int s = q / Ls, t = q - s * Ls;
  const size_t src = (size_t)(p * 2 + c) * 256 + s * 16 + t;
  double v = 0;
  for (int b = 0; b < nblocks; ++b) v += partial[(size_t)b * NP * 2 * 256 + src];
  out[e] = v;
}

size_t gram_partial_doubles() { return (size_t)kGramBlocks * kGramMaxPairs * 2 * 256; }

template <int F>
static hipError_t gram_launch_f(int prec, const GramFields& fs, int Vh, int Ls, double* partial,
                                int nb, hipStream_t st) {
  if (prec == P32 && Ls <= 16) {
    hipLaunchKernelGGL(k_gram_mfma<F>, dim3(nb), dim3(256), 0, st, fs, Vh, Ls, partial);
    return hipGetLastError();
  }
  switch (prec) {
    case P64: hipLaunchKernelGGL((k_gram_valu<double, F>), dim3(nb), dim3(256), 0, st, fs, Vh, Ls, partial); break;
    case P32: hipLaunchKernelGGL((k_gram_valu<float, F>), dim3(nb), dim3(256), 0, st, fs, Vh, Ls, partial); break;
    case P16: hipLaunchKernelGGL((k_gram_valu<__half, F>), dim3(nb), dim3(256), 0, st, fs, Vh, Ls, partial); break;
    case PBF16:
      hipLaunchKernelGGL((k_gram_valu<__hip_bfloat16, F>), dim3(nb), dim3(256), 0, st, fs, Vh, Ls,
                         partial);
      break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

// All pairs (i <= j, row-major over i) of F fields; out: F(F+1)/2 x 2 x Ls x Ls doubles.
hipError_t launch_chiral_gram(int prec, const void* const* fields, int F, int Vh, int Ls,
                              double* partial, double* out, hipStream_t st) {
  if (F < 1 || F > kGramMaxF || Ls < 1 || Ls > 16) return hipErrorInvalidValue;
  GramFields fs{};
  for (int i = 0; i < F; ++i) fs.f[i] = fields[i];
  const size_t groups = (size_t)Vh * 3;
  int nb = (int)((groups + 3) / 4);
  if (nb > kGramBlocks) nb = kGramBlocks;
  if (nb < 1) nb = 1;
  hipError_t e = hipErrorInvalidValue;
  switch (F) {
    case 1: e = gram_launch_f<1>(prec, fs, Vh, Ls, partial, nb, st); break;
    case 2: e = gram_launch_f<2>(prec, fs, Vh, Ls, partial, nb, st); break;
    case 3: e = gram_launch_f<3>(prec, fs, Vh, Ls, partial, nb, st); break;
    case 4: e = gram_launch_f<4>(prec, fs, Vh, Ls, partial, nb, st); break;
    case 5: e = gram_launch_f<5>(prec, fs, Vh, Ls, partial, nb, st); break;
    case 6: e = gram_launch_f<6>(prec, fs, Vh, Ls, partial, nb, st); break;
  }
  if (e != hipSuccess) return e;
  const int NP = gram_pairs(F);
  const int total = NP * 2 * Ls * Ls;
  hipLaunchKernelGGL(k_gram_finish, dim3((total + 255) / 256), dim3(256), 0, st,
                     (const double*)partial, nb, NP, Ls, out);
  return hipGetLastError();
}

}  // namespace dwf
