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
#include "dwf_device.h"
#include "dwf_kernels.h"

namespace dwf {

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
      h1[r] = C(2) * w2[r]; h1[i] = C(2) * w2[i];
      h2[r] = C(2) * w3[r]; h2[i] = C(2) * w3[i];
    } else {                // h1 = 2 w0, h2 = 2 w1
      h1[r] = C(2) * w0[r]; h1[i] = C(2) * w0[i];
      h2[r] = C(2) * w1[r]; h2[i] = C(2) * w1[i];
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

template <int DIR, bool DAG, typename T, typename C>
__device__ __forceinline__ void hop_dir(C* acc, const T* __restrict__ in,
                                        const T* __restrict__ L, const int* __restrict__ nbr,
                                        int site, int ls, int Vh, int Ls) {
  using S = St<T>;
  const int nb = nbr[DIR * Vh + site];
  const T* wp = in + (size_t)nb * 24 * Ls + ls;
  C w[24];
  if (!DAG) {
#pragma unroll
    for (int k = 0; k < 24; ++k) w[k] = S::ld(wp + k * Ls);
  } else {  // gamma_5 (Dirac basis) exchanges the upper and lower spin pairs
#pragma unroll
    for (int k = 0; k < 12; ++k) {
      w[k] = S::ld(wp + (k + 12) * Ls);
      w[k + 12] = S::ld(wp + k * Ls);
    }
  }
  C h1[6], h2[6];
  project<DIR>(w, h1, h2);
  const T* up = L + (size_t)DIR * 18 * Vh + site;
  C u[18];
#pragma unroll
  for (int e = 0; e < 18; ++e) u[e] = S::ld(up + (size_t)e * Vh);
  C a1[6], a2[6];
  su3_mul(u, h1, a1);
  su3_mul(u, h2, a2);
  reconstruct<DIR>(acc, a1, a2);
}

template <typename T, int OP, bool MAT, bool ADDX, bool OUT2>
__global__ void __launch_bounds__(256)
k_dwf(T* __restrict__ out, const T* __restrict__ in, const T* __restrict__ links,
      const int* __restrict__ nbr, int parity, int Vh, int Ls, int spb,
      const typename St<T>::C* __restrict__ E, typename St<T>::C scale,
      const T* __restrict__ X, typename St<T>::C xcoef, T* __restrict__ out2,
      const typename St<T>::C* __restrict__ F) {
  using S = St<T>;
  using C = typename S::C;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  C* sh_u = reinterpret_cast<C*>(smem_raw);
  C* sh_l = sh_u + spb * 12 * Ls;
  C* sh_E = sh_l + spb * 12 * Ls;
  C* sh_F = sh_E + 2 * Ls * Ls;

  const int tid = threadIdx.x;
  const int lsite = tid / Ls;
  const int ls = tid - lsite * Ls;
  const int site = blockIdx.x * spb + lsite;
  const bool valid = (lsite < spb) && (site < Vh);

  if (MAT) {
    const int LL = Ls * Ls;
    for (int i = tid; i < 2 * LL; i += blockDim.x) {
      const int chi = i / LL, r = i - chi * LL;
      const int so = r / Ls, si = r - so * Ls;
      sh_E[chi * LL + si * Ls + so] = E[i];
      if (OUT2) sh_F[chi * LL + si * Ls + so] = F[i];
    }
  }

  C acc[24];
#pragma unroll
  for (int k = 0; k < 24; ++k) acc[k] = 0;
  if (valid) {
    if (OP == K_SITE) {
      const T* wp = in + (size_t)site * 24 * Ls + ls;
#pragma unroll
      for (int k = 0; k < 24; ++k) acc[k] = S::ld(wp + k * Ls);
    } else {
      const bool dag = (OP == K_HOPDAG);
      const T* L = links + (size_t)parity * 8 * 18 * Vh;
      const int* nb = nbr + (size_t)parity * 8 * Vh;
      if (!dag) {
        hop_dir<0, false>(acc, in, L, nb, site, ls, Vh, Ls);
        hop_dir<1, false>(acc, in, L, nb, site, ls, Vh, Ls);
        hop_dir<2, false>(acc, in, L, nb, site, ls, Vh, Ls);
        hop_dir<3, false>(acc, in, L, nb, site, ls, Vh, Ls);
        hop_dir<4, false>(acc, in, L, nb, site, ls, Vh, Ls);
        hop_dir<5, false>(acc, in, L, nb, site, ls, Vh, Ls);
        hop_dir<6, false>(acc, in, L, nb, site, ls, Vh, Ls);
        hop_dir<7, false>(acc, in, L, nb, site, ls, Vh, Ls);
      } else {
        hop_dir<0, true>(acc, in, L, nb, site, ls, Vh, Ls);
        hop_dir<1, true>(acc, in, L, nb, site, ls, Vh, Ls);
        hop_dir<2, true>(acc, in, L, nb, site, ls, Vh, Ls);
        hop_dir<3, true>(acc, in, L, nb, site, ls, Vh, Ls);
        hop_dir<4, true>(acc, in, L, nb, site, ls, Vh, Ls);
        hop_dir<5, true>(acc, in, L, nb, site, ls, Vh, Ls);
        hop_dir<6, true>(acc, in, L, nb, site, ls, Vh, Ls);
        hop_dir<7, true>(acc, in, L, nb, site, ls, Vh, Ls);
#pragma unroll
        for (int k = 0; k < 12; ++k) {
          C t = acc[k];
          acc[k] = acc[k + 12];
          acc[k + 12] = t;
        }
      }
    }
  }

  const size_t base = (size_t)site * 24 * Ls + ls;
  if (!MAT) {
    if (valid) {
#pragma unroll
      for (int k = 0; k < 24; ++k) {
        C v = scale * acc[k];
        if (ADDX) v += xcoef * S::ld(X + base + k * Ls);
        S::st(out + base + k * Ls, v);
      }
    }
    return;
  }

  // chiral components of the core, per (site, k < 12, s): P+ ~ u, P- ~ l
  if (valid) {
    C* u = sh_u + lsite * 12 * Ls + ls;
    C* l = sh_l + lsite * 12 * Ls + ls;
#pragma unroll
    for (int k = 0; k < 12; ++k) {
      u[k * Ls] = acc[k] + acc[k + 12];
      l[k * Ls] = acc[k] - acc[k + 12];
    }
  }
  __syncthreads();
  C o[24];
  if (valid) {
    C ou[12], ol[12];
#pragma unroll
    for (int k = 0; k < 12; ++k) { ou[k] = 0; ol[k] = 0; }
    const C* u = sh_u + lsite * 12 * Ls;
    const C* l = sh_l + lsite * 12 * Ls;
    const int LL = Ls * Ls;
    for (int s = 0; s < Ls; ++s) {
      const C ep = sh_E[s * Ls + ls];
      const C em = sh_E[LL + s * Ls + ls];
#pragma unroll
      for (int k = 0; k < 12; ++k) {
        ou[k] += ep * u[k * Ls + s];
        ol[k] += em * l[k * Ls + s];
      }
    }
#pragma unroll
    for (int k = 0; k < 12; ++k) {
      o[k] = ou[k] + ol[k];
      o[k + 12] = ou[k] - ol[k];
    }
    if (ADDX) {
#pragma unroll
      for (int k = 0; k < 24; ++k) o[k] += xcoef * S::ld(X + base + k * Ls);
    }
#pragma unroll
    for (int k = 0; k < 24; ++k) S::st(out + base + k * Ls, o[k]);
  }
  if (OUT2) {
    __syncthreads();
    if (valid) {
      C* u = sh_u + lsite * 12 * Ls + ls;
      C* l = sh_l + lsite * 12 * Ls + ls;
#pragma unroll
      for (int k = 0; k < 12; ++k) {
        u[k * Ls] = o[k] + o[k + 12];
        l[k * Ls] = o[k] - o[k + 12];
      }
    }
    __syncthreads();
    if (valid) {
      C ou[12], ol[12];
#pragma unroll
      for (int k = 0; k < 12; ++k) { ou[k] = 0; ol[k] = 0; }
      const C* u = sh_u + lsite * 12 * Ls;
      const C* l = sh_l + lsite * 12 * Ls;
      const int LL = Ls * Ls;
      for (int s = 0; s < Ls; ++s) {
        const C fp = sh_F[s * Ls + ls];
        const C fm = sh_F[LL + s * Ls + ls];
#pragma unroll
        for (int k = 0; k < 12; ++k) {
          ou[k] += fp * u[k * Ls + s];
          ol[k] += fm * l[k * Ls + s];
        }
      }
#pragma unroll
      for (int k = 0; k < 12; ++k) {
        S::st(out2 + base + k * Ls, ou[k] + ol[k]);
        S::st(out2 + base + (k + 12) * Ls, ou[k] - ol[k]);
      }
    }
  }
}

template <typename T>
static hipError_t launch_dwf_t(int op, void* out, const void* in, const void* links,
                               const int* nbr, int parity, int Vh, int Ls, const Epi& e,
                               hipStream_t st) {
  using C = typename St<T>::C;
  const bool mat = e.E != nullptr;
  const bool addx = e.X != nullptr;
  const bool out2 = e.out2 != nullptr;
  if (out2 && !mat) return hipErrorInvalidValue;
  int spb = 256 / Ls;
  if (spb < 1) return hipErrorInvalidValue;
  size_t matb = (mat ? 2 : 0) * (size_t)Ls * Ls * sizeof(C) * (out2 ? 2 : 1);
  const size_t lds_cap = 64 * 1024;
  if (mat) {
    if (matb >= lds_cap) return hipErrorInvalidValue;
    int cap = (int)((lds_cap - matb) / (24 * (size_t)Ls * sizeof(C)));
    if (cap < spb) spb = cap;
    if (spb < 1) return hipErrorInvalidValue;
  }
  const size_t lds = mat ? (24 * (size_t)spb * Ls * sizeof(C) + matb) : 0;
  dim3 block(spb * Ls), grid((Vh + spb - 1) / spb);
  T* o = (T*)out;
  const T* i = (const T*)in;
  const T* L = (const T*)links;
  const C* E = (const C*)e.E;
  const C* F = (const C*)e.F;
  const T* X = (const T*)e.X;
  T* o2 = (T*)e.out2;
  const C sc = (C)e.scale, xc = (C)e.xcoef;
#define DWF_LAUNCH(OPV, M, A, O2)                                                          \
  hipLaunchKernelGGL((k_dwf<T, OPV, M, A, O2>), grid, block, lds, st, o, i, L, nbr, parity, \
                     Vh, Ls, spb, E, sc, X, xc, o2, F)
#define DWF_BY_OP(M, A, O2)                                \
  do {                                                     \
    if (op == K_SITE) DWF_LAUNCH(K_SITE, M, A, O2);        \
    else if (op == K_HOP) DWF_LAUNCH(K_HOP, M, A, O2);     \
    else DWF_LAUNCH(K_HOPDAG, M, A, O2);                   \
  } while (0)
  if (!mat) {
    if (addx) DWF_BY_OP(false, true, false);
    else DWF_BY_OP(false, false, false);
  } else if (!out2) {
    if (addx) DWF_BY_OP(true, true, false);
    else DWF_BY_OP(true, false, false);
  } else {
    if (addx) DWF_BY_OP(true, true, true);
    else DWF_BY_OP(true, false, true);
  }
#undef DWF_BY_OP
#undef DWF_LAUNCH
  return hipGetLastError();
}

hipError_t launch_dwf(int prec, int op, void* out, const void* in, const void* links,
                      const int* nbr, int parity, int Vh, int Ls, const Epi& e,
                      hipStream_t st) {
  switch (prec) {
    case P64: return launch_dwf_t<double>(op, out, in, links, nbr, parity, Vh, Ls, e, st);
    case P32: return launch_dwf_t<float>(op, out, in, links, nbr, parity, Vh, Ls, e, st);
    case P16: return launch_dwf_t<__half>(op, out, in, links, nbr, parity, Vh, Ls, e, st);
    case PBF16:
      return launch_dwf_t<__hip_bfloat16>(op, out, in, links, nbr, parity, Vh, Ls, e, st);
  }
  return hipErrorInvalidValue;
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
// BLAS-1
// ---------------------------------------------------------------------------
template <typename T>
__global__ void k_set_zero(T* __restrict__ x, size_t n) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x)
    St<T>::st(x + i, 0);
}
template <typename T>
__global__ void k_copy(T* __restrict__ y, const T* __restrict__ x, size_t n) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x)
    y[i] = x[i];
}
template <typename T>
__global__ void k_axpy(T* __restrict__ y, typename St<T>::C a, const T* __restrict__ x, size_t n) {
  using S = St<T>;
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x)
    S::st(y + i, S::ld(y + i) + a * S::ld(x + i));
}
template <typename T>
__global__ void k_aypx(typename St<T>::C a, T* __restrict__ y, const T* __restrict__ x, size_t n) {
  using S = St<T>;
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x)
    S::st(y + i, a * S::ld(y + i) + S::ld(x + i));
}
template <typename T>
__global__ void k_scal(T* __restrict__ x, typename St<T>::C a, size_t n) {
  using S = St<T>;
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x)
    S::st(x + i, a * S::ld(x + i));
}
template <typename TO, typename TI>
__global__ void k_convert(TO* __restrict__ y, const TI* __restrict__ x, size_t n) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x)
    St<TO>::st(y + i, (typename St<TO>::C)St<TI>::ld(x + i));
}

template <typename T>
static void setz_t(void* x, size_t n, hipStream_t st) {
  hipLaunchKernelGGL(k_set_zero<T>, grid_for(n), dim3(256), 0, st, (T*)x, n);
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
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  return v;
}

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
    __threadfence();
    unsigned prev = atomicAdd(counter, 1u);
    amlast = (prev == gridDim.x - 1);
  }
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

static dim3 red_grid(size_t n) {
  size_t g = (n + 255) / 256;
  if (g > (size_t)kRedBlocks) g = kRedBlocks;
  if (g < 1) g = 1;
  return dim3((unsigned)g);
}

template <typename T>
__global__ void __launch_bounds__(256)
k_norm2(const T* __restrict__ x, size_t n, double* partial, unsigned* counter, double* result) {
  double v[1] = {0};
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x) {
    double a = (double)St<T>::ld(x + i);
    v[0] += a * a;
  }
  block_finish<1>(v, partial, counter, result);
}
template <typename T>
__global__ void __launch_bounds__(256)
k_dot(const T* __restrict__ x, const T* __restrict__ y, size_t n, double* partial,
      unsigned* counter, double* result) {
  double v[1] = {0};
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x)
    v[0] += (double)St<T>::ld(x + i) * (double)St<T>::ld(y + i);
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
template <typename T>
__global__ void __launch_bounds__(256)
k_cg_xr(T* __restrict__ x, T* __restrict__ r, const T* __restrict__ p, const T* __restrict__ s,
        typename St<T>::C a, size_t n, double* partial, unsigned* counter, double* result) {
  using S = St<T>;
  using C = typename S::C;
  double v[1] = {0};
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x) {
    S::st(x + i, S::ld(x + i) + a * S::ld(p + i));
    C rn = S::ld(r + i) - a * S::ld(s + i);
    S::st(r + i, rn);
    double rs = (double)S::ld(r + i);
    v[0] += rs * rs;
  }
  block_finish<1>(v, partial, counter, result);
}
template <typename T>
__global__ void __launch_bounds__(256)
k_axpy_dot(T* __restrict__ y, typename St<T>::C a, const T* __restrict__ x, size_t n,
           double* partial, unsigned* counter, double* result) {
  using S = St<T>;
  double v[1] = {0};
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x) {
    const auto xv = S::ld(x + i);
    S::st(y + i, S::ld(y + i) + a * xv);
    v[0] += (double)S::ld(y + i) * (double)xv;
  }
  block_finish<1>(v, partial, counter, result);
}
template <typename T>
__global__ void __launch_bounds__(256)
k_sub_norm2(T* __restrict__ r, const T* __restrict__ ax, const T* __restrict__ b, size_t n,
            double* partial, unsigned* counter, double* result) {
  using S = St<T>;
  double v[1] = {0};
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x) {
    S::st(r + i, S::ld(ax + i) - S::ld(b + i));
    double t = (double)S::ld(r + i);
    v[0] += t * t;
  }
  block_finish<1>(v, partial, counter, result);
}

template <typename T>
static void norm2_t(const void* x, size_t n, const RedScratch& r, int slot, hipStream_t st) {
  hipLaunchKernelGGL(k_norm2<T>, red_grid(n), dim3(256), 0, st, (const T*)x, n, r.partial,
                     r.counter, r.result + slot);
}
template <typename T>
static void dot_t(const void* x, const void* y, size_t n, const RedScratch& r, int slot,
                  hipStream_t st) {
  hipLaunchKernelGGL(k_dot<T>, red_grid(n), dim3(256), 0, st, (const T*)x, (const T*)y, n,
                     r.partial, r.counter, r.result + slot);
}
template <typename T>
static void cg_xr_t(void* x, void* rr, const void* p, const void* s, double a, size_t n,
                    const RedScratch& r, int slot, hipStream_t st) {
  hipLaunchKernelGGL(k_cg_xr<T>, red_grid(n), dim3(256), 0, st, (T*)x, (T*)rr, (const T*)p,
                     (const T*)s, (typename St<T>::C)a, n, r.partial, r.counter, r.result + slot);
}
template <typename T>
static void axpy_dot_t(void* y, double a, const void* x, size_t n, const RedScratch& r, int slot,
                       hipStream_t st) {
  hipLaunchKernelGGL(k_axpy_dot<T>, red_grid(n), dim3(256), 0, st, (T*)y, (typename St<T>::C)a,
                     (const T*)x, n, r.partial, r.counter, r.result + slot);
}
template <typename T>
static void sub_norm2_t(void* rr, const void* ax, const void* b, size_t n, const RedScratch& r,
                        int slot, hipStream_t st) {
  hipLaunchKernelGGL(k_sub_norm2<T>, red_grid(n), dim3(256), 0, st, (T*)rr, (const T*)ax,
                     (const T*)b, n, r.partial, r.counter, r.result + slot);
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
  return hipGetLastError();
}
hipError_t launch_cg_xr(int prec, void* x, void* r, const void* p, const void* s, double a,
                        size_t n, const RedScratch& red, int slot, hipStream_t st) {
  DWF_DISPATCH(prec, cg_xr_t, (x, r, p, s, a, n, red, slot, st));
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
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x) {
    const double pv = pj[i];
    xj[i] += c0 * pv;
    const double np = c1 * pv + c2 * r[i];
    pj[i] = np;
    v[0] += np * np;
  }
  block_finish<1>(v, partial + (size_t)blockIdx.y * kRedBlocks, counters + blockIdx.y,
                  pp + j);
}

hipError_t launch_shift_update(void* const* x, void* const* p, const void* r,
                               const double* coef, int m, size_t n, double* partial,
                               unsigned* counters, double* pp, hipStream_t st) {
  if (m <= 1) return hipSuccess;
  dim3 g = red_grid(n);
  g.y = m - 1;
  hipLaunchKernelGGL(k_shift_update, g, dim3(256), 0, st, (double* const*)x, (double* const*)p,
                     (const double*)r, coef, n, partial, counters, pp);
  return hipGetLastError();
}

}  // namespace dwf
