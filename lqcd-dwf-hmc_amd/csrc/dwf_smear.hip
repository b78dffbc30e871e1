// dwf_smear.hip -- stout smearing and plaquette kernels (FP64).
//
// One smearing level (APE staples with stout projection):
//   C_mu(x) = rho sum_{nu != mu} [ U_nu(x) U_mu(x+nu) U_nu(x+mu)^+
//                                 + U_nu(x-nu)^+ U_mu(x-nu) U_nu(x-nu+mu) ]
//   iQ = at(C_mu U_mu^+)   (anti-hermitian traceless part)
//   U'_mu(x) = exp(iQ) U_mu(x),  exp by the Cayley-Hamilton form of
//   Morningstar and Peardon, Phys. Rev. D 69 (2004) 054501.
// Same steps as src/lib_alt_QXS/Smear/asmear_APE-tmpl.h:183-223 and
// aprojection_Stout_SU3-tmpl.h:477-553, 897-977; one thread per (site, mu).
#include "dwf_device.h"
#include "dwf_kernels.h"

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

__device__ M3 exp_iQ(const M3& iQ) {
  double n2 = 0;
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) n2 += iQ.a[i][j].r * iQ.a[i][j].r + iQ.a[i][j].i * iQ.a[i][j].i;
  M3 r;
  if (!(n2 > 1.0e-10)) {
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int j = 0; j < 3; ++j) r.a[i][j] = {i == j ? 1.0 : 0.0, 0.0};
    return r;
  }
  const M3 iQ2 = mul_nn(iQ, iQ);
  const M3 iQ3 = mul_nn(iQ, iQ2);
  // Q = -i iQ: c0 = det Q = -Im tr(iQ^3)/3, c1 = tr(Q^2)/2 = -Re tr(iQ^2)/2
  const double c0 = -(iQ3.a[0][0].i + iQ3.a[1][1].i + iQ3.a[2][2].i) / 3.0;
  const double c1 = -0.5 * (iQ2.a[0][0].r + iQ2.a[1][1].r + iQ2.a[2][2].r);
  const double c13r = sqrt(c1 / 3.0);
  const double c0max = 2.0 * c13r * c13r * c13r;
  double ratio = c0 / c0max;
  ratio = fmin(1.0, fmax(-1.0, ratio));
  const double theta = acos(ratio);
  const double u = c13r * cos(theta / 3.0);
  const double w = sqrt(c1) * sin(theta / 3.0);
  const double xi0 = (w == 0.0) ? 1.0 : sin(w) / w;
  const double u2 = u * u, w2 = w * w, cw = cos(w);
  const cplx emiu = {cos(u), -sin(u)};
  const cplx e2iu = {cos(2.0 * u), sin(2.0 * u)};
  const cplx h0 = cadd(cscal(u2 - w2, e2iu),
                       cmul(emiu, cplx{8.0 * u2 * cw, 2.0 * u * (3.0 * u2 + w2) * xi0}));
  const cplx t1 = cmul(emiu, cplx{2.0 * u * cw, -(3.0 * u2 - w2) * xi0});
  const cplx h1 = {2.0 * u * e2iu.r - t1.r, 2.0 * u * e2iu.i - t1.i};
  const cplx t2 = cmul(emiu, cplx{cw, 3.0 * u * xi0});
  const cplx h2 = {e2iu.r - t2.r, e2iu.i - t2.i};
  const double fden = 1.0 / (9.0 * u2 - w2);
  const cplx f0 = cscal(fden, h0), f1 = cscal(fden, h1), f2 = cscal(fden, h2);
  // exp(iQ) = f0 + f1 Q + f2 Q^2 with Q = -i iQ, Q^2 = -iQ^2
  const cplx f1mi = {f1.i, -f1.r};  // f1 * (-i)
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      cplx v = cmul(f1mi, iQ.a[i][j]);
      const cplx q2 = cmul(f2, iQ2.a[i][j]);
      v.r -= q2.r;
      v.i -= q2.i;
      if (i == j) v = cadd(v, f0);
      r.a[i][j] = v;
    }
  return r;
}

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

__global__ void __launch_bounds__(128)
k_stout(double* __restrict__ Uout, const double* __restrict__ Uin, int Nx, int Ny, int Nz,
        int Nt, double rho) {
  const size_t V = (size_t)Nx * Ny * Nz * Nt;
  const size_t idx = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  if (idx >= 4 * V) return;
  const int mu = (int)(idx / V);
  const size_t site = idx - (size_t)mu * V;
  M3 Cst;
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) Cst.a[i][j] = {0, 0};
  const size_t xpmu = shift_site(site, mu, +1, Nx, Ny, Nz, Nt);
  for (int nu = 0; nu < 4; ++nu) {
    if (nu == mu) continue;
    const size_t xpnu = shift_site(site, nu, +1, Nx, Ny, Nz, Nt);
    const size_t xmnu = shift_site(site, nu, -1, Nx, Ny, Nz, Nt);
    const size_t xmnupmu = shift_site(xmnu, mu, +1, Nx, Ny, Nz, Nt);
    // upper: U_nu(x) U_mu(x+nu) U_nu(x+mu)^+
    const M3 up = mul_nn(load3(Uin + 18 * (site + V * nu)),
                         mul_nd(load3(Uin + 18 * (xpnu + V * mu)), load3(Uin + 18 * (xpmu + V * nu))));
    // lower: U_nu(x-nu)^+ U_mu(x-nu) U_nu(x-nu+mu)
    const M3 lo = mul_dn(load3(Uin + 18 * (xmnu + V * nu)),
                         mul_nn(load3(Uin + 18 * (xmnu + V * mu)), load3(Uin + 18 * (xmnupmu + V * nu))));
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int j = 0; j < 3; ++j) {
        Cst.a[i][j].r += rho * up.a[i][j].r + rho * lo.a[i][j].r;
        Cst.a[i][j].i += rho * up.a[i][j].i + rho * lo.a[i][j].i;
      }
  }
  const M3 U = load3(Uin + 18 * (site + V * mu));
  const M3 Om = mul_nd(Cst, U);
  M3 iQ;
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j)
      iQ.a[i][j] = {0.5 * (Om.a[i][j].r - Om.a[j][i].r), 0.5 * (Om.a[i][j].i + Om.a[j][i].i)};
  const double tr = (iQ.a[0][0].i + iQ.a[1][1].i + iQ.a[2][2].i) / 3.0;
#pragma unroll
  for (int i = 0; i < 3; ++i) iQ.a[i][i].i -= tr;
  store3(Uout + 18 * (site + V * mu), mul_nn(exp_iQ(iQ), U));
}

hipError_t launch_stout(double* Uout, const double* Uin, int Nx, int Ny, int Nz, int Nt,
                        double rho, hipStream_t st) {
  const size_t n = 4 * (size_t)Nx * Ny * Nz * Nt;
  hipLaunchKernelGGL(k_stout, dim3((unsigned)((n + 127) / 128)), dim3(128), 0, st, Uout, Uin,
                     Nx, Ny, Nz, Nt, rho);
  return hipGetLastError();
}

__device__ __forceinline__ double wave_sum_g(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  return v;
}

__global__ void __launch_bounds__(256)
k_plaquette(const double* __restrict__ U, int Nx, int Ny, int Nz, int Nt, double* partial,
            unsigned* counter, double* result) {
  const size_t V = (size_t)Nx * Ny * Nz * Nt;
  double acc = 0;
  for (size_t site = blockIdx.x * (size_t)blockDim.x + threadIdx.x; site < V;
       site += (size_t)gridDim.x * blockDim.x) {
    for (int mu = 0; mu < 4; ++mu)
      for (int nu = mu + 1; nu < 4; ++nu) {
        const size_t xpmu = shift_site(site, mu, +1, Nx, Ny, Nz, Nt);
        const size_t xpnu = shift_site(site, nu, +1, Nx, Ny, Nz, Nt);
        const M3 a = mul_nn(load3(U + 18 * (site + V * mu)), load3(U + 18 * (xpmu + V * nu)));
        const M3 b = mul_nn(load3(U + 18 * (site + V * nu)), load3(U + 18 * (xpnu + V * mu)));
        // Re tr(a b^+)
        for (int i = 0; i < 3; ++i)
          for (int k = 0; k < 3; ++k) acc += a.a[i][k].r * b.a[i][k].r + a.a[i][k].i * b.a[i][k].i;
      }
  }
  // block and grid reduction (same scheme as the field reductions)
  __shared__ double sh[4];
  __shared__ bool amlast;
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  double t = wave_sum_g(acc);
  if (lane == 0) sh[wid] = t;
  __syncthreads();
  if (threadIdx.x == 0) {
    partial[blockIdx.x] = sh[0] + sh[1] + sh[2] + sh[3];
    __threadfence();
    unsigned prev = atomicAdd(counter, 1u);
    amlast = (prev == gridDim.x - 1);
  }
  __syncthreads();
  if (!amlast) return;
  __threadfence();
  double w = 0;
  for (unsigned i = threadIdx.x; i < gridDim.x; i += blockDim.x)
    w += __hip_atomic_load(partial + i, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  __syncthreads();
  t = wave_sum_g(w);
  if (lane == 0) sh[wid] = t;
  __syncthreads();
  if (threadIdx.x == 0) {
    *result = sh[0] + sh[1] + sh[2] + sh[3];
    __hip_atomic_store(counter, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
}

hipError_t launch_plaquette(const double* U, int Nx, int Ny, int Nz, int Nt,
                            const RedScratch& red, int slot, hipStream_t st) {
  const size_t V = (size_t)Nx * Ny * Nz * Nt;
  size_t g = (V + 255) / 256;
  if (g > (size_t)kRedBlocks) g = kRedBlocks;
  hipLaunchKernelGGL(k_plaquette, dim3((unsigned)g), dim3(256), 0, st, U, Nx, Ny, Nz, Nt,
                     red.partial, red.counter, red.result + slot);
  return hipGetLastError();
}

}  // namespace dwf
