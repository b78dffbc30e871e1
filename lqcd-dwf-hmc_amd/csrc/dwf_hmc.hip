// dwf_hmc.hip -- FP64 gauge kernels of the HMC: conjugate momenta, the link
// update exp(alpha iP) U, and the plaquette + rectangle (Iwasaki-type) gauge
// action and force.  Gauge fields are the corelib Field_G layout
// [mu][lexical site][18]; one thread per link or per site.
#include "dwf_device.h"
#include "dwf_kernels.h"
#include "dwf_su3.h"

namespace dwf {

__device__ __forceinline__ double hmc_block_sum(double v, double* sh) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = v;
  __syncthreads();
  double r = 0;
  for (int w = 0; w < (int)(blockDim.x >> 6); ++w) r += sh[w];
  return r;
}

// iP = i P^a lambda^a from 8 Gaussian numbers h_1..h_8 per link
// (Langevin_Momentum::set_iP_SU3, src/lib/HMC/langevin_Momentum.cpp:146-200):
//   P = | h3 + h8/sqrt3   h1 - i h2        h4 - i h5    |
//       | h1 + i h2       -h3 + h8/sqrt3   h6 - i h7    |
//       | h4 + i h5       h6 + i h7        -2 h8/sqrt3  |,  iP = i P;
// block partials of |iP|^2.
__global__ void __launch_bounds__(256)
k_set_iP(double* __restrict__ P, const double* __restrict__ H, size_t nlinks,
         double* __restrict__ partial) {
  const double sq3r = 1.0 / sqrt(3.0), sq3r2 = 2.0 * sq3r;
  double acc = 0;
  for (size_t l = blockIdx.x * (size_t)blockDim.x + threadIdx.x; l < nlinks;
       l += (size_t)gridDim.x * blockDim.x) {
    const double* h = H + 8 * l;
    const double h1 = h[0], h2 = h[1], h3 = h[2], h4 = h[3], h5 = h[4], h6 = h[5], h7 = h[6],
                 h8 = h[7];
    const double v[18] = {0.0, h3 + h8 * sq3r, h2, h1, h5, h4,
                          -h2, h1, 0.0, -h3 + h8 * sq3r, h7, h6,
                          -h5, h4, -h7, h6, 0.0, -h8 * sq3r2};
    double* o = P + 18 * l;
#pragma unroll
    for (int e = 0; e < 18; ++e) {
      o[e] = v[e];
      acc += v[e] * v[e];
    }
  }
  __shared__ double sh[4];
  const double t = hmc_block_sum(acc, sh);
  if (threadIdx.x == 0) partial[blockIdx.x] = t;
}

// Gram-Schmidt of the first two rows, third row = (row0 x row1)^*
// (Mat_SU_N::reunit_SU3, src/lib/Tools/mat_SU_N.cpp:18-86)
__device__ __forceinline__ void reunit_su3(double* va) {
  const double sn1 = sqrt(1.0 / (va[0] * va[0] + va[1] * va[1] + va[2] * va[2] + va[3] * va[3] +
                                 va[4] * va[4] + va[5] * va[5]));
#pragma unroll
  for (int i = 0; i < 6; ++i) va[i] *= sn1;
  const double sp1r = va[0] * va[6] + va[1] * va[7] + va[2] * va[8] + va[3] * va[9] +
                      va[4] * va[10] + va[5] * va[11];
  const double sp1i = va[0] * va[7] - va[1] * va[6] + va[2] * va[9] - va[3] * va[8] +
                      va[4] * va[11] - va[5] * va[10];
  va[6] = va[6] - sp1r * va[0] + sp1i * va[1];
  va[7] = va[7] - sp1r * va[1] - sp1i * va[0];
  va[8] = va[8] - sp1r * va[2] + sp1i * va[3];
  va[9] = va[9] - sp1r * va[3] - sp1i * va[2];
  va[10] = va[10] - sp1r * va[4] + sp1i * va[5];
  va[11] = va[11] - sp1r * va[5] - sp1i * va[4];
  const double sn2 = sqrt(1.0 / (va[6] * va[6] + va[7] * va[7] + va[8] * va[8] + va[9] * va[9] +
                                 va[10] * va[10] + va[11] * va[11]));
#pragma unroll
  for (int i = 6; i < 12; ++i) va[i] *= sn2;
  va[12] = va[2] * va[10] - va[3] * va[11] - va[4] * va[8] + va[5] * va[9];
  va[13] = -va[2] * va[11] - va[3] * va[10] + va[4] * va[9] + va[5] * va[8];
  va[14] = va[4] * va[6] - va[5] * va[7] - va[0] * va[10] + va[1] * va[11];
  va[15] = -va[4] * va[7] - va[5] * va[6] + va[0] * va[11] + va[1] * va[10];
  va[16] = va[0] * va[8] - va[1] * va[9] - va[2] * va[6] + va[3] * va[7];
  va[17] = -va[0] * va[9] - va[1] * va[8] + va[2] * va[7] + va[3] * va[6];
}

// W = exp(alpha iP) U by the nested Taylor form of order n,
//   U + alpha iP (U + alpha/2 iP (... (U + alpha/n iP U))),
// then reunitarised (SU_N::mat_exp, src/lib/Tools/mat_SU_N.h:637-655;
// mult_exp_Field_G, src/lib/Field/Imp/field_G_imp.cpp:496-520).  W may alias U.
__global__ void __launch_bounds__(128)
k_exp_update(double* W, double alpha, const double* __restrict__ P, const double* U,
             size_t nlinks, int n) {
  for (size_t l = blockIdx.x * (size_t)blockDim.x + threadIdx.x; l < nlinks;
       l += (size_t)gridDim.x * blockDim.x) {
    const M3 u0 = load3(U + 18 * l);
    const M3 iv = load3(P + 18 * l);
    M3 tmp = u0;
    for (int k = 0; k < n; ++k) {
      const double f = alpha / (n - k);
      M3 p = mul_nn(iv, tmp);
#pragma unroll
      for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) {
          tmp.a[i][j].r = p.a[i][j].r * f + u0.a[i][j].r;
          tmp.a[i][j].i = p.a[i][j].i * f + u0.a[i][j].i;
        }
    }
    double va[18];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int j = 0; j < 3; ++j) {
        va[2 * (3 * i + j)] = tmp.a[i][j].r;
        va[2 * (3 * i + j) + 1] = tmp.a[i][j].i;
      }
    reunit_su3(va);
    double* o = W + 18 * l;
#pragma unroll
    for (int e = 0; e < 18; ++e) o[e] = va[e];
  }
}

struct Lat {
  int N[4];
  size_t V;
  // site + d * mu_hat (periodic)
  __device__ __forceinline__ size_t mv(size_t s, int mu, int d) const {
    int c[4];
    size_t q = s;
    c[0] = (int)(q % N[0]);
    q /= N[0];
    c[1] = (int)(q % N[1]);
    q /= N[1];
    c[2] = (int)(q % N[2]);
    c[3] = (int)(q / N[2]);
    c[mu] = ((c[mu] + d) % N[mu] + N[mu]) % N[mu];
    return c[0] + (size_t)N[0] * (c[1] + (size_t)N[1] * (c[2] + (size_t)N[2] * c[3]));
  }
  __device__ __forceinline__ M3 u(const double* U, int mu, size_t s) const {
    return load3(U + 18 * (s + V * mu));
  }
};

__device__ __forceinline__ double retr_nd(const M3& a, const M3& b) {   // Re tr(a b^+)
  double t = 0;
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) t += a.a[i][j].r * b.a[i][j].r + a.a[i][j].i * b.a[i][j].i;
  return t;
}

// Per site: sum over mu < nu of Re tr(U_mu(x) S^+) for the plaquette staple and
// the two rectangle staples (2 x 1 along mu, 1 x 2 along nu) that
// Action_G_Rectangle::calcH sums (src/lib/Action/Gauge/action_G_Rectangle.cpp:
// 113-190); block partials [2 * block] = plaquette sum, [2 * block + 1] = rectangles.
__global__ void __launch_bounds__(128)
k_rect_action(const double* __restrict__ U, Lat L, double* __restrict__ partial) {
  double plaq = 0, rect = 0;
  for (size_t x = blockIdx.x * (size_t)blockDim.x + threadIdx.x; x < L.V;
       x += (size_t)gridDim.x * blockDim.x) {
    for (int mu = 0; mu < 4; ++mu)
      for (int nu = mu + 1; nu < 4; ++nu) {
        const size_t xm = L.mv(x, mu, 1), xn = L.mv(x, nu, 1);
        const size_t xmn = L.mv(xm, nu, 1), xmm = L.mv(xm, mu, 1), xnn = L.mv(xn, nu, 1);
        const M3 umu = L.u(U, mu, x);
        const M3 unu_x = L.u(U, nu, x);
        // upper plaquette staple U_nu(x) U_mu(x+nu) U_nu(x+mu)^+
        const M3 cup = mul_nn(unu_x, mul_nd(L.u(U, mu, xn), L.u(U, nu, xm)));
        plaq += retr_nd(umu, cup);
        // 2 x 1: U_nu(x) U_mu(x+nu) U_mu(x+mu+nu) U_nu(x+2mu)^+ U_mu(x+mu)^+
        const M3 ra = mul_nn(mul_nn(unu_x, L.u(U, mu, xn)),
                             mul_nd(mul_nd(L.u(U, mu, xmn), L.u(U, nu, xmm)), L.u(U, mu, xm)));
        // 1 x 2: U_nu(x) U_nu(x+nu) U_mu(x+2nu) U_nu(x+mu+nu)^+ U_nu(x+mu)^+
        const M3 rb = mul_nn(mul_nn(unu_x, L.u(U, nu, xn)),
                             mul_nd(mul_nd(L.u(U, mu, xnn), L.u(U, nu, xmn)), L.u(U, nu, xm)));
        rect += retr_nd(umu, ra) + retr_nd(umu, rb);
      }
  }
  __shared__ double sh[2];
  __shared__ double shw[2][4];
  double v[2] = {plaq, rect};
#pragma unroll
  for (int k = 0; k < 2; ++k) {
    double t = v[k];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) t += __shfl_down(t, off, 64);
    if ((threadIdx.x & 63) == 0) shw[k][threadIdx.x >> 6] = t;
  }
  __syncthreads();
  if (threadIdx.x < 2) {
    double r = 0;
    for (int w = 0; w < (int)(blockDim.x >> 6); ++w) r += shw[threadIdx.x][w];
    sh[threadIdx.x] = r;
    partial[2 * blockIdx.x + threadIdx.x] = r;
  }
}

__device__ __forceinline__ void m3_axpy(M3& y, double a, const M3& x) {
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      y.a[i][j].r += a * x.a[i][j].r;
      y.a[i][j].i += a * x.a[i][j].i;
    }
}

// F_mu(x) = -(beta / 3) at(U_mu(x) S^+), S = sum_{nu != mu} of c_plaq times the
// two plaquette staples and c_rect times the six rectangle staples
// (Force_G_Rectangle::force_core, src/lib/Force/Gauge/force_G_Rectangle.cpp:84-216).
// One thread per link; F overwritten.
__global__ void __launch_bounds__(128)
k_rect_force(double* __restrict__ F, const double* __restrict__ U, Lat L, double beta,
             double c_plaq, double c_rect) {
  const size_t nl = 4 * L.V;
  for (size_t l = blockIdx.x * (size_t)blockDim.x + threadIdx.x; l < nl;
       l += (size_t)gridDim.x * blockDim.x) {
    const int mu = (int)(l / L.V);
    const size_t x = l - (size_t)mu * L.V;
    M3 S;
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int j = 0; j < 3; ++j) S.a[i][j] = {0, 0};
    const size_t xpm = L.mv(x, mu, 1), xmm = L.mv(x, mu, -1), xpm2 = L.mv(xpm, mu, 1);
    for (int nu = 0; nu < 4; ++nu) {
      if (nu == mu) continue;
      const size_t xpn = L.mv(x, nu, 1), xmn = L.mv(x, nu, -1);
      const size_t xpmpn = L.mv(xpm, nu, 1), xmnpm = L.mv(xmn, mu, 1);
      // plaquette staples
      const M3 cup = mul_nn(L.u(U, nu, x), mul_nd(L.u(U, mu, xpn), L.u(U, nu, xpm)));
      const M3 cdn = mul_dn(L.u(U, nu, xmn), mul_nn(L.u(U, mu, xmn), L.u(U, nu, xmnpm)));
      m3_axpy(S, c_plaq, cup);
      m3_axpy(S, c_plaq, cdn);
      if (c_rect != 0.0) {
        const size_t xpn2 = L.mv(xpn, nu, 1), xmn2 = L.mv(xmn, nu, -1);
        const size_t xmmpn = L.mv(xmm, nu, 1), xmmmn = L.mv(xmm, nu, -1);
        const size_t xmnpm2 = L.mv(xmnpm, mu, 1), xmn2pm = L.mv(xmn2, mu, 1);
        // U_nu(x) U_mu(x+nu) U_mu(x+mu+nu) U_nu(x+2mu)^+ U_mu(x+mu)^+
        const M3 a = mul_nn(mul_nn(L.u(U, nu, x), L.u(U, mu, xpn)),
                            mul_nd(mul_nd(L.u(U, mu, xpmpn), L.u(U, nu, xpm2)), L.u(U, mu, xpm)));
        // U_nu(x) U_nu(x+nu) U_mu(x+2nu) U_nu(x+mu+nu)^+ U_nu(x+mu)^+
        const M3 b = mul_nn(mul_nn(L.u(U, nu, x), L.u(U, nu, xpn)),
                            mul_nd(mul_nd(L.u(U, mu, xpn2), L.u(U, nu, xpmpn)), L.u(U, nu, xpm)));
        // U_mu(x-mu)^+ U_nu(x-mu) U_mu(x-mu+nu) U_mu(x+nu) U_nu(x+mu)^+
        const M3 c = mul_nn(mul_dn(L.u(U, mu, xmm), L.u(U, nu, xmm)),
                            mul_nn(L.u(U, mu, xmmpn), mul_nd(L.u(U, mu, xpn), L.u(U, nu, xpm))));
        // U_nu(x-nu)^+ U_mu(x-nu) U_mu(x-nu+mu) U_nu(x-nu+2mu) U_mu(x+mu)^+
        const M3 d = mul_nn(mul_dn(L.u(U, nu, xmn), L.u(U, mu, xmn)),
                            mul_nn(L.u(U, mu, xmnpm), mul_nd(L.u(U, nu, xmnpm2), L.u(U, mu, xpm))));
        // U_nu(x-nu)^+ U_nu(x-2nu)^+ U_mu(x-2nu) U_nu(x-2nu+mu) U_nu(x-nu+mu)
        const M3 e = mul_nn(mul_dn(L.u(U, nu, xmn), dag3(L.u(U, nu, xmn2))),
                            mul_nn(L.u(U, mu, xmn2), mul_nn(L.u(U, nu, xmn2pm), L.u(U, nu, xmnpm))));
        // U_mu(x-mu)^+ U_nu(x-mu-nu)^+ U_mu(x-mu-nu) U_mu(x-nu) U_nu(x-nu+mu)
        const M3 f = mul_nn(mul_dn(L.u(U, mu, xmm), dag3(L.u(U, nu, xmmmn))),
                            mul_nn(L.u(U, mu, xmmmn), mul_nn(L.u(U, mu, xmn), L.u(U, nu, xmnpm))));
        m3_axpy(S, c_rect, a);
        m3_axpy(S, c_rect, b);
        m3_axpy(S, c_rect, c);
        m3_axpy(S, c_rect, d);
        m3_axpy(S, c_rect, e);
        m3_axpy(S, c_rect, f);
      }
    }
    const M3 g = at3(mul_nd(L.u(U, mu, x), S));
    M3 out;
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int j = 0; j < 3; ++j) out.a[i][j] = cscal(-beta / 3.0, g.a[i][j]);
    store3(F + 18 * l, out);
  }
}

static unsigned hmc_grid(size_t n, int block) {
  size_t g = (n + block - 1) / block;
  if (g > 2048) g = 2048;
  return (unsigned)(g < 1 ? 1 : g);
}

hipError_t launch_set_iP(double* P, const double* H, size_t nlinks, double* partial,
                         unsigned* nblk, hipStream_t st) {
  const unsigned g = hmc_grid(nlinks, 256) > (unsigned)kRedBlocks ? (unsigned)kRedBlocks
                                                                   : hmc_grid(nlinks, 256);
  *nblk = g;
  hipLaunchKernelGGL(k_set_iP, dim3(g), dim3(256), 0, st, P, H, nlinks, partial);
  return hipGetLastError();
}

hipError_t launch_exp_update(double* W, double alpha, const double* P, const double* U,
                             size_t nlinks, int nprec, hipStream_t st) {
  hipLaunchKernelGGL(k_exp_update, dim3(hmc_grid(nlinks, 128)), dim3(128), 0, st, W, alpha, P,
                     U, nlinks, nprec);
  return hipGetLastError();
}

hipError_t launch_rect_action(const double* U, const int dims[4], double* partial,
                              unsigned* nblk, hipStream_t st) {
  Lat L;
  for (int i = 0; i < 4; ++i) L.N[i] = dims[i];
  L.V = (size_t)dims[0] * dims[1] * dims[2] * dims[3];
  unsigned g = hmc_grid(L.V, 128);
  if (g > (unsigned)kRedBlocks) g = kRedBlocks;
  *nblk = g;
  hipLaunchKernelGGL(k_rect_action, dim3(g), dim3(128), 0, st, U, L, partial);
  return hipGetLastError();
}

hipError_t launch_rect_force(double* F, const double* U, const int dims[4], double beta,
                             double c_plaq, double c_rect, hipStream_t st) {
  Lat L;
  for (int i = 0; i < 4; ++i) L.N[i] = dims[i];
  L.V = (size_t)dims[0] * dims[1] * dims[2] * dims[3];
  hipLaunchKernelGGL(k_rect_force, dim3(hmc_grid(4 * L.V, 128)), dim3(128), 0, st, F, U, L,
                     beta, c_plaq, c_rect);
  return hipGetLastError();
}

}  // namespace dwf
