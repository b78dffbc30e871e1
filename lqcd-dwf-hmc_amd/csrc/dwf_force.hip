// dwf_force.hip -- DWF fermion force kernels (FP64).
//
// Tensor product of the even-odd DWF force (the reference's lexical
// force_udiv1_D / _Ddag followed by tensorProd_5din,
// aforce_F_Domainwall_5din-tmpl.h:338-933, afield_Spinor-inc.h:322-370):
//   F_mu(x) += scale * sum_{s,d} W(x)_{a,d,s} conj(Z(x)_{b,d,s}),
//   W(x) = (1 + sgn gamma_mu) V(x + mu) * bc   (bc when x + mu wraps),
// for the sites x of one parity; V is the field of the other parity.  One
// thread per (site, s); the sum over s runs through LDS inside the block.
// mult_generator (lib_alt_QXS/Force/Fermion/aforce_F.cpp:97-110):
//   F <- -2 at(U F).
#include "dwf_device.h"
#include "dwf_kernels.h"

namespace dwf {

// (1 + sgn gamma_mu) v, Dirac basis; v, w: 24 reals, k = ri + 2*(c + 3*d)
__device__ __forceinline__ void one_pm_gamma(double* w, const double* v, int mu, double sgn) {
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    const int r = 2 * c, i = 2 * c + 1;
    const double* v0 = v; const double* v1 = v + 6; const double* v2 = v + 12; const double* v3 = v + 18;
    double g[4][2];
    if (mu == 0) {         // g1: (-i v3, -i v2, i v1, i v0)
      g[0][0] = v3[i];  g[0][1] = -v3[r];
      g[1][0] = v2[i];  g[1][1] = -v2[r];
      g[2][0] = -v1[i]; g[2][1] = v1[r];
      g[3][0] = -v0[i]; g[3][1] = v0[r];
    } else if (mu == 1) {  // g2: (-v3, v2, v1, -v0)
      g[0][0] = -v3[r]; g[0][1] = -v3[i];
      g[1][0] = v2[r];  g[1][1] = v2[i];
      g[2][0] = v1[r];  g[2][1] = v1[i];
      g[3][0] = -v0[r]; g[3][1] = -v0[i];
    } else if (mu == 2) {  // g3: (-i v2, i v3, i v0, -i v1)
      g[0][0] = v2[i];  g[0][1] = -v2[r];
      g[1][0] = -v3[i]; g[1][1] = v3[r];
      g[2][0] = -v0[i]; g[2][1] = v0[r];
      g[3][0] = v1[i];  g[3][1] = -v1[r];
    } else {               // g4: (v0, v1, -v2, -v3)
      g[0][0] = v0[r];  g[0][1] = v0[i];
      g[1][0] = v1[r];  g[1][1] = v1[i];
      g[2][0] = -v2[r]; g[2][1] = -v2[i];
      g[3][0] = -v3[r]; g[3][1] = -v3[i];
    }
#pragma unroll
    for (int d = 0; d < 4; ++d) {
      w[r + 6 * d] = v[r + 6 * d] + sgn * g[d][0];
      w[i + 6 * d] = v[i + 6 * d] + sgn * g[d][1];
    }
  }
}

__global__ void __launch_bounds__(256)
k_force_tensor(double* __restrict__ F, const double* __restrict__ Z,
               const double* __restrict__ Vn, const int* __restrict__ nbr, int parity, int Nx,
               int Ny, int Nz, int Nt, int bcx, int bcy, int bcz, int bct, int Ls, int spb,
               double sgn, double scale, int accumulate) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  double* sh = reinterpret_cast<double*>(smem_raw);   // [spb][Ls][18]
  const int Vh = Nx / 2 * Ny * Nz * Nt;
  const size_t V = 2 * (size_t)Vh;
  const int tid = threadIdx.x;
  const int lsite = tid / Ls, ls = tid - lsite * Ls;
  const int site = blockIdx.x * spb + lsite;
  const bool valid = lsite < spb && site < Vh;
  // coordinates of the output site (parity p)
  int xc[4] = {0, 0, 0, 0};
  size_t lex = 0;
  if (valid) {
    const int Nx2 = Nx / 2;
    int r = site / Nx2;
    const int x2 = site - r * Nx2;
    xc[1] = r % Ny;
    r /= Ny;
    xc[2] = r % Nz;
    xc[3] = r / Nz;
    xc[0] = 2 * x2 + ((xc[1] + xc[2] + xc[3] + parity) & 1);
    lex = xc[0] + (size_t)Nx * (xc[1] + (size_t)Ny * (xc[2] + (size_t)Nz * xc[3]));
  }
  const int N[4] = {Nx, Ny, Nz, Nt};
  const int bc[4] = {bcx, bcy, bcz, bct};
  double z[24];
  if (valid) {
    const double* zp = Z + (size_t)site * 24 * Ls + ls;
#pragma unroll
    for (int k = 0; k < 24; ++k) z[k] = zp[k * Ls];
  }
  for (int mu = 0; mu < 4; ++mu) {
    double acc[18];
#pragma unroll
    for (int e = 0; e < 18; ++e) acc[e] = 0;
    if (valid) {
      const int nb = nbr[((size_t)parity * 8 + 2 * mu) * Vh + site];
      const double* vp = Vn + (size_t)nb * 24 * Ls + ls;
      double v[24], w[24];
#pragma unroll
      for (int k = 0; k < 24; ++k) v[k] = vp[k * Ls];
      one_pm_gamma(w, v, mu, sgn);
#pragma unroll
      for (int a = 0; a < 3; ++a)
#pragma unroll
        for (int b = 0; b < 3; ++b) {
          double re = 0, im = 0;
#pragma unroll
          for (int d = 0; d < 4; ++d) {
            const double wr = w[2 * a + 6 * d], wi = w[2 * a + 6 * d + 1];
            const double zr = z[2 * b + 6 * d], zi = z[2 * b + 6 * d + 1];
            re += wr * zr + wi * zi;
            im += wi * zr - wr * zi;
          }
          acc[2 * (3 * a + b)] = re;
          acc[2 * (3 * a + b) + 1] = im;
        }
    }
    __syncthreads();
    if (valid) {
#pragma unroll
      for (int e = 0; e < 18; ++e) sh[(lsite * Ls + ls) * 18 + e] = acc[e];
    }
    __syncthreads();
    // the Ls threads of a site sum over s and write the 18 reals of F_mu(x)
    if (valid) {
      const double f = scale * ((xc[mu] == N[mu] - 1) ? (double)bc[mu] : 1.0);
      for (int e = ls; e < 18; e += Ls) {
        double t = 0;
        for (int s = 0; s < Ls; ++s) t += sh[(lsite * Ls + s) * 18 + e];
        double* dst = F + 18 * (lex + V * mu) + e;
        *dst = accumulate ? *dst + f * t : f * t;
      }
    }
  }
}

hipError_t launch_force_tensor(double* F, const double* Z, const double* Vn, const int* nbr,
                               int parity, int Nx, int Ny, int Nz, int Nt, const int* bc,
                               int Ls, double sgn, double scale, bool accumulate,
                               hipStream_t st) {
  const int Vh = Nx / 2 * Ny * Nz * Nt;
  int spb = 256 / Ls;
  if (spb < 1) return hipErrorInvalidValue;
  while (spb > 1 && (size_t)spb * Ls * 18 * 8 > 64 * 1024) --spb;
  const size_t lds = (size_t)spb * Ls * 18 * 8;
  hipLaunchKernelGGL(k_force_tensor, dim3((Vh + spb - 1) / spb), dim3(spb * Ls), lds, st, F, Z,
                     Vn, nbr, parity, Nx, Ny, Nz, Nt, bc[0], bc[1], bc[2], bc[3], Ls, spb, sgn,
                     scale, accumulate ? 1 : 0);
  return hipGetLastError();
}

// F_mu(x) <- -2 at(U_mu(x) F_mu(x)); the links are taken from the packed
// forward entries (which carry the boundary factor bc) times bc, i.e. U_mu(x)
// for bc = +-1.
__global__ void __launch_bounds__(256)
k_mult_generator(double* __restrict__ F, const double* __restrict__ links, int Nx, int Ny,
                 int Nz, int Nt, int bcx, int bcy, int bcz, int bct) {
  const int Vh = Nx / 2 * Ny * Nz * Nt;
  const size_t V = 2 * (size_t)Vh;
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= 2 * Vh) return;
  const int p = idx / Vh, s2 = idx - p * Vh;
  const int Nx2 = Nx / 2;
  int r = s2 / Nx2;
  const int x2 = s2 - r * Nx2;
  const int y = r % Ny;
  r /= Ny;
  const int z = r % Nz;
  const int t = r / Nz;
  const int x = 2 * x2 + ((y + z + t + p) & 1);
  const int xc[4] = {x, y, z, t};
  const int N[4] = {Nx, Ny, Nz, Nt};
  const int bc[4] = {bcx, bcy, bcz, bct};
  const size_t lex = x + (size_t)Nx * (y + (size_t)Ny * (z + (size_t)Nz * t));
  for (int mu = 0; mu < 4; ++mu) {
    const double f = (xc[mu] == N[mu] - 1) ? (double)bc[mu] : 1.0;
    const double* L = links + ((size_t)(p * 8 + 2 * mu) * 18) * Vh + s2;
    double u[18], fm[18], m[18];
    for (int e = 0; e < 18; ++e) u[e] = f * L[(size_t)e * Vh];
    double* Fp = F + 18 * (lex + V * mu);
    for (int e = 0; e < 18; ++e) fm[e] = Fp[e];
    for (int a = 0; a < 3; ++a)
      for (int b = 0; b < 3; ++b) {
        double re = 0, im = 0;
        for (int c = 0; c < 3; ++c) {
          const double ur = u[2 * (3 * a + c)], ui = u[2 * (3 * a + c) + 1];
          const double vr = fm[2 * (3 * c + b)], vi = fm[2 * (3 * c + b) + 1];
          re += ur * vr - ui * vi;
          im += ur * vi + ui * vr;
        }
        m[2 * (3 * a + b)] = re;
        m[2 * (3 * a + b) + 1] = im;
      }
    // at(m) = (m - m^+)/2 - tr(m - m^+)/6, times -2
    double o[18];
    for (int a = 0; a < 3; ++a)
      for (int b = 0; b < 3; ++b) {
        o[2 * (3 * a + b)] = 0.5 * (m[2 * (3 * a + b)] - m[2 * (3 * b + a)]);
        o[2 * (3 * a + b) + 1] = 0.5 * (m[2 * (3 * a + b) + 1] + m[2 * (3 * b + a) + 1]);
      }
    const double tr = (o[1] + o[9] + o[17]) / 3.0;
    o[1] -= tr;
    o[9] -= tr;
    o[17] -= tr;
    for (int e = 0; e < 18; ++e) Fp[e] = -2.0 * o[e];
  }
}

hipError_t launch_mult_generator(double* F, const double* links, int Nx, int Ny, int Nz, int Nt,
                                 const int* bc, hipStream_t st) {
  const int Vh = Nx / 2 * Ny * Nz * Nt;
  hipLaunchKernelGGL(k_mult_generator, dim3((2 * Vh + 255) / 256), dim3(256), 0, st, F, links, Nx,
                     Ny, Nz, Nt, bc[0], bc[1], bc[2], bc[3]);
  return hipGetLastError();
}

}  // namespace dwf
