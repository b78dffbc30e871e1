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
#include "dwf_su3.h"

namespace dwf {

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

// C_mu(x) = rho sum_{nu != mu} (upper + lower staple)
__device__ M3 staple_sum(const double* __restrict__ Uin, size_t site, int mu, size_t V, int Nx,
                         int Ny, int Nz, int Nt, double rho) {
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
  return Cst;
}

__global__ void __launch_bounds__(128)
k_stout(double* __restrict__ Uout, const double* __restrict__ Uin, int Nx, int Ny, int Nz,
        int Nt, double rho) {
  const size_t V = (size_t)Nx * Ny * Nz * Nt;
  const size_t idx = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  if (idx >= 4 * V) return;
  const int mu = (int)(idx / V);
  const size_t site = idx - (size_t)mu * V;
  const M3 Cst = staple_sum(Uin, site, mu, V, Nx, Ny, Nz, Nt, rho);
  const M3 U = load3(Uin + 18 * (site + V * mu));
  const M3 iQ = at3(mul_nd(Cst, U));
  store3(Uout + 18 * (site + V * mu), mul_nn(exp_iQ(iQ), U));
}

hipError_t launch_stout(double* Uout, const double* Uin, int Nx, int Ny, int Nz, int Nt,
                        double rho, hipStream_t st) {
  const size_t n = 4 * (size_t)Nx * Ny * Nz * Nt;
  hipLaunchKernelGGL(k_stout, dim3((unsigned)((n + 127) / 128)), dim3(128), 0, st, Uout, Uin,
                     Nx, Ny, Nz, Nt, rho);
  return hipGetLastError();
}

// ---- stout force recursion --------------------------------------------------
// One level of AForceSmear_APE::force_udiv (lib_alt_QXS/Smear/aforceSmear_APE-tmpl.h:
// 157-206): Sigma = dS/dU from Sigma' = dS/dU' of U' = exp(iQ) U.  Per link, the
// Morningstar-Peardon chain rule (AProjection_Stout_SU3::force_recursive,
// aprojection_Stout_SU3-tmpl.h:691-895):
//   Gamma = tr(U S' B1) Q + tr(U S' B2) Q^2 + f1 U S' + f2 (Q U S' + U S' Q),
//   iLambda = at(i Gamma),  iTheta = iLambda U,  Xi = S' exp(iQ) + C^+ iLambda,
// then Sigma_mu(x) = Xi + rho sum_nu (six staple products of U and iTheta).

__device__ __forceinline__ M3 m3_poly(cplx c0, cplx c1, cplx c2, const M3& Q, const M3& Q2) {
  M3 r;
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      cplx v = cadd(cmul(c1, Q.a[i][j]), cmul(c2, Q2.a[i][j]));
      if (i == j) v = cadd(v, c0);
      r.a[i][j] = v;
    }
  return r;
}
__device__ __forceinline__ cplx m3_tr(const M3& x) {
  return cadd(cadd(x.a[0][0], x.a[1][1]), x.a[2][2]);
}
__device__ __forceinline__ cplx cexp_i(double phi) { return {cos(phi), sin(phi)}; }

__device__ double stout_xi1(double w) {
  if (w < 0.25) {
    const double w2 = w * w;
    return -1.0 / 3.0 +
           w2 * (1.0 / 30.0 + w2 * (-1.0 / 840.0 + w2 * (1.0 / 45360.0 - w2 / 3991680.0)));
  }
  return (w * cos(w) - sin(w)) / (w * w * w);
}

__global__ void __launch_bounds__(128)
k_stout_force_link(double* __restrict__ Sigma, double* __restrict__ iTheta,
                   const double* __restrict__ Sigmap, const double* __restrict__ Uin, int Nx,
                   int Ny, int Nz, int Nt, double rho) {
  const size_t V = (size_t)Nx * Ny * Nz * Nt;
  const size_t idx = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  if (idx >= 4 * V) return;
  const int mu = (int)(idx / V);
  const size_t site = idx - (size_t)mu * V;
  const size_t off = 18 * (site + V * mu);
  const M3 C = staple_sum(Uin, site, mu, V, Nx, Ny, Nz, Nt, rho);
  const M3 U = load3(Uin + off);
  const M3 Sp = load3(Sigmap + off);
  const M3 iQ = at3(mul_nd(C, U));
  double n2 = 0;
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) n2 += iQ.a[i][j].r * iQ.a[i][j].r + iQ.a[i][j].i * iQ.a[i][j].i;
  M3 eiQ, iGam;
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      eiQ.a[i][j] = {i == j ? 1.0 : 0.0, 0.0};
      iGam.a[i][j] = {0.0, 0.0};
    }
  if (n2 > 1.0e-10) {
    const M3 iQ2 = mul_nn(iQ, iQ);
    const M3 iQ3 = mul_nn(iQ, iQ2);
    M3 Q, Q2;   // Q = -i iQ, Q^2 = -iQ^2
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int j = 0; j < 3; ++j) {
        Q.a[i][j] = {iQ.a[i][j].i, -iQ.a[i][j].r};
        Q2.a[i][j] = {-iQ2.a[i][j].r, -iQ2.a[i][j].i};
      }
    const double c0 = -(iQ3.a[0][0].i + iQ3.a[1][1].i + iQ3.a[2][2].i) / 3.0;
    const double c1 = -0.5 * (iQ2.a[0][0].r + iQ2.a[1][1].r + iQ2.a[2][2].r);
    const double c13r = sqrt(c1 / 3.0);
    const double ratio = fmin(1.0, fmax(-1.0, c0 / (2.0 * c13r * c13r * c13r)));
    const double theta = acos(ratio);
    const double u = c13r * cos(theta / 3.0), w = sqrt(c1) * sin(theta / 3.0);
    const double u2 = u * u, w2 = w * w, cw = cos(w);
    const double xi0 = (w == 0.0) ? 1.0 : sin(w) / w;
    const double xi1 = stout_xi1(w);
    const cplx emiu = cexp_i(-u), e2iu = cexp_i(2.0 * u);
    // f_j
    const double fd = 1.0 / (9.0 * u2 - w2);
    const cplx f0 = cscal(fd, cadd(cscal(u2 - w2, e2iu),
                                   cmul(emiu, cplx{8.0 * u2 * cw, 2.0 * u * (3.0 * u2 + w2) * xi0})));
    const cplx f1 = cscal(fd, cadd(cscal(2.0 * u, e2iu),
                                   cmul(emiu, cplx{-2.0 * u * cw, (3.0 * u2 - w2) * xi0})));
    const cplx f2 = cscal(fd, cadd(e2iu, cmul(emiu, cplx{-cw, -3.0 * u * xi0})));
    // r_j^(1), r_j^(2)
    const cplx r01 = cadd(cmul(cplx{2.0 * u, 2.0 * (u2 - w2)}, e2iu),
                          cmul(emiu, cplx{16.0 * u * cw + 2.0 * u * (3.0 * u2 + w2) * xi0,
                                          -8.0 * u2 * cw + 2.0 * (9.0 * u2 + w2) * xi0}));
    const cplx r11 = cadd(cmul(cplx{2.0, 4.0 * u}, e2iu),
                          cmul(emiu, cplx{-2.0 * cw + (3.0 * u2 - w2) * xi0,
                                          2.0 * u * cw + 6.0 * u * xi0}));
    const cplx r21 = cadd(cmul(cplx{0.0, 2.0}, e2iu),
                          cmul(emiu, cplx{-3.0 * u * xi0, cw - 3.0 * xi0}));
    const cplx r02 = cadd(cscal(-2.0, e2iu),
                          cmul(emiu, cplx{-8.0 * u2 * xi0, 2.0 * u * (cw + xi0 + 3.0 * u2 * xi1)}));
    const cplx r12 = cmul(emiu, cplx{2.0 * u * xi0, -cw - xi0 + 3.0 * u2 * xi1});
    const cplx r22 = cmul(emiu, cplx{xi0, -3.0 * u * xi1});
    // b_ij
    const double bd = 1.0 / (2.0 * (9.0 * u2 - w2) * (9.0 * u2 - w2));
    const double g1 = 3.0 * u2 - w2, g2 = 30.0 * u2 + 2.0 * w2;
    auto bj1 = [&](cplx ra, cplx rb, cplx f) {
      return cscal(bd, cadd(cadd(cscal(2.0 * u, ra), cscal(g1, rb)), cscal(-g2, f)));
    };
    auto bj2 = [&](cplx ra, cplx rb, cplx f) {
      return cscal(bd, cadd(cadd(ra, cscal(-3.0 * u, rb)), cscal(-24.0 * u, f)));
    };
    const M3 US = mul_nn(U, Sp);
    const cplx t1 = m3_tr(mul_nn(US, m3_poly(bj1(r01, r02, f0), bj1(r11, r12, f1),
                                             bj1(r21, r22, f2), Q, Q2)));
    const cplx t2 = m3_tr(mul_nn(US, m3_poly(bj2(r01, r02, f0), bj2(r11, r12, f1),
                                             bj2(r21, r22, f2), Q, Q2)));
    const M3 QUS = mul_nn(Q, US), USQ = mul_nn(US, Q);
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int j = 0; j < 3; ++j) {
        cplx g = cadd(cmul(t1, Q.a[i][j]), cmul(t2, Q2.a[i][j]));
        g = cadd(g, cmul(f1, US.a[i][j]));
        g = cadd(g, cmul(f2, cadd(QUS.a[i][j], USQ.a[i][j])));
        iGam.a[i][j] = {-g.i, g.r};   // i Gamma
      }
    eiQ = m3_poly(f0, f1, f2, Q, Q2);
  }
  const M3 iLam = at3(iGam);
  store3(iTheta + off, mul_nn(iLam, U));
  M3 Xi = mul_nn(Sp, eiQ);
  const M3 CdL = mul_dn(C, iLam);
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) Xi.a[i][j] = cadd(Xi.a[i][j], CdL.a[i][j]);
  store3(Sigma + off, Xi);
}

// Sigma_mu(x) += rho sum_{nu != mu} of the six staple products (force_each,
// aforceSmear_APE-tmpl.h:280-360)
__global__ void __launch_bounds__(128)
k_stout_force_staple(double* __restrict__ Sigma, const double* __restrict__ Uin,
                     const double* __restrict__ iTheta, int Nx, int Ny, int Nz, int Nt,
                     double rho) {
  const size_t V = (size_t)Nx * Ny * Nz * Nt;
  const size_t idx = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  if (idx >= 4 * V) return;
  const int mu = (int)(idx / V);
  const size_t site = idx - (size_t)mu * V;
  auto Ul = [&](size_t s, int d) { return load3(Uin + 18 * (s + V * d)); };
  auto Th = [&](size_t s, int d) { return load3(iTheta + 18 * (s + V * d)); };
  M3 acc;
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) acc.a[i][j] = {0, 0};
  const size_t xpmu = shift_site(site, mu, +1, Nx, Ny, Nz, Nt);
  for (int nu = 0; nu < 4; ++nu) {
    if (nu == mu) continue;
    const size_t xpnu = shift_site(site, nu, +1, Nx, Ny, Nz, Nt);
    const size_t xmnu = shift_site(site, nu, -1, Nx, Ny, Nz, Nt);
    const size_t xmnupmu = shift_site(xmnu, mu, +1, Nx, Ny, Nz, Nt);
    const M3 Vnu_xpmu = Ul(xpmu, nu), Vmu_xpnu = Ul(xpnu, mu), Vnu_x = Ul(site, nu);
    const M3 Vnu_b = Ul(xmnupmu, nu), Vmu_mn = Ul(xmnu, mu), Vnu_mn = Ul(xmnu, nu);
    // forward-staple products (x, x+mu, x+nu)
    const M3 a = mul_nd(Vnu_xpmu, Vmu_xpnu);                       // V_nu(x+mu) V_mu(x+nu)^+
    M3 t = mul_nd(a, Th(site, nu));                                // ... iTheta_nu(x)^+
    const M3 t4 = mul_nd(mul_nd(Th(xpmu, nu), Vmu_xpnu), Vnu_x);   // iTheta_nu(x+mu) V_mu(x+nu)^+ V_nu(x)^+
    const M3 t6 = mul_nd(mul_nd(Vnu_xpmu, Th(xpnu, mu)), Vnu_x);   // V_nu(x+mu) iTheta_mu(x+nu)^+ V_nu(x)^+
    // backward-staple products (x-nu, x-nu+mu)
    const M3 t2 = mul_nn(mul_dn(Vnu_b, dag3(Th(xmnu, mu))), Vnu_mn);  // V_nu^+ iTheta_mu^+ V_nu
    const M3 b = mul_dn(Vnu_b, dag3(Vmu_mn));                          // V_nu(x-nu+mu)^+ V_mu(x-nu)^+
    const M3 t3 = mul_nn(b, Th(xmnu, nu));                             // ... iTheta_nu(x-nu)
    const M3 t5 = mul_nn(mul_dn(Th(xmnupmu, nu), dag3(Vmu_mn)), Vnu_mn); // iTheta_nu^+ V_mu^+ V_nu
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int j = 0; j < 3; ++j) {
        cplx v = cadd(cadd(t.a[i][j], t2.a[i][j]), cadd(t3.a[i][j], t4.a[i][j]));
        v = cadd(v, cadd(t5.a[i][j], t6.a[i][j]));
        acc.a[i][j] = cadd(acc.a[i][j], cscal(rho, v));
      }
  }
  double* out = Sigma + 18 * (site + V * mu);
  M3 s0 = load3(out);
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) s0.a[i][j] = cadd(s0.a[i][j], acc.a[i][j]);
  store3(out, s0);
}

hipError_t launch_stout_force(double* Sigma, double* iTheta, const double* Sigmap,
                              const double* Uin, int Nx, int Ny, int Nz, int Nt, double rho,
                              hipStream_t st) {
  const size_t n = 4 * (size_t)Nx * Ny * Nz * Nt;
  const dim3 grid((unsigned)((n + 127) / 128));
  hipLaunchKernelGGL(k_stout_force_link, grid, dim3(128), 0, st, Sigma, iTheta, Sigmap, Uin, Nx,
                     Ny, Nz, Nt, rho);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_stout_force_staple, grid, dim3(128), 0, st, Sigma, Uin,
                     (const double*)iTheta, Nx, Ny, Nz, Nt, rho);
  return hipGetLastError();
}

// F <- -2 at(U F) with a lexical gauge field U (the thin links of a smeared force)
__global__ void __launch_bounds__(128)
k_mult_generator_gauge(double* __restrict__ F, const double* __restrict__ U, size_t n) {
  const size_t idx = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  if (idx >= n) return;
  const M3 A = at3(mul_nn(load3(U + 18 * idx), load3(F + 18 * idx)));
  M3 r;
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) r.a[i][j] = cscal(-2.0, A.a[i][j]);
  store3(F + 18 * idx, r);
}

hipError_t launch_mult_generator_gauge(double* F, const double* U, size_t nlinks,
                                       hipStream_t st) {
  hipLaunchKernelGGL(k_mult_generator_gauge, dim3((unsigned)((nlinks + 127) / 128)), dim3(128), 0,
                     st, F, U, nlinks);
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
