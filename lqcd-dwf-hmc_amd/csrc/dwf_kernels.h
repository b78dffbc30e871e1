// dwf_kernels.h -- host-side launchers of the libdwfhip kernels.
//
// Device field layout of a 5D even-odd spinor (one parity):
//   element (site, k, s) at ((site * 24) + k) * Ls + s
//   site in [0, V/2) (even-odd half-site order of the reference), k = ri + 2*(c + 3*d)
//   (Dirac basis), s in [0, Ls).  Ls is innermost so that the Ls threads of one
//   4D site read contiguous memory and share that site's links.
// Device links (per output parity p, hop direction dir = 2*mu + fb):
//   element (p, dir, ab, site) at ((p * 8 + dir) * 18 + ab) * Vh + site
//   fb = 0: U_mu(x) times the boundary factor when x + mu wraps
//   fb = 1: U_mu(x - mu)^dagger times the boundary factor when x - mu wraps
// Neighbour table: nbr[(p * 8 + dir) * Vh + site] = half-site of x +/- mu.
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>

namespace dwf {

enum Prec { P64 = 0, P32 = 1, P16 = 2, PBF16 = 3 };
enum KOp { K_SITE = 0, K_HOP = 1, K_HOPDAG = 2 };

size_t prec_bytes(int prec);

// Epilogue of the fused kernel:  out = xcoef * X + E(core),  out2 = F(out)
// core = hop(in) (K_HOP), gamma5 hop gamma5 (in) (K_HOPDAG) or in (K_SITE).
// E, F: per-chirality Ls x Ls real matrices in the compute type, laid out
// [chi][s_out][s_in] with chi = 0 for P+ and 1 for P-, each already scaled by
// 1/2 (the chiral recombination factor).  E == nullptr means out = scale*core.
struct Epi {
  const void* E = nullptr;
  double scale = 1.0;
  const void* X = nullptr;
  double xcoef = 0.0;
  void* out2 = nullptr;
  const void* F = nullptr;
  const void* G = nullptr;    // with X: out = E core + G X (chiral matrix) instead of + xcoef X
  const void* X2 = nullptr;   // with out2: out2 = F out + X2 (X2 may alias out2)
  double* nrm_partial = nullptr;   // with E: |out|^2 -> nrm_result (block partials, >= Vh doubles)
  double* nrm_result = nullptr;
  double* nrm_host = nullptr;      // with nrm_*: also written to this coherent host memory
  // with X: xcoef = xnum[0] / xden[0], read by the kernel from device memory
  const double* xnum = nullptr;
  const double* xden = nullptr;
  // out == nullptr (Ls = 16, with out2): out is formed for out2 and the norm
  // but not stored.
  // CG updates fused into the epilogue (Ls = 16 only), a = unum[0] / uden[0]:
  // K_SITE with X: upd += a X (besides out, out2);  hop without out2:
  // upd -= a out, written instead of out, with nrm_* taken of the new upd
  void* upd = nullptr;
  const double* unum = nullptr;
  const double* uden = nullptr;
  // nrm_defer: leave the block partials of the norm for a later launch to add
  // (no finishing kernel); nblocks_out receives the launch's block count
  bool nrm_defer = false;
  unsigned* nblocks_out = nullptr;
  // add fin_n partials into *fin_result in the launch's last block (the
  // finishing of an earlier launch's deferred norm)
  const double* fin_partial = nullptr;
  unsigned fin_n = 0;
  double* fin_result = nullptr;
  // several right-hand sides of one lattice in one launch: 2^nrs fields, each
  // rstride reals after the previous one, for in, out, X, out2, X2 and upd;
  // their spinors of one 4D site share the staged links (Vh stays the 4D count)
  int nrs = 0;
  size_t rstride = 0;
};

// Grouped launch: up to kGroupMax lattices (same Ls, precision and epilogue
// form; compile-time Ls 12 or 16; FP64 or FP32) in one launch of the fused
// kernel.  The member table lives in device memory and is built once.
constexpr int kGroupMax = 4;
struct GroupMember {
  void* out = nullptr;
  const void* in = nullptr;
  const void* links = nullptr;
  const int* nbr = nullptr;
  int parity = 0;
  int Vh = 0;
  Epi e;
};
struct GroupLaunch {
  int prec = 0, op = 0, Ls = 0, n = 0;
  bool mat = false, addx = false, out2 = false;
  unsigned grid = 0, block = 0;
  size_t lds = 0;
  void* table = nullptr;   // device memory (owned)
};
// Krylov-block fold of the multishift CG: the shifts' x and p are updated once
// per block of at most kFoldMax iterations from the block's residuals.
constexpr int kFoldMax = 16;
constexpr int kFoldCoef = 2 + 2 * kFoldMax;   // per shift: a, b, c_1..c_K, d_1..d_K
// Batched multishift CG (dwf_batch.hip): device state of one system (one
// right-hand side with its shifts), the scalars of AShiftsolver_CG
// (ashiftsolver_CG-tmpl.h:216-286) and the block fold's bookkeeping.  The
// residuals live in a ring of kFoldMax + 1 fields per system: q0 is the slot of
// the residual at the start of the current block, m the iterations done in it,
// rcur the slot of the newest residual.  fold != 0: this iteration ends a block
// (start slot fq0, length fm, shifts 1..fnfold-1, coefficients fc[cbuf ^ 1]).
constexpr int kMsShifts = 32;
struct MsState {
  double rr, rr_new, pa, alpha_p, beta_p, snorm, sigma0, stop, alpha;
  int active, upd, conv, nshift2, nconv, max_iter;
  int q0, m, rcur, nfold, fold, fq0, fm, fnfold, cbuf, pad_;
  double csh2[kMsShifts], zeta1[kMsShifts], zeta2[kMsShifts], pp[kMsShifts], ppe[kMsShifts];
  double fc[2][kMsShifts][kFoldCoef];
};
// Device pointer tables of one batched solve (all in device memory):
// s, p0, x0, u (= D p0, the operator's intermediate): [nsys]; ring: [nsys][kFoldMax + 1];
// x, p: [nsys][kMsShifts] (slot 0 unused); n: reals per field of each system;
// partial: nsys * kMsShifts * nblk doubles; pp_res: nsys * kMsShifts (slot 0: |p0|^2);
// status: coherent host memory (2 doubles).
struct MsLaunch {
  int nsys = 0, nblk = 0;
  void** s = nullptr;
  void** p0 = nullptr;
  void** x0 = nullptr;
  void** ring = nullptr;
  void** x = nullptr;
  void** p = nullptr;
  void** u = nullptr;
  const size_t* n = nullptr;
  MsState* state = nullptr;
  double* partial = nullptr;
  double* pp_res = nullptr;
  double* status = nullptr;
  int* counter = nullptr;   // iterations done (device), copied to status[0]
};
// stage 0: pa = |u|^2 + sigma0 |p0|^2;  1: x0/r update, rr_new;  2: p0 (and |p0|^2)
// and the shifts
hipError_t launch_ms_blas(int prec, const MsLaunch& L, int stage, hipStream_t st);
hipError_t launch_ms_ctrl(const MsLaunch& L, hipStream_t st);

// result[0..nr) = sums of nblocks block partials of nr values each
// (partial[block * nr + r]), in a fixed order; nr = 1 or 2
hipError_t launch_finish_partials(int nr, const double* partial, unsigned nblocks,
                                  double* result, hipStream_t st);
// parallel-transport convolution network (dwf_ptc.hip), FP32 activations
hipError_t launch_ptc_layer(bool adj, float* out, const float* in, int cin, int cout,
                            const unsigned short* W, int wcin, int wcout, const float* links,
                            const int* nbr, int Vh, int Ls, hipStream_t st);
hipError_t launch_ptc_wgrad(double* G, double* partial, int nblk, const float* g, const float* in,
                            int cin, int cout, const float* links, const int* nbr, int Vh, int Ls,
                            hipStream_t st);
hipError_t launch_ptc_embed(float* act, const float* x, size_t n, hipStream_t st);
hipError_t launch_ptc_residual(float* g, const float* act, const float* t, size_t n,
                               double* partial, unsigned* nblk, hipStream_t st);
// HMC gauge kernels (dwf_hmc.hip), FP64 Field_G arrays [mu][site][18]
hipError_t launch_set_iP(double* P, const double* H, size_t nlinks, double* partial,
                         unsigned* nblk, hipStream_t st);
hipError_t launch_exp_update(double* W, double alpha, const double* P, const double* U,
                             size_t nlinks, int nprec, hipStream_t st);
hipError_t launch_rect_action(const double* U, const int dims[4], double* partial,
                              unsigned* nblk, hipStream_t st);
hipError_t launch_rect_force(double* F, const double* U, const int dims[4], double beta,
                             double c_plaq, double c_rect, hipStream_t st);

hipError_t prepare_dwf_group(int prec, int op, int Ls, int n, const GroupMember* m,
                             GroupLaunch* gl);
hipError_t launch_dwf_group(const GroupLaunch& gl, hipStream_t st);
void free_dwf_group(GroupLaunch* gl);

hipError_t launch_dwf(int prec, int op, void* out, const void* in, const void* links,
                      const int* nbr, int parity, int Vh, int Ls, const Epi& e,
                      hipStream_t st);

hipError_t launch_to_device_layout(int prec, void* out, const double* in, int Vh, int Ls,
                                   hipStream_t st);
hipError_t launch_to_corelib_layout(int prec, double* out, const void* in, int Vh, int Ls,
                                    hipStream_t st);
// U: lexical Field_G (4, V, 18) double on the device.
hipError_t launch_pack_links(int prec, void* links, const double* U, int Nx, int Ny, int Nz,
                             int Nt, int bcx, int bcy, int bcz, int bct, hipStream_t st);

// BLAS-1 on n elements of the storage type (arithmetic in the compute type,
// reductions in double).
hipError_t launch_set_zero(int prec, void* x, size_t n, hipStream_t st);
hipError_t launch_set_value(int prec, void* x, double a, size_t n, hipStream_t st);
hipError_t launch_copy(int prec, void* y, const void* x, size_t n, hipStream_t st);
hipError_t launch_axpy(int prec, void* y, double a, const void* x, size_t n, hipStream_t st);
hipError_t launch_aypx(int prec, double a, void* y, const void* x, size_t n, hipStream_t st);
hipError_t launch_scal(int prec, void* x, double a, size_t n, hipStream_t st);
hipError_t launch_convert(int prec_out, void* y, int prec_in, const void* x, size_t n,
                          hipStream_t st);
// y = a x (x double, y any storage type; one rounding)
hipError_t launch_convert_scaled(int prec_out, void* y, int prec_in, const void* x, double a,
                                 size_t n, hipStream_t st);
// y += a x (y double, x any storage type)
hipError_t launch_axpy_from(int prec_y, void* y, double a, int prec_x, const void* x, size_t n,
                            hipStream_t st);

// Reductions: write to result[0..] (device), using the scratch partial/counter.
struct RedScratch {
  double* partial;   // kRedBlocks * 4 doubles
  unsigned* counter; // 1
  double* result;    // 8
};
#ifndef DWF_RED_BLOCKS
#define DWF_RED_BLOCKS 1024
#endif
constexpr int kRedBlocks = DWF_RED_BLOCKS;
// Reductions: block partials, then a one-block kernel adds them in a fixed order
// (1), or the last block of the same kernel does so behind a fence and an atomic
// counter (0).
#ifndef DWF_RED_TWO_KERNELS
#define DWF_RED_TWO_KERNELS 1
#endif
constexpr bool kRedTwoKernels = DWF_RED_TWO_KERNELS != 0;
hipError_t launch_norm2(int prec, const void* x, size_t n, const RedScratch& r, int slot,
                        hipStream_t st);
hipError_t launch_dot(int prec, const void* x, const void* y, size_t n, const RedScratch& r,
                      int slot, hipStream_t st);
// complex <x, y> (result[slot] = Re, result[slot+1] = Im); n = Vh * 24 * Ls
hipError_t launch_dotc_ls(int prec, const void* x, const void* y, size_t n, int Ls,
                          const RedScratch& r, int slot, hipStream_t st);
// CG step:  x += a p;  r -= a s;  result[slot] = |r|^2
// x += a p, r_out = r_in - a (s + sig p), |r_out|^2 -> result[slot] (r_out != r_in)
hipError_t launch_cg_xr_to(int prec, void* x, void* rout, const void* rin, const void* p,
                           const void* s, double a, double sig, size_t n, const RedScratch& red,
                           int slot, hipStream_t st);
// y = a y + x, |y|^2 -> result[slot]
hipError_t launch_aypx_norm(int prec, double a, void* y, const void* x, size_t n,
                            const RedScratch& red, int slot, hipStream_t st);
hipError_t launch_cg_xr(int prec, void* x, void* r, const void* p, const void* s, double a,
                        size_t n, const RedScratch& red, int slot, hipStream_t st);
// y += a x; result[slot] = Re<y, x>   (used as s += sigma0 p; pap = <s, p>)
hipError_t launch_axpy_dot(int prec, void* y, double a, const void* x, size_t n,
                           const RedScratch& red, int slot, hipStream_t st);
// r = (A x) - b, result[slot] = |r|^2   (r may alias Ax)
hipError_t launch_sub_norm2(int prec, void* r, const void* ax, const void* b, size_t n,
                            const RedScratch& red, int slot, hipStream_t st);

// CG updates with device-resident coefficients a = num[0] / den[0]:
//   x += a p, r -= a s, |r|^2 -> result[slot];   y = a y + x
hipError_t launch_cg_xr_dev(int prec, void* x, void* r, const void* p, const void* s,
                            const double* num, const double* den, size_t n,
                            const RedScratch& red, int slot, hipStream_t st);
hipError_t launch_aypx_dev(int prec, void* y, const void* x, const double* num, const double* den,
                           size_t n, hipStream_t st);
// x += (rr / pap) p,  p = r + (rr_new / rr) p  (scalars in device memory)
hipError_t launch_cg_px_dev(int prec, void* x, void* p, const void* r, const double* rr,
                            const double* pap, const double* rr_new, size_t n, hipStream_t st);

// Multishift update of active shifts j = 1..m-1 (arrays of device pointers on
// the device; coefficients on the device, 3 per shift: -beta_s, alpha_s, zeta):
//   x_j += c0 p_j;  p_j = c1 p_j + c2 r;  pp[j] = |p_j|^2
hipError_t launch_shift_update(void* const* x, void* const* p, const void* r,
                               const double* coef, int m, size_t n, double* partial,
                               unsigned* counters, double* pp, hipStream_t st);

// Krylov-block fold of the multishift CG for shifts j = 1..nsh-1, with the
// block's residuals R[0..m) (m <= kFoldMax) and per-shift coefficients
// [a, b, c_1..c_K, d_1..d_K] (kFoldCoef doubles, K = kFoldMax) on the device:
//   x_j += b p_j + sum_l d_l R_l;  p_j = a p_j + sum_l c_l R_l;  pp[j] = |p_j|^2
hipError_t launch_ms_fold(void* const* x, void* const* p, const void* const* R,
                          const double* coef, int nsh, int m, size_t n, double* partial,
                          double* pp, hipStream_t st);

// v = a0 b + sum_{j<m} c[j] x[j] (double; x and c device arrays)
hipError_t launch_lincomb(void* v, const void* b, double a0, void* const* x, const double* c,
                          int m, size_t n, hipStream_t st);

// Force: F_mu(x) (+)= scale * sum_{s,d} [(1 + sgn g_mu) Vn(x+mu) bc]_{a,d,s} conj(Z(x)_{b,d,s})
// for the sites x of one parity; F lexical Field_G (4, V, 18) double.
hipError_t launch_force_tensor(double* F, const double* Z, const double* Vn, const int* nbr,
                               int parity, int Nx, int Ny, int Nz, int Nt, const int* bc,
                               int Ls, double sgn, double scale, bool accumulate,
                               hipStream_t st);
// F <- -2 at(U F) with U from the packed forward links
hipError_t launch_mult_generator(double* F, const double* links, int Nx, int Ny, int Nz, int Nt,
                                 const int* bc, hipStream_t st);

// Learned preconditioner: x_i = uniform [-1, 1) from a counter hash of (seed, i)
hipError_t launch_fill_uniform(int prec, void* x, size_t n, unsigned long long seed,
                               hipStream_t st);
// Chiral Gram matrices of F <= 6 fields (Ls <= 16), all pairs i <= j (row-major
// over i): out[p][chi][s][t] = sum_{site, k < 12} p_chi(f_i)(s) p_chi(f_j)(t),
// p_+ = f_k + f_{k+12}, p_- = f_k - f_{k+12}.  partial: gram_partial_doubles().
hipError_t launch_chiral_gram(int prec, const void* const* fields, int F, int Vh, int Ls,
                              double* partial, double* out, hipStream_t st);
size_t gram_partial_doubles();

// Gauge: one stout smearing level, lexical Field_G (4, V, 18) double device.
hipError_t launch_stout(double* Uout, const double* Uin, int Nx, int Ny, int Nz, int Nt,
                        double rho, hipStream_t st);
// One level of the stout force recursion: Sigma = dS/dU from Sigmap = dS/dU'
// (U' = stout(U)); iTheta: scratch gauge field.  Lexical Field_G double device.
hipError_t launch_stout_force(double* Sigma, double* iTheta, const double* Sigmap,
                              const double* Uin, int Nx, int Ny, int Nz, int Nt, double rho,
                              hipStream_t st);
// F <- -2 at(U F) over nlinks = 4 V links of a lexical gauge field U
hipError_t launch_mult_generator_gauge(double* F, const double* U, size_t nlinks,
                                       hipStream_t st);
// plaquette sum (un-normalised) into result[slot]
hipError_t launch_plaquette(const double* U, int Nx, int Ny, int Nz, int Nt,
                            const RedScratch& red, int slot, hipStream_t st);

}  // namespace dwf
