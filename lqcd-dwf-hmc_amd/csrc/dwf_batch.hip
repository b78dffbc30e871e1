// dwf_batch.hip -- kernels of the batched multishift CG.
//
// Several right-hand sides ("systems": pseudofermion sources, possibly on
// different lattices), each with up to kMsShifts shifts, advance together: one
// launch per stage covers every system (blockIdx.y = system), and the scalar
// recurrences of AShiftsolver_CG::solve_step (ashiftsolver_CG-tmpl.h:216-286)
// run on the device (k_ms_ctrl), so that an iteration needs no host
// synchronisation and can be replayed from a captured graph.
//
// Per iteration, for every active system k:
//   [operator]  s_k = A p0_k = D^+ u_k, u_k = D p0_k   (multi-RHS / grouped hops)
//   k_ms_pa     pa_k = <p0_k, (A + sigma0) p0_k> = |u_k|^2 + sigma0 |p0_k|^2
//               (|p0_k|^2 from the p0 update of the previous iteration)
//   k_ms_xr     x0_k += (rr_k / pa_k) p0_k; r_new = r - (rr_k / pa_k) (s_k + sigma0 p0_k)
//               into the next slot of the system's residual ring; |r_new|^2
//   k_ms_ctrl   truncation of the shifts by the previous |p_j|^2, beta, alpha,
//               zeta_j and the shifts' block coefficients, convergence, and
//               whether this iteration ends a block
//   k_ms_shift  p0_k = r + alpha p0_k, |p0_k|^2; at the end of a block, for the shifts
//               j >= 1 (k_ms_fold's combination):
//               x_kj += b p_kj + sum_l d_l R_l; p_kj = a p_kj + sum_l c_l R_l; |p_kj|^2
// so the shifts' fields are touched once per block of kFoldMax iterations
// (|p_j|^2 in between: |p_j|^2 <- alpha_j^2 |p_j|^2 + zeta_j^2 |r|^2, r being
// orthogonal to the previous p_j).  Reductions are deterministic (block
// partials, then k_ms_finish adds them in a fixed order).
#include "dwf_device.h"
#include "dwf_kernels.h"

namespace dwf {

// two consecutive reals of a field (FP64: one 16-byte access, FP32: 8 bytes)
template <typename T> struct Vec2;
template <> struct Vec2<double> {
  static __device__ __forceinline__ void ld(const double* p, double* o) {
    const dbl2_t v = *reinterpret_cast<const dbl2_t*>(p);
    o[0] = v[0];
    o[1] = v[1];
  }
  static __device__ __forceinline__ void st(double* p, const double* v) {
    *reinterpret_cast<dbl2_t*>(p) = dbl2_t{v[0], v[1]};
  }
};
template <> struct Vec2<float> {
  static __device__ __forceinline__ void ld(const float* p, float* o) {
    const flt2_t v = *reinterpret_cast<const flt2_t*>(p);
    o[0] = v[0];
    o[1] = v[1];
  }
  static __device__ __forceinline__ void st(float* p, const float* v) {
    *reinterpret_cast<flt2_t*>(p) = flt2_t{v[0], v[1]};
  }
};

__device__ __forceinline__ double ms_wave_sum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  return v;
}

// 256-thread block: the block's sum of v, valid in thread 0
__device__ __forceinline__ double ms_block_sum(double v, double* sh) {
  const double t = ms_wave_sum(v);
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = t;
  __syncthreads();
  return sh[0] + sh[1] + sh[2] + sh[3];
}

#define MS_LOOP(n)                                                                     \
  for (size_t i = (blockIdx.x * (size_t)blockDim.x + threadIdx.x) * 2; i < (n);        \
       i += (size_t)gridDim.x * blockDim.x * 2)

// |u|^2, u = D p0 (fields of an even number of reals)
template <typename T>
__global__ void __launch_bounds__(256)
k_ms_pa(const T* const* __restrict__ u, const size_t* __restrict__ n,
        const MsState* __restrict__ st, double* __restrict__ partial) {
  using C = typename St<T>::C;
  const int k = blockIdx.y;
  if (!st[k].active) return;
  const T* uk = u[k];
  double v = 0;
  MS_LOOP(n[k]) {
    C uv[2];
    Vec2<T>::ld(uk + i, uv);
    v += (double)uv[0] * (double)uv[0] + (double)uv[1] * (double)uv[1];
  }
  __shared__ double sh[4];
  const double t = ms_block_sum(v, sh);
  if (threadIdx.x == 0) partial[(size_t)k * gridDim.x + blockIdx.x] = t;
}

// x0 += a p0;  r -= a s;  |r|^2,  a = rr / pa
template <typename T>
__global__ void __launch_bounds__(256)
k_ms_xr(T* const* __restrict__ x0, T* const* __restrict__ ring, T* const* __restrict__ p0,
        T* const* __restrict__ s, const size_t* __restrict__ n, const MsState* __restrict__ st,
        double* __restrict__ partial) {
  using C = typename St<T>::C;
  const int k = blockIdx.y;
  const MsState& S = st[k];
  if (!S.active) return;
  const C a = (C)(S.rr / S.pa);
  const C sg = (C)S.sigma0;
  constexpr int NR = kFoldMax + 1;
  T* xk = x0[k];
  const T* rin = ring[(size_t)k * NR + (S.q0 + S.m) % NR];
  T* rout = ring[(size_t)k * NR + (S.q0 + S.m + 1) % NR];
  const T* pk = p0[k];
  const T* sk = s[k];
  double v = 0;
  MS_LOOP(n[k]) {
    C xv[2], rv[2], pv[2], sv[2];
    Vec2<T>::ld(xk + i, xv);
    Vec2<T>::ld(rin + i, rv);
    Vec2<T>::ld(pk + i, pv);
    Vec2<T>::ld(sk + i, sv);
#pragma unroll
    for (int j = 0; j < 2; ++j) {
      xv[j] += a * pv[j];
      rv[j] = (rv[j] - a * (sv[j] + sg * pv[j]));
      v += (double)rv[j] * (double)rv[j];
    }
    Vec2<T>::st(xk + i, xv);
    Vec2<T>::st(rout + i, rv);
  }
  __shared__ double sh[4];
  const double t = ms_block_sum(v, sh);
  if (threadIdx.x == 0) partial[(size_t)k * gridDim.x + blockIdx.x] = t;
}

// p0 = r + alpha p0 (r the newest residual); at the end of a block also the
// fold of shifts j = 1 .. fnfold - 1 with the block's residuals R_1 .. R_fm
// and coefficients fc[cbuf ^ 1][j]; |p_j|^2 block partials (j = 0: the new p0)
// at partial[((k * kMsShifts) + j) * gridDim.x + block].
template <typename T>
__global__ void __launch_bounds__(256)
k_ms_shift(T* const* __restrict__ x, T* const* __restrict__ p, T* const* __restrict__ p0,
           T* const* __restrict__ ring, const size_t* __restrict__ n,
           const MsState* __restrict__ st, double* __restrict__ partial) {
  using C = typename St<T>::C;
  constexpr int NR = kFoldMax + 1;
  const int k = blockIdx.y;
  const MsState& S = st[k];
  if (!S.upd) return;
  const C alpha = (C)S.alpha;
  T* p0k = p0[k];
  T* const* rk = ring + (size_t)k * NR;
  const T* rc = rk[S.rcur];
  __shared__ double sh[4];
  if (!S.fold) {
    double v0 = 0;
    MS_LOOP(n[k]) {
      C rv[2], qv[2];
      Vec2<T>::ld(rc + i, rv);
      Vec2<T>::ld(p0k + i, qv);
#pragma unroll
      for (int e = 0; e < 2; ++e) {
        qv[e] = rv[e] + alpha * qv[e];
        v0 += (double)qv[e] * (double)qv[e];
      }
      Vec2<T>::st(p0k + i, qv);
    }
    const double t = ms_block_sum(v0, sh);
    if (threadIdx.x == 0) partial[(size_t)k * kMsShifts * gridDim.x + blockIdx.x] = t;
    return;
  }
  const int nsh = S.fnfold, m = S.fm, q0 = S.fq0;
  const double(*fc)[kFoldCoef] = S.fc[S.cbuf ^ 1];
  T* const* xk = x + (size_t)k * kMsShifts;
  T* const* pk = p + (size_t)k * kMsShifts;
  double v[kMsShifts];
#pragma unroll
  for (int j = 0; j < kMsShifts; ++j) v[j] = 0;
  MS_LOOP(n[k]) {
    C rv[kFoldMax][2], qv[2];
#pragma unroll
    for (int l = 0; l < kFoldMax; ++l) {
      if (l < m) Vec2<T>::ld(rk[(q0 + 1 + l) % NR] + i, rv[l]);
      else rv[l][0] = rv[l][1] = C(0);
    }
    Vec2<T>::ld(p0k + i, qv);
    {
      C rn[2];   // the newest residual (R_fm, read again rather than indexed)
      Vec2<T>::ld(rc + i, rn);
#pragma unroll
      for (int e = 0; e < 2; ++e) {
        qv[e] = rn[e] + alpha * qv[e];
        v[0] += (double)qv[e] * (double)qv[e];
      }
    }
    Vec2<T>::st(p0k + i, qv);
#pragma unroll
    for (int j = 1; j < kMsShifts; ++j) {
      if (j < nsh) {
        const double* c = fc[j];
        T* xj = xk[j];
        T* pj = pk[j];
        C xv[2], pv[2];
        Vec2<T>::ld(xj + i, xv);
        Vec2<T>::ld(pj + i, pv);
#pragma unroll
        for (int e = 0; e < 2; ++e) {
          double pn = c[0] * (double)pv[e], xn = (double)xv[e] + c[1] * (double)pv[e];
#pragma unroll
          for (int l = 0; l < kFoldMax; ++l) {
            pn += c[2 + l] * (double)rv[l][e];
            xn += c[2 + kFoldMax + l] * (double)rv[l][e];
          }
          xv[e] = (C)xn;
          pv[e] = (C)pn;
          v[j] += (double)pv[e] * (double)pv[e];
        }
        Vec2<T>::st(xj + i, xv);
        Vec2<T>::st(pj + i, pv);
      }
    }
  }
#pragma unroll
  for (int j = 0; j < kMsShifts; ++j) {
    if (j == 0 || j < nsh) {
      const double t = ms_block_sum(v[j], sh);
      if (threadIdx.x == 0)
        partial[((size_t)k * kMsShifts + j) * gridDim.x + blockIdx.x] = t;
      __syncthreads();
    }
  }
}

// out[k * ostride + o] = sum of the nblk partials of (system k, output o),
// in a fixed order; block (o, k)
__global__ void __launch_bounds__(256)
k_ms_finish(const double* __restrict__ partial, int nblk, int nout, double* out, int ostride) {
  const int o = blockIdx.x, k = blockIdx.y;
  const double* p = partial + ((size_t)k * nout + o) * nblk;
  double w = 0;
  for (int i = threadIdx.x; i < nblk; i += blockDim.x) w += p[i];
  __shared__ double sh[4];
  const double t = ms_block_sum(w, sh);
  if (threadIdx.x == 0) out[(size_t)k * ostride + o] = t;
}
// pa_k = (sum of the nblk partials of system k) + sigma0 |p0_k|^2, |p0_k|^2 =
// pp_res[k][0] (the previous p0 update, or |b_k|^2 at the start); block k
__global__ void __launch_bounds__(256)
k_ms_finish_pa(const double* __restrict__ partial, int nblk, const double* __restrict__ pp_res,
               MsState* __restrict__ st) {
  const int k = blockIdx.x;
  const double* p = partial + (size_t)k * nblk;
  double w = 0;
  for (int i = threadIdx.x; i < nblk; i += blockDim.x) w += p[i];
  __shared__ double sh[4];
  const double t = ms_block_sum(w, sh);
  if (threadIdx.x == 0 && st[k].active) st[k].pa = t + st[k].sigma0 * pp_res[(size_t)k * kMsShifts];
}

// The scalar part of solve_step for every system, one wave per system with
// lane j holding shift j (the values a lane needs are kept in registers and
// the state is written back at the end, so no lane reads what another lane
// wrote), then the status for the host: status[0] = iteration, status[1] =
// active systems (coherent host memory, written after a system-scope fence).
// pp_res: |p_j|^2 of the previous fold (unscaled), [k][kMsShifts].
__global__ void __launch_bounds__(1024)
k_ms_ctrl(MsState* __restrict__ st, const double* __restrict__ pp_res, int nsys, int* counter,
          double* status) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nw = blockDim.x >> 6;
  const int j = lane;
  const bool js = j < kMsShifts;
  __shared__ int nact;
  if (threadIdx.x == 0) nact = 0;
  __syncthreads();
  for (int k = wave; k < nsys; k += nw) {
    MsState& S = st[k];
    // the previous iteration folded: its computed |p_j|^2 replace the estimates
    double ppe = js ? S.ppe[j] : 0.0;
    if (S.fold && j >= 1 && j < S.fnfold) ppe = pp_res[k * kMsShifts + j];
    const int was_active = S.active;
    if (was_active) {
      const double snorm = S.snorm, stop = S.stop;
      int nsh2 = S.nshift2;
      if (S.nconv >= 0) {
        // shift truncation after the previous update (from the top): the
        // highest j < nshift2 with |p_j|^2 / |b|^2 > stop, + 1 (unchanged if none)
        const double ppj = (j == 0) ? S.pp[0] : ppe * snorm;
        const unsigned long long mask = __ballot(j < nsh2 && ppj > stop);
        if (js && j >= 1 && j < nsh2) S.pp[j] = ppj;
        if (mask) nsh2 = 64 - __clzll(mask);
      }
      const double rr_p = S.rr;
      const double beta = -rr_p / S.pa;
      const double rr = S.rr_new;
      const double alpha = rr / rr_p;
      const double alpha_h = 1.0 + S.alpha_p * beta / S.beta_p;
      const int m = S.m + 1;   // this iteration's residual is R_m of the block
      const int cb = S.cbuf;
      if (j >= 1 && j < nsh2) {
        double zeta = (alpha_h - S.csh2[j] * beta) / S.zeta1[j] + (1.0 - alpha_h) / S.zeta2[j];
        zeta = 1.0 / zeta;
        const double zr = zeta / S.zeta1[j];
        const double beta_s = beta * zr, alpha_s = alpha * zr * zr;
        double* c = S.fc[cb][j];
        // x_j -= beta_s p_j (p_j before its update), then p_j = alpha_s p_j + zeta R_m
        c[1] -= beta_s * c[0];
        for (int l = 0; l < m - 1; ++l) c[2 + kFoldMax + l] -= beta_s * c[2 + l];
        c[0] *= alpha_s;
        for (int l = 0; l < m - 1; ++l) c[2 + l] *= alpha_s;
        c[2 + m - 1] = zeta;
        ppe = alpha_s * alpha_s * ppe + zeta * zeta * rr;
        S.zeta2[j] = S.zeta1[j];
        S.zeta1[j] = zeta;
      }
      const int nconv = S.nconv + 1;
      const bool conv = rr * snorm < stop;
      const int active = (conv || nconv + 1 >= S.max_iter) ? 0 : 1;
      const bool fold = (m == kFoldMax) || !active;
      const int q0 = S.q0, nfold = S.nfold;
      const int rcur = (q0 + m) % (kFoldMax + 1);
      if (fold && js) {   // the next block's coefficients start as p_j = p_j'
        double* c = S.fc[cb ^ 1][j];
        for (int e = 0; e < kFoldCoef; ++e) c[e] = (e == 0) ? 1.0 : 0.0;
      }
      if (lane == 0) {
        S.pp[0] = rr;
        S.alpha = alpha;
        S.rr = rr;
        S.alpha_p = alpha;
        S.beta_p = beta;
        S.nconv = nconv;
        S.nshift2 = nsh2;
        S.upd = 1;
        S.rcur = rcur;
        S.active = active;
        if (conv) S.conv = 1;
        if (fold) {
          S.fold = 1;
          S.fq0 = q0;
          S.fm = m;
          S.fnfold = nfold;
          S.cbuf = cb ^ 1;
          S.q0 = rcur;
          S.m = 0;
          S.nfold = nsh2;
        } else {
          S.fold = 0;
          S.m = m;
        }
        if (active) atomicAdd(&nact, 1);
      }
    } else if (lane == 0) {
      S.upd = 0;
      S.fold = 0;
    }
    if (js) S.ppe[j] = ppe;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    const int iter = counter[0];
    counter[0] = iter + 1;
    status[1] = (double)nact;
    __threadfence_system();
    status[0] = (double)iter;
    __threadfence_system();
  }
}

template <typename T>
static hipError_t ms_iter_blas_t(const MsLaunch& L, int stage, hipStream_t st) {
  const dim3 g(L.nblk, L.nsys), b(256);
  switch (stage) {
    case 0:
      hipLaunchKernelGGL(k_ms_pa<T>, g, b, 0, st, (const T* const*)L.u, L.n, L.state, L.partial);
      hipLaunchKernelGGL(k_ms_finish_pa, dim3(L.nsys), b, 0, st, L.partial, L.nblk, L.pp_res,
                         L.state);
      break;
    case 1:
      hipLaunchKernelGGL(k_ms_xr<T>, g, b, 0, st, (T* const*)L.x0, (T* const*)L.ring,
                         (T* const*)L.p0, (T* const*)L.s, L.n, L.state, L.partial);
      hipLaunchKernelGGL(k_ms_finish, dim3(1, L.nsys), b, 0, st, L.partial, L.nblk, 1,
                         &L.state->rr_new, (int)(sizeof(MsState) / sizeof(double)));
      break;
    case 2:
      hipLaunchKernelGGL(k_ms_shift<T>, g, b, 0, st, (T* const*)L.x, (T* const*)L.p,
                         (T* const*)L.p0, (T* const*)L.ring, L.n, L.state, L.partial);
      hipLaunchKernelGGL(k_ms_finish, dim3(kMsShifts, L.nsys), b, 0, st, L.partial, L.nblk,
                         kMsShifts, L.pp_res, kMsShifts);
      break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t launch_ms_blas(int prec, const MsLaunch& L, int stage, hipStream_t st) {
  if (prec == P64) return ms_iter_blas_t<double>(L, stage, st);
  if (prec == P32) return ms_iter_blas_t<float>(L, stage, st);
  return hipErrorInvalidValue;
}

hipError_t launch_ms_ctrl(const MsLaunch& L, hipStream_t st) {
  if (L.nsys > 64) return hipErrorInvalidValue;
  const int waves = L.nsys < 16 ? L.nsys : 16;
  hipLaunchKernelGGL(k_ms_ctrl, dim3(1), dim3(64 * waves), 0, st, L.state, L.pp_res, L.nsys,
                     L.counter, L.status);
  return hipGetLastError();
}

}  // namespace dwf
