// dwf_api.cpp -- the extern "C" ABI of libdwfhip (see include/dwf_hip.h).
//
// Host side of the operator, solvers and smearing.  Each operator mode is a
// short chain of the fused k_dwf kernel (hop or site-local core with a
// per-chirality Ls x Ls epilogue), chosen below so that every Ls-mixing part
// of the even-odd Moebius operator is folded into a neighbouring hop:
//
//   D     = 1 - Di k Hop_eo Mp Di k Hop_oe Mp              (k = -kappa (4 - M0))
//   Ddag  = 1 - Mp^T k Hop^+_eo Di^T Mp^T k Hop^+_oe Di^T
// with Mp the 5th-dimension part of D_eo (afopr_Domainwall_eo-tmpl.h:765-828),
// Di = D_ee^-1 (its L_inv/U_inv, :1022-1117) and Hop^+ = g5 Hop g5.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <utility>
#include <string>
#include <chrono>
#include <vector>

#include "dwf_hip.h"
#include "dwf_kernels.h"

using namespace dwf;

// ---------------------------------------------------------------------------
// error handling
// ---------------------------------------------------------------------------
static thread_local std::string g_err;

static int fail(int code, const std::string& msg) {
  g_err = msg;
  return code;
}

#define HIPCHK(expr)                                                                       \
  do {                                                                                     \
    hipError_t e_ = (expr);                                                                \
    if (e_ != hipSuccess)                                                                  \
      return fail(DWF_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_));        \
  } while (0)

// ---------------------------------------------------------------------------
// library context: stream and reduction scratch
// ---------------------------------------------------------------------------
struct Ctx {
  bool ok = false;
  int device = -1;
  hipStream_t stream = nullptr;
  RedScratch red{};
  double* host_result = nullptr;  // pinned
  // multishift scratch
  double* sh_partial = nullptr;
  unsigned* sh_counters = nullptr;
  double* sh_pp = nullptr;
  double* sh_coef = nullptr;
  void** sh_ptrs = nullptr;       // device array: x[0..n), p[0..n)
  double* host_coef = nullptr;    // pinned
  int sh_cap = 0;
  hipEvent_t ring_ev[2] = {nullptr, nullptr};   // convergence checks of the sync-free CG
  double* host_ring = nullptr;    // coherent pinned memory written by kernels (2 doubles)
};
static Ctx g_ctx;
static std::mutex g_mtx;

// Scratch of dwf_fopr_mult_noalloc: the operator's intermediate fields for
// callers that cannot allocate.  One copy per GPU, held until the process
// ends; host code uses the per-device address from hipGetSymbolAddress, never
// the array's name.
constexpr size_t kNoAllocScratch = (size_t)256 << 20;   // 256 MiB
static __device__ __align__(256) unsigned char scratch[kNoAllocScratch];
constexpr int kMaxDevices = 64;
static void* g_scratch_ptr[kMaxDevices] = {};

// address of this device's scratch array (the current device must be dev)
static int scratch_init(int dev) {
  if (dev < 0 || dev >= kMaxDevices) return fail(DWF_ERR_NO_DEVICE, "device index out of range");
  if (g_scratch_ptr[dev]) return DWF_OK;
  void* ptr = nullptr;
  const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !ptr)
    return fail(DWF_ERR_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
  g_scratch_ptr[dev] = ptr;
  return DWF_OK;
}

static int ctx_init() {
  if (g_ctx.ok) return DWF_OK;
  std::lock_guard<std::mutex> lk(g_mtx);
  if (g_ctx.ok) return DWF_OK;
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n < 1)
    return fail(DWF_ERR_NO_DEVICE, "no HIP device visible");
  HIPCHK(hipGetDevice(&g_ctx.device));
  HIPCHK(hipStreamCreateWithFlags(&g_ctx.stream, hipStreamNonBlocking));
  HIPCHK(hipMalloc(&g_ctx.red.partial, sizeof(double) * kRedBlocks * 4));
  HIPCHK(hipMalloc(&g_ctx.red.counter, sizeof(unsigned) * 4));
  HIPCHK(hipMemset(g_ctx.red.counter, 0, sizeof(unsigned) * 4));
  HIPCHK(hipMalloc(&g_ctx.red.result, sizeof(double) * 16));
  HIPCHK(hipHostMalloc(&g_ctx.host_result, sizeof(double) * 64, hipHostMallocDefault));
  for (auto& e : g_ctx.ring_ev) HIPCHK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
  HIPCHK(hipHostMalloc(&g_ctx.host_ring, sizeof(double) * 8, hipHostMallocCoherent));
  {
    const int rc = scratch_init(g_ctx.device);
    if (rc != DWF_OK) return rc;
  }
  g_ctx.ok = true;
  return DWF_OK;
}

#define CTX()                           \
  do {                                  \
    int rc_ = ctx_init();               \
    if (rc_ != DWF_OK) return rc_;      \
  } while (0)

static int fetch(int n, double* out, int slot = 0) {
  HIPCHK(hipMemcpyAsync(g_ctx.host_result, g_ctx.red.result + slot, sizeof(double) * n,
                        hipMemcpyDeviceToHost, g_ctx.stream));
  HIPCHK(hipStreamSynchronize(g_ctx.stream));
  for (int i = 0; i < n; ++i) out[i] = g_ctx.host_result[i];
  return DWF_OK;
}

static int ensure_shift_scratch(int nshift) {
  if (g_ctx.sh_cap >= nshift) return DWF_OK;
  if (g_ctx.sh_partial) {
    (void)hipFree(g_ctx.sh_partial);
    (void)hipFree(g_ctx.sh_counters);
    (void)hipFree(g_ctx.sh_pp);
    (void)hipFree(g_ctx.sh_coef);
    (void)hipFree(g_ctx.sh_ptrs);
    (void)hipHostFree(g_ctx.host_coef);
  }
  HIPCHK(hipMalloc(&g_ctx.sh_partial, sizeof(double) * kRedBlocks * nshift));
  HIPCHK(hipMalloc(&g_ctx.sh_counters, sizeof(unsigned) * nshift));
  HIPCHK(hipMemset(g_ctx.sh_counters, 0, sizeof(unsigned) * nshift));
  HIPCHK(hipMalloc(&g_ctx.sh_pp, sizeof(double) * nshift));
  HIPCHK(hipMalloc(&g_ctx.sh_coef, sizeof(double) * 3 * nshift));
  HIPCHK(hipMalloc(&g_ctx.sh_ptrs, sizeof(void*) * 2 * nshift));
  HIPCHK(hipHostMalloc(&g_ctx.host_coef, sizeof(double) * (3 * nshift + 2 * nshift + nshift),
                       hipHostMallocDefault));
  g_ctx.sh_cap = nshift;
  return DWF_OK;
}

// ---------------------------------------------------------------------------
// per-chirality Ls x Ls matrices (host, long double)
// ---------------------------------------------------------------------------
struct ChiMat {
  int Ls = 0;
  std::vector<long double> m[2];  // [chi][so * Ls + si]
  explicit ChiMat(int L = 0) : Ls(L) {
    m[0].assign((size_t)L * L, 0.0L);
    m[1].assign((size_t)L * L, 0.0L);
  }
  long double& at(int chi, int so, int si) { return m[chi][(size_t)so * Ls + si]; }
  long double at(int chi, int so, int si) const { return m[chi][(size_t)so * Ls + si]; }
};

static ChiMat cm_identity(int L) {
  ChiMat r(L);
  for (int c = 0; c < 2; ++c)
    for (int s = 0; s < L; ++s) r.at(c, s, s) = 1.0L;
  return r;
}
static ChiMat cm_mul(const ChiMat& a, const ChiMat& b) {
  const int L = a.Ls;
  ChiMat r(L);
  for (int c = 0; c < 2; ++c)
    for (int i = 0; i < L; ++i)
      for (int j = 0; j < L; ++j) {
        long double s = 0;
        for (int k = 0; k < L; ++k) s += a.at(c, i, k) * b.at(c, k, j);
        r.at(c, i, j) = s;
      }
  return r;
}
static ChiMat cm_scale(const ChiMat& a, long double f) {
  ChiMat r = a;
  for (int c = 0; c < 2; ++c)
    for (auto& v : r.m[c]) v *= f;
  return r;
}
static ChiMat cm_transpose(const ChiMat& a) {
  const int L = a.Ls;
  ChiMat r(L);
  for (int c = 0; c < 2; ++c)
    for (int i = 0; i < L; ++i)
      for (int j = 0; j < L; ++j) r.at(c, i, j) = a.at(c, j, i);
  return r;
}
// Gauss-Jordan with partial pivoting
static bool cm_inverse(const ChiMat& a, ChiMat& out) {
  const int L = a.Ls;
  out = ChiMat(L);
  for (int c = 0; c < 2; ++c) {
    std::vector<long double> A(a.m[c]), B((size_t)L * L, 0.0L);
    for (int i = 0; i < L; ++i) B[(size_t)i * L + i] = 1.0L;
    for (int col = 0; col < L; ++col) {
      int piv = col;
      for (int r = col + 1; r < L; ++r)
        if (std::fabs(A[(size_t)r * L + col]) > std::fabs(A[(size_t)piv * L + col])) piv = r;
      if (A[(size_t)piv * L + col] == 0.0L) return false;
      if (piv != col)
        for (int k = 0; k < L; ++k) {
          std::swap(A[(size_t)piv * L + k], A[(size_t)col * L + k]);
          std::swap(B[(size_t)piv * L + k], B[(size_t)col * L + k]);
        }
      const long double inv = 1.0L / A[(size_t)col * L + col];
      for (int k = 0; k < L; ++k) {
        A[(size_t)col * L + k] *= inv;
        B[(size_t)col * L + k] *= inv;
      }
      for (int r = 0; r < L; ++r) {
        if (r == col) continue;
        const long double f = A[(size_t)r * L + col];
        if (f == 0.0L) continue;
        for (int k = 0; k < L; ++k) {
          A[(size_t)r * L + k] -= f * A[(size_t)col * L + k];
          B[(size_t)r * L + k] -= f * B[(size_t)col * L + k];
        }
      }
    }
    out.m[c] = B;
  }
  return true;
}
static bool cm_chirally_symmetric(const ChiMat& a) {
  for (size_t i = 0; i < a.m[0].size(); ++i)
    if (a.m[0][i] != a.m[1][i]) return false;
  return true;
}

// ---------------------------------------------------------------------------
// operator object
// ---------------------------------------------------------------------------
enum MatId {
  M_PRE,        // Mp (site-local prologue when Mp is not spin-scalar)
  M_E1,         // first hop of D
  M_E2,         // second hop of D (with the -1)
  M_DIT,        // Di^T
  M_G1,         // first hop of Ddag: Di^T Mp^T k
  M_G2,         // second hop of Ddag: -Mp^T k
  M_DEO,        // epilogue of the Deo/Doe modes (k Mp when spin-scalar)
  M_DEODAG,     // Mp^T k
  M_DEE, M_DEET, M_DI,
  M_G5, M_G5R,
  M_M5F,        // -1/2 Mp: M5 weight of the force (aforce_F_Domainwall_5din-tmpl.h:355-365)
  M_HE2,        // -M_E2: second hop of H = 1 - D
  M_HG2,        // -M_G2: second hop of H^+ = 1 - D^+
  M_ID,         // 1 (site launches that only add a multiple of a field)
  M_COUNT
};

struct dwf_fopr_s {
  int N[4];
  int Ls;
  int V, Vh;
  int prec;
  double mq, M0, alpha;
  std::vector<double> b, c;
  int bc[4];
  std::string mode = "D";
  bool mp_scalar = false;
  bool deo_is_scalar = false;   // the Deo/Doe epilogue is a multiple of 1
  double deo_scale = 0.0;
  size_t nelem = 0;         // reals per field
  void* links = nullptr;
  int* nbr = nullptr;
  void* mats = nullptr;     // M_COUNT * 2 * Ls * Ls compute-type values
  double* stage = nullptr;  // corelib staging (double), max(field, gauge)
  size_t stage_bytes = 0;
  void* work[5] = {nullptr, nullptr, nullptr, nullptr, nullptr};
  bool have_config = false;
  std::vector<ChiMat> hmats;   // host copies of the matrices (long double)
  double* hop_partial = nullptr;   // block partials of fused norms (Vh doubles)
};

static size_t cbytes(int prec) { return prec == P64 ? 8 : 4; }

static const void* mat(const dwf_fopr_s* op, int id) {
  return (const char*)op->mats + (size_t)id * 2 * op->Ls * op->Ls * cbytes(op->prec);
}

// ChiMats to the device in the compute type of prec, each scaled by 1/2 (the
// chiral recombination factor of the kernel epilogue)
static int upload_chimats(int prec, void* dst, const std::vector<ChiMat>& ms) {
  if (ms.empty()) return DWF_OK;
  const int L = ms[0].Ls;
  const size_t per = 2 * (size_t)L * L;
  if (prec == P64) {
    std::vector<double> h(per * ms.size());
    for (size_t id = 0; id < ms.size(); ++id)
      for (int c = 0; c < 2; ++c)
        for (int i = 0; i < L * L; ++i)
          h[id * per + c * L * L + i] = (double)(0.5L * ms[id].m[c][i]);
    HIPCHK(hipMemcpy(dst, h.data(), h.size() * sizeof(double), hipMemcpyHostToDevice));
  } else {
    std::vector<float> h(per * ms.size());
    for (size_t id = 0; id < ms.size(); ++id)
      for (int c = 0; c < 2; ++c)
        for (int i = 0; i < L * L; ++i)
          h[id * per + c * L * L + i] = (float)(0.5L * ms[id].m[c][i]);
    HIPCHK(hipMemcpy(dst, h.data(), h.size() * sizeof(float), hipMemcpyHostToDevice));
  }
  return DWF_OK;
}

static int upload_mats(dwf_fopr_s* op, const std::vector<ChiMat>& ms) {
  op->hmats = ms;
  return upload_chimats(op->prec, op->mats, ms);
}

// Builds the per-chirality matrices of the operator.  The 5th-dimension
// structure is that of afopr_Domainwall_eo-tmpl.h: D_eo uses
//   w'_s = b_s A_s(w_s) + c_s [ P- w_{s+1} (alpha, or -mq at s = Ls-1)
//                             + P+ w_{s-1} (alpha, or -mq at s = 0) ]
// (:765-828), D_ee the same with b -> b(4-M0)+1, c -> c(4-M0)-1 (:900-959);
// A_s is alpha in the bulk and (alpha P+ + P-) / (P+ + alpha P-) at s = 0 /
// Ls-1.  Chirality 0 is P+ (gamma_5 = +1), 1 is P-.
static int build_mats(dwf_fopr_s* op) {
  const int L = op->Ls;
  const long double al = op->alpha, mq = op->mq, M0 = op->M0;
  ChiMat Mp(L), Dee(L);
  for (int s = 0; s < L; ++s) {
    const long double Ap = (s == 0) ? al : (s == L - 1 ? 1.0L : al);
    const long double Am = (s == 0) ? 1.0L : (s == L - 1 ? al : al);
    const long double cdn = (s == 0) ? -mq : al;      // coefficient of P+ w_{s-1}
    const long double cup = (s == L - 1) ? -mq : al;  // coefficient of P- w_{s+1}
    const int sd = (s - 1 + L) % L, su = (s + 1) % L;
    const long double b = op->b[s], c = op->c[s];
    Mp.at(0, s, s) += b * Ap;
    Mp.at(0, s, sd) += c * cdn;
    Mp.at(1, s, s) += b * Am;
    Mp.at(1, s, su) += c * cup;
    const long double F1 = b * (4.0L - M0) + 1.0L, F2 = c * (4.0L - M0) - 1.0L;
    Dee.at(0, s, s) += F1 * Ap;
    Dee.at(0, s, sd) += F2 * cdn;
    Dee.at(1, s, s) += F1 * Am;
    Dee.at(1, s, su) += F2 * cup;
  }
  ChiMat Di;
  if (!cm_inverse(Dee, Di)) return fail(DWF_ERR_INVALID_ARG, "D_ee is singular");
  const long double kappa = 1.0L / (8.0L - 2.0L * M0);
  const long double k = -kappa * (4.0L - M0);
  op->mp_scalar = cm_chirally_symmetric(Mp);
  const ChiMat I = cm_identity(L);
  const ChiMat MpT = cm_transpose(Mp), DiT = cm_transpose(Di);
  std::vector<ChiMat> ms(M_COUNT, ChiMat(L));
  ms[M_PRE] = Mp;
  if (op->mp_scalar) {
    ms[M_E1] = cm_scale(cm_mul(Di, Mp), k);
    ms[M_E2] = cm_scale(cm_mul(Di, Mp), -k);
    ms[M_DEO] = cm_scale(Mp, k);
  } else {
    ms[M_E1] = cm_scale(cm_mul(Mp, Di), k);
    ms[M_E2] = cm_scale(Di, -k);
    ms[M_DEO] = cm_scale(I, k);
  }
  ms[M_DIT] = DiT;
  // when the Deo epilogue is s * 1 (Shamir-type, uniform b), the hop kernel
  // applies the scalar instead of the Ls x Ls matrix
  {
    const ChiMat& E = ms[M_DEO];
    const long double d0 = E.at(0, 0, 0);
    bool scalar = true;
    for (int c = 0; c < 2 && scalar; ++c)
      for (int i = 0; i < L && scalar; ++i)
        for (int j = 0; j < L; ++j)
          if (E.at(c, i, j) != (i == j ? d0 : 0.0L)) {
            scalar = false;
            break;
          }
    op->deo_is_scalar = scalar;
    op->deo_scale = (double)d0;
  }
  ms[M_G1] = cm_scale(cm_mul(DiT, MpT), k);
  ms[M_G2] = cm_scale(MpT, -k);
  ms[M_DEODAG] = cm_scale(MpT, k);
  ms[M_DEE] = Dee;
  ms[M_DEET] = cm_transpose(Dee);
  ms[M_DI] = Di;
  ChiMat G5(L), G5R(L);
  for (int s = 0; s < L; ++s) {
    G5.at(0, s, s) = 1.0L;
    G5.at(1, s, s) = -1.0L;
    G5R.at(0, s, L - 1 - s) = 1.0L;
    G5R.at(1, s, L - 1 - s) = -1.0L;
  }
  ms[M_G5] = G5;
  ms[M_G5R] = G5R;
  ms[M_M5F] = cm_scale(Mp, -0.5L);
  ms[M_HE2] = cm_scale(ms[M_E2], -1.0L);
  ms[M_HG2] = cm_scale(ms[M_G2], -1.0L);
  ChiMat Id(L);
  for (int s = 0; s < L; ++s) Id.at(0, s, s) = Id.at(1, s, s) = 1.0L;
  ms[M_ID] = Id;
  return upload_mats(op, ms);
}

static int build_nbr(dwf_fopr_s* op) {
  const int Nx = op->N[0], Ny = op->N[1], Nz = op->N[2], Nt = op->N[3];
  const int Nx2 = Nx / 2, Vh = op->Vh;
  std::vector<int> h((size_t)2 * 8 * Vh);
  for (int p = 0; p < 2; ++p)
    for (int s2 = 0; s2 < Vh; ++s2) {
      int r = s2 / Nx2;
      const int x2 = s2 - r * Nx2;
      const int y = r % Ny;
      r /= Ny;
      const int z = r % Nz;
      const int t = r / Nz;
      const int x = 2 * x2 + ((y + z + t + p) & 1);
      const int xc[4] = {x, y, z, t};
      for (int mu = 0; mu < 4; ++mu)
        for (int fb = 0; fb < 2; ++fb) {
          int n[4] = {xc[0], xc[1], xc[2], xc[3]};
          n[mu] = (n[mu] + (fb == 0 ? 1 : -1) + op->N[mu]) % op->N[mu];
          h[((size_t)p * 8 + 2 * mu + fb) * Vh + s2] =
              n[0] / 2 + Nx2 * (n[1] + Ny * (n[2] + Nz * n[3]));
        }
    }
  HIPCHK(hipMemcpy(op->nbr, h.data(), h.size() * sizeof(int), hipMemcpyHostToDevice));
  return DWF_OK;
}

static int get_work(dwf_fopr_s* op, int i, void** out) {
  if (!op->work[i]) HIPCHK(hipMalloc(&op->work[i], op->nelem * prec_bytes(op->prec)));
  *out = op->work[i];
  return DWF_OK;
}

// launch helpers ------------------------------------------------------------
static int run_site(dwf_fopr_s* op, void* out, const void* in, int mid, void* out2 = nullptr,
                    int fid = -1) {
  Epi e;
  e.E = mat(op, mid);
  if (out2) {
    e.out2 = out2;
    e.F = mat(op, fid);
  }
  HIPCHK(launch_dwf(op->prec, K_SITE, out, in, op->links, op->nbr, 0, op->Vh, op->Ls, e,
                    g_ctx.stream));
  return DWF_OK;
}
static int run_hop(dwf_fopr_s* op, bool dag, int parity, void* out, const void* in, int mid,
                   const void* X = nullptr, double xcoef = 0.0, void* out2 = nullptr,
                   int fid = -1) {
  Epi e;
  e.E = mat(op, mid);
  e.X = X;
  e.xcoef = xcoef;
  if (out2) {
    e.out2 = out2;
    e.F = mat(op, fid);
  }
  HIPCHK(launch_dwf(op->prec, dag ? K_HOPDAG : K_HOP, out, in, op->links, op->nbr, parity,
                    op->Vh, op->Ls, e, g_ctx.stream));
  return DWF_OK;
}

#define RC(expr)                  \
  do {                            \
    int rc__ = (expr);            \
    if (rc__ != DWF_OK) return rc__; \
  } while (0)

// D v = w - Di k Hop_eo Mp Di k Hop_oe Mp w; optional v2 = Di^T v (fused)
static int op_D(dwf_fopr_s* op, void* v, const void* w, void* v2 = nullptr) {
  void *t0, *t1;
  RC(get_work(op, 0, &t0));
  RC(get_work(op, 1, &t1));
  const void* src = w;
  if (!op->mp_scalar) {
    RC(run_site(op, t0, w, M_PRE));
    src = t0;
  }
  RC(run_hop(op, false, 1, t1, src, M_E1));
  RC(run_hop(op, false, 0, v, t1, M_E2, w, 1.0, v2, M_DIT));
  return DWF_OK;
}
// Ddag v = w - Mp^T k Hop+_eo Di^T Mp^T k Hop+_oe Di^T w; pre = Di^T w if known
static int op_Ddag(dwf_fopr_s* op, void* v, const void* w, const void* pre = nullptr) {
  void *t0, *t1;
  RC(get_work(op, 0, &t0));
  RC(get_work(op, 1, &t1));
  if (!pre) {
    RC(run_site(op, t0, w, M_DIT));
    pre = t0;
  }
  RC(run_hop(op, true, 1, t1, pre, M_G1));
  RC(run_hop(op, true, 0, v, t1, M_G2, w, 1.0));
  return DWF_OK;
}
// Ls = 16 and 12 (the matrix-core epilogue): u = D w need not be stored in
// DdagD.  The last hop of D forms u for u2 = Di^T u (and a norm) without storing
// it, and the last hop of Ddag adds u as Dee^T u2 through its second chiral
// matrix, which saves one field write per DdagD.
static bool u_unstored(const dwf_fopr_s* op) { return op->Ls == 16 || op->Ls == 12; }
// Ddag u from u2 = Di^T u alone (its prologue already done, u not stored)
static int op_Ddag_u2(dwf_fopr_s* op, void* v, const void* u2) {
  void* t1;
  RC(get_work(op, 1, &t1));
  RC(run_hop(op, true, 1, t1, u2, M_G1));
  Epi g;
  g.E = mat(op, M_G2);
  g.X = u2;
  g.G = mat(op, M_DEET);
  HIPCHK(launch_dwf(op->prec, K_HOPDAG, v, t1, op->links, op->nbr, 0, op->Vh, op->Ls, g,
                    g_ctx.stream));
  return DWF_OK;
}
static int op_DdagD(dwf_fopr_s* op, void* v, const void* w) {
  void *u, *u2;
  RC(get_work(op, 3, &u2));
  if (u_unstored(op)) {
    RC(op_D(op, nullptr, w, u2));
    return op_Ddag_u2(op, v, u2);
  }
  RC(get_work(op, 2, &u));
  RC(op_D(op, u, w, u2));
  RC(op_Ddag(op, v, u, u2));
  return DWF_OK;
}
// DdagD with |D w|^2 (= <w, DdagD w>) reduced into the device scalar *nrm: the
// norm is fused into the epilogue of the last hop of D
// mpw: Mp w already computed by the caller (non-spin-scalar Mp), or nullptr.
static int op_DdagD_nrm(dwf_fopr_s* op, void* v, const void* w, double* nrm,
                        const void* mpw = nullptr) {
  void *t0, *t1, *u, *u2;
  RC(get_work(op, 0, &t0));
  RC(get_work(op, 1, &t1));
  RC(get_work(op, 2, &u));
  RC(get_work(op, 3, &u2));
  if (!op->hop_partial) HIPCHK(hipMalloc(&op->hop_partial, (size_t)op->Vh * sizeof(double)));
  const void* src = w;
  if (mpw) {
    src = mpw;
  } else if (!op->mp_scalar) {
    RC(run_site(op, t0, w, M_PRE));
    src = t0;
  }
  RC(run_hop(op, false, 1, t1, src, M_E1));
  Epi e;
  e.E = mat(op, M_E2);
  e.X = w;
  e.xcoef = 1.0;
  e.out2 = u2;
  e.F = mat(op, M_DIT);
  e.nrm_partial = op->hop_partial;
  e.nrm_result = nrm;
  const bool keep_u = !u_unstored(op);
  HIPCHK(launch_dwf(op->prec, K_HOP, keep_u ? u : nullptr, t1, op->links, op->nbr, 0, op->Vh,
                    op->Ls, e, g_ctx.stream));
  return keep_u ? op_Ddag(op, v, u, u2) : op_Ddag_u2(op, v, u2);
}
// One CG step's operator part with the residual update fused into the last
// hop (Ls = 16): mpw = Mp p (non-spin-scalar Mp) or p itself (Mp folded into
// the hops); |D p|^2 -> *pap; then r -= (*rr / *pap) DdagD p, and |r|^2 of the
// new r -> *rr_new (and *rr_host, coherent host memory).  DdagD p itself is
// not stored.
// |D p|^2 is added up by the last block of the third hop (no finishing launch).
static int op_DdagD_cg(dwf_fopr_s* op, const void* p, const void* mpw, double* pap, void* r,
                       const double* rr, double* rr_new, double* rr_host) {
  void *t1, *u2;
  RC(get_work(op, 1, &t1));
  RC(get_work(op, 3, &u2));
  if (!op->hop_partial) HIPCHK(hipMalloc(&op->hop_partial, (size_t)op->Vh * sizeof(double)));
  RC(run_hop(op, false, 1, t1, mpw, M_E1));
  // u = D p is needed only through u2 = Di^T u (the prologue of D^dag) and as
  // the added term of the last hop, which takes it as Dee^T u2: u is never stored
  Epi e;
  e.E = mat(op, M_E2);
  e.X = p;
  e.xcoef = 1.0;
  e.out2 = u2;
  e.F = mat(op, M_DIT);
  e.nrm_partial = op->hop_partial;
  e.nrm_result = pap;
  unsigned nb2 = 0;
  e.nrm_defer = true;   // added up by the last block of the next hop
  e.nblocks_out = &nb2;
  HIPCHK(launch_dwf(op->prec, K_HOP, nullptr, t1, op->links, op->nbr, 0, op->Vh, op->Ls, e,
                    g_ctx.stream));
  {
    Epi h;   // first hop of D^dag; its last block adds up |D p|^2
    h.E = mat(op, M_G1);
    h.fin_partial = op->hop_partial;
    h.fin_n = nb2;
    h.fin_result = pap;
    HIPCHK(launch_dwf(op->prec, K_HOPDAG, t1, u2, op->links, op->nbr, 1, op->Vh, op->Ls, h,
                      g_ctx.stream));
  }
  Epi g;
  g.E = mat(op, M_G2);
  g.X = u2;
  g.G = mat(op, M_DEET);
  g.upd = r;
  g.unum = rr;
  g.uden = pap;
  g.nrm_partial = op->hop_partial;
  g.nrm_result = rr_new;
  g.nrm_host = rr_host;
  HIPCHK(launch_dwf(op->prec, K_HOPDAG, nullptr, t1, op->links, op->nbr, 0, op->Vh, op->Ls, g,
                    g_ctx.stream));
  return DWF_OK;
}
static int op_DDdag(dwf_fopr_s* op, void* v, const void* w) {
  void* u;
  RC(get_work(op, 2, &u));
  RC(op_Ddag(op, u, w));
  RC(op_D(op, v, u));
  return DWF_OK;
}
static int op_H(dwf_fopr_s* op, void* v, const void* w) {
  void* u;
  RC(get_work(op, 2, &u));
  RC(op_D(op, u, w));
  RC(run_site(op, v, u, M_G5R));
  return DWF_OK;
}
static int op_Hdag(dwf_fopr_s* op, void* v, const void* w) {
  void *u, *u2;
  RC(get_work(op, 2, &u));
  RC(get_work(op, 3, &u2));
  RC(run_site(op, u, w, M_G5R, u2, M_DIT));
  RC(op_Ddag(op, v, u, u2));
  return DWF_OK;
}
// D_eo / D_oe (mult modes "Deo" = output even, "Doe" = output odd)
static int op_Deo(dwf_fopr_s* op, void* v, const void* w, int parity) {
  const void* src = w;
  if (!op->mp_scalar) {
    void* t0;
    RC(get_work(op, 0, &t0));
    RC(run_site(op, t0, w, M_PRE));
    src = t0;
  }
  if (op->deo_is_scalar) {
    Epi e;
    e.scale = op->deo_scale;
    HIPCHK(launch_dwf(op->prec, K_HOP, v, src, op->links, op->nbr, parity, op->Vh, op->Ls, e,
                      g_ctx.stream));
    return DWF_OK;
  }
  RC(run_hop(op, false, parity, v, src, M_DEO));
  return DWF_OK;
}

static bool hop_last(const std::string& m) {
  return m == "D" || m == "Ddag" || m == "DdagD" || m == "DDdag" || m == "H" || m == "Hdag" ||
         m == "Deo" || m == "Doe";
}

static int mult_impl(dwf_fopr_s* op, void* v, const void* w, const std::string& m, bool dag) {
  if (!op->have_config) return fail(DWF_ERR_NO_CONFIG, "set_config has not been called");
  if (!v || !w) return fail(DWF_ERR_INVALID_ARG, "null field");
  // kernels that gather neighbours must not write their own input: an aliased
  // input is first copied to a buffer of its own
  const void* src = w;
  if (v == w && hop_last(m)) {
    void* tmp;
    RC(get_work(op, 4, &tmp));
    HIPCHK(launch_copy(op->prec, tmp, w, op->nelem, g_ctx.stream));
    src = tmp;
  }
  int rc = DWF_OK;
  if (!dag) {
    if (m == "D") rc = op_D(op, v, src);
    else if (m == "Ddag") rc = op_Ddag(op, v, src);
    else if (m == "DdagD") rc = op_DdagD(op, v, src);
    else if (m == "DDdag") rc = op_DDdag(op, v, src);
    else if (m == "H") rc = op_H(op, v, src);
    else if (m == "Hdag") rc = op_Hdag(op, v, src);
    else if (m == "Deo") rc = op_Deo(op, v, src, 0);
    else if (m == "Doe") rc = op_Deo(op, v, src, 1);
    else if (m == "Dee" || m == "Doo") rc = run_site(op, v, src, M_DEE);
    else if (m == "Dee_inv" || m == "Doo_inv") rc = run_site(op, v, src, M_DI);
    else rc = fail(DWF_ERR_UNKNOWN_MODE, "mode undefined: " + m);
  } else {
    if (m == "D") rc = op_Ddag(op, v, src);
    else if (m == "Ddag") rc = op_D(op, v, src);
    else if (m == "DdagD") rc = op_DdagD(op, v, src);
    else if (m == "DDdag") rc = op_DDdag(op, v, src);
    else if (m == "H") rc = op_Hdag(op, v, src);
    else if (m == "Hdag") rc = op_H(op, v, src);
    else if (m == "Deo") rc = run_hop(op, true, 1, v, src, M_DEODAG);   // Ddag_eo(v, w, 1)
    else if (m == "Doe") rc = run_hop(op, true, 0, v, src, M_DEODAG);   // Ddag_eo(v, w, 0)
    else if (m == "Dee" || m == "Doo") rc = run_site(op, v, src, M_DEET);
    else if (m == "Dee_inv" || m == "Doo_inv") rc = run_site(op, v, src, M_DIT);
    else rc = fail(DWF_ERR_UNKNOWN_MODE, "mode undefined: " + m);
  }
  return rc;
}

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------
extern "C" {

const char* dwf_last_error(void) { return g_err.c_str(); }
const char* dwf_version(void) { return "dwfhip 0.1 (gfx950)"; }
void* dwf_stream(void) {
  if (ctx_init() != DWF_OK) return nullptr;
  return (void*)g_ctx.stream;
}
int dwf_synchronize(void) {
  CTX();
  HIPCHK(hipStreamSynchronize(g_ctx.stream));
  return DWF_OK;
}

int dwf_fopr_create(const dwf_fopr_params* p, dwf_fopr* out) {
  CTX();
  if (!p || !out) return fail(DWF_ERR_INVALID_ARG, "null argument");
  *out = nullptr;
  for (int mu = 0; mu < 4; ++mu)
    if (p->lattice_size[mu] < 2 || (p->lattice_size[mu] & 1))
      return fail(DWF_ERR_INVALID_ARG, "lattice extents must be even and >= 2");
  const int Ls = p->extent_of_5th_dimension;
  if (Ls < 2 || Ls > 48) return fail(DWF_ERR_UNSUPPORTED, "extent_of_5th_dimension must be in [2, 48]");
  if (p->precision < DWF_F64 || p->precision > DWF_BF16)
    return fail(DWF_ERR_INVALID_ARG, "unknown precision");
  if (p->domain_wall_height == 4.0) return fail(DWF_ERR_INVALID_ARG, "M0 = 4 is singular");
  dwf_fopr_s* op = new dwf_fopr_s;
  for (int mu = 0; mu < 4; ++mu) {
    op->N[mu] = p->lattice_size[mu];
    op->bc[mu] = p->boundary_condition[mu];
  }
  op->Ls = Ls;
  op->V = op->N[0] * op->N[1] * op->N[2] * op->N[3];
  op->Vh = op->V / 2;
  op->prec = p->precision;
  op->mq = p->quark_mass;
  op->M0 = p->domain_wall_height;
  op->alpha = p->parameter_alpha == 0.0 ? 1.0 : p->parameter_alpha;
  op->b.resize(Ls);
  op->c.resize(Ls);
  for (int s = 0; s < Ls; ++s) {
    op->b[s] = p->coefficient_b ? p->coefficient_b[s] : p->b_uniform;
    op->c[s] = p->coefficient_c ? p->coefficient_c[s] : p->c_uniform;
  }
  op->nelem = (size_t)op->Vh * 24 * Ls;
  auto cleanup = [&](int rc) {
    dwf_fopr_destroy(op);
    return rc;
  };
  if (hipMalloc(&op->links, (size_t)2 * 8 * 18 * op->Vh * prec_bytes(op->prec)) != hipSuccess ||
      hipMalloc((void**)&op->nbr, (size_t)2 * 8 * op->Vh * sizeof(int)) != hipSuccess ||
      hipMalloc(&op->mats, (size_t)M_COUNT * 2 * Ls * Ls * cbytes(op->prec)) != hipSuccess)
    return cleanup(fail(DWF_ERR_HIP, "device allocation failed"));
  op->stage_bytes = std::max(op->nelem, (size_t)op->V * 4 * 18) * sizeof(double);
  if (hipMalloc(&op->stage, op->stage_bytes) != hipSuccess)
    return cleanup(fail(DWF_ERR_HIP, "device allocation failed"));
  int rc = build_nbr(op);
  if (rc != DWF_OK) return cleanup(rc);
  rc = build_mats(op);
  if (rc != DWF_OK) return cleanup(rc);
  *out = op;
  return DWF_OK;
}

int dwf_fopr_destroy(dwf_fopr op) {
  if (!op) return DWF_OK;
  if (g_ctx.stream) (void)hipStreamSynchronize(g_ctx.stream);
  if (op->links) (void)hipFree(op->links);
  if (op->nbr) (void)hipFree(op->nbr);
  if (op->mats) (void)hipFree(op->mats);
  if (op->stage) (void)hipFree(op->stage);
  if (op->hop_partial) (void)hipFree(op->hop_partial);
  for (auto& w : op->work)
    if (w) (void)hipFree(w);
  delete op;
  return DWF_OK;
}

int dwf_fopr_set_config_device(dwf_fopr op, const double* U) {
  CTX();
  if (!op || !U) return fail(DWF_ERR_INVALID_ARG, "null argument");
  HIPCHK(launch_pack_links(op->prec, op->links, U, op->N[0], op->N[1], op->N[2], op->N[3],
                           op->bc[0], op->bc[1], op->bc[2], op->bc[3], g_ctx.stream));
  op->have_config = true;
  return DWF_OK;
}

int dwf_fopr_set_config(dwf_fopr op, const double* U_host) {
  CTX();
  if (!op || !U_host) return fail(DWF_ERR_INVALID_ARG, "null argument");
  const size_t bytes = (size_t)op->V * 4 * 18 * sizeof(double);
  HIPCHK(hipMemcpyAsync(op->stage, U_host, bytes, hipMemcpyHostToDevice, g_ctx.stream));
  RC(dwf_fopr_set_config_device(op, op->stage));
  HIPCHK(hipStreamSynchronize(g_ctx.stream));
  return DWF_OK;
}

int dwf_fopr_set_mode(dwf_fopr op, const char* mode) {
  if (!op || !mode) return fail(DWF_ERR_INVALID_ARG, "null argument");
  static const char* known[] = {"D", "Ddag", "DdagD", "DDdag", "H", "Hdag", "Deo", "Doe",
                                "Dee", "Doo", "Dee_inv", "Doo_inv"};
  for (const char* k : known)
    if (std::strcmp(k, mode) == 0) {
      op->mode = mode;
      return DWF_OK;
    }
  return fail(DWF_ERR_UNKNOWN_MODE, std::string("mode undefined: ") + mode);
}
const char* dwf_fopr_get_mode(dwf_fopr op) { return op ? op->mode.c_str() : ""; }

int dwf_fopr_mult(dwf_fopr op, void* v, const void* w) {
  CTX();
  if (!op) return fail(DWF_ERR_INVALID_ARG, "null operator");
  return mult_impl(op, v, w, op->mode, false);
}
// mult(v, w) with the intermediate fields in the library's static scratch
// (5 fields of 256-byte aligned slices): no allocation.  Calls that overlap on
// one GPU share the scratch and race.
int dwf_fopr_mult_noalloc(dwf_fopr op, void* v, const void* w) {
  if (!g_ctx.ok) return fail(DWF_ERR_NO_DEVICE, "library not initialised (create an operator first)");
  if (!op) return fail(DWF_ERR_INVALID_ARG, "null operator");
  int dev = -1;
  HIPCHK(hipGetDevice(&dev));
  if (dev != g_ctx.device)
    return fail(DWF_ERR_INVALID_ARG, "the current device is not the library's device");
  RC(scratch_init(dev));
  const size_t fb = (op->nelem * prec_bytes(op->prec) + 255) & ~(size_t)255;
  if (5 * fb > kNoAllocScratch)
    return fail(DWF_ERR_INVALID_ARG,
                "dwf_fopr_mult_noalloc: needs 5 fields of " + std::to_string(fb) +
                    " bytes of scratch, the maximum is " + std::to_string(kNoAllocScratch) +
                    " bytes in total");
  unsigned char* base = static_cast<unsigned char*>(g_scratch_ptr[dev]);
  void* saved[5];
  for (int i = 0; i < 5; ++i) {
    saved[i] = op->work[i];
    op->work[i] = base + i * fb;
  }
  const int rc = mult_impl(op, v, w, op->mode, false);
  for (int i = 0; i < 5; ++i) op->work[i] = saved[i];
  return rc;
}
int dwf_fopr_mult_dag(dwf_fopr op, void* v, const void* w) {
  CTX();
  if (!op) return fail(DWF_ERR_INVALID_ARG, "null operator");
  return mult_impl(op, v, w, op->mode, true);
}
int dwf_fopr_mult_mode(dwf_fopr op, void* v, const void* w, const char* mode) {
  CTX();
  if (!op || !mode) return fail(DWF_ERR_INVALID_ARG, "null argument");
  std::string m(mode);
  if (m != "Deo" && m != "Doe" && m != "Dee" && m != "Doo" && m != "Dee_inv" && m != "Doo_inv")
    return fail(DWF_ERR_UNKNOWN_MODE, "mode undefined for mult(v, w, mode): " + m);
  return mult_impl(op, v, w, m, false);
}
int dwf_fopr_mult_dag_mode(dwf_fopr op, void* v, const void* w, const char* mode) {
  CTX();
  if (!op || !mode) return fail(DWF_ERR_INVALID_ARG, "null argument");
  std::string m(mode);
  if (m != "Deo" && m != "Doe" && m != "Dee" && m != "Doo" && m != "Dee_inv" && m != "Doo_inv")
    return fail(DWF_ERR_UNKNOWN_MODE, "mode undefined for mult_dag(v, w, mode): " + m);
  return mult_impl(op, v, w, m, true);
}
int dwf_fopr_mult_gm5(dwf_fopr op, void* v, const void* w) {
  CTX();
  if (!op || !v || !w) return fail(DWF_ERR_INVALID_ARG, "null argument");
  return run_site(op, v, w, M_G5);
}
int dwf_fopr_mult_gm5R(dwf_fopr op, void* v, const void* w) {
  CTX();
  if (!op || !v || !w) return fail(DWF_ERR_INVALID_ARG, "null argument");
  if (v == w) return fail(DWF_ERR_INVALID_ARG, "mult_gm5R needs distinct fields");
  return run_site(op, v, w, M_G5R);
}

int dwf_fopr_field_nin(dwf_fopr op) { return op ? 24 * op->Ls : 0; }
int dwf_fopr_field_nvol(dwf_fopr op) { return op ? op->Vh : 0; }
int dwf_fopr_field_nex(dwf_fopr op) { return op ? 1 : 0; }
size_t dwf_fopr_field_bytes(dwf_fopr op) { return op ? op->nelem * prec_bytes(op->prec) : 0; }
int dwf_fopr_precision(dwf_fopr op) { return op ? op->prec : -1; }

double dwf_fopr_flop_count(dwf_fopr op, const char* mode) {
  // afopr_Domainwall_5din_eo-tmpl.h:1196-1250 (Dirac representation)
  if (!op || !mode) return 0.0;
  const double vsite = op->V, Ns = op->Ls, Nc = 3, Nd = 4;
  const double hop = Ns * Nc * Nd * (6 * (4 * Nc + 2) + 2 * (4 * Nc + 1)) * vsite;
  const double pre = Ns * Nc * Nd * 14 * vsite;
  const double lu = Nc * Nd * (2 + (Ns - 1) * 26) * vsite;
  const double axpy = Ns * 2 * Nc * Nd * vsite;
  const double deo = hop + pre;
  std::string m(mode);
  if (m == "D" || m == "Ddag") return deo + lu + 0.5 * axpy;
  if (m == "DdagD") return 2 * deo + 2 * lu + axpy;
  if (m == "Dee_inv" || m == "Doo_inv") return 0.5 * lu;
  if (m == "Dee" || m == "Doo") return 0.5 * pre;
  if (m == "Deo" || m == "Doe") return 0.5 * deo;
  return 0.0;
}

int dwf_fopr_convert_device(dwf_fopr op, void* v, const double* w) {
  CTX();
  if (!op || !v || !w) return fail(DWF_ERR_INVALID_ARG, "null argument");
  HIPCHK(launch_to_device_layout(op->prec, v, w, op->Vh, op->Ls, g_ctx.stream));
  return DWF_OK;
}
int dwf_fopr_reverse_device(dwf_fopr op, double* v, const void* w) {
  CTX();
  if (!op || !v || !w) return fail(DWF_ERR_INVALID_ARG, "null argument");
  HIPCHK(launch_to_corelib_layout(op->prec, v, w, op->Vh, op->Ls, g_ctx.stream));
  return DWF_OK;
}
int dwf_fopr_convert(dwf_fopr op, void* v, const double* w_host) {
  CTX();
  if (!op || !v || !w_host) return fail(DWF_ERR_INVALID_ARG, "null argument");
  HIPCHK(hipMemcpyAsync(op->stage, w_host, op->nelem * sizeof(double), hipMemcpyHostToDevice,
                        g_ctx.stream));
  RC(dwf_fopr_convert_device(op, v, op->stage));
  HIPCHK(hipStreamSynchronize(g_ctx.stream));
  return DWF_OK;
}
int dwf_fopr_reverse(dwf_fopr op, double* v_host, const void* w) {
  CTX();
  if (!op || !v_host || !w) return fail(DWF_ERR_INVALID_ARG, "null argument");
  RC(dwf_fopr_reverse_device(op, op->stage, w));
  HIPCHK(hipMemcpyAsync(v_host, op->stage, op->nelem * sizeof(double), hipMemcpyDeviceToHost,
                        g_ctx.stream));
  HIPCHK(hipStreamSynchronize(g_ctx.stream));
  return DWF_OK;
}

// ---- fields ---------------------------------------------------------------
int dwf_field_alloc(dwf_fopr op, void** f) {
  CTX();
  if (!op || !f) return fail(DWF_ERR_INVALID_ARG, "null argument");
  HIPCHK(hipMalloc(f, op->nelem * prec_bytes(op->prec)));
  HIPCHK(hipMemsetAsync(*f, 0, op->nelem * prec_bytes(op->prec), g_ctx.stream));
  HIPCHK(hipStreamSynchronize(g_ctx.stream));
  return DWF_OK;
}
int dwf_field_free(void* f) {
  if (!f) return DWF_OK;
  if (g_ctx.stream) (void)hipStreamSynchronize(g_ctx.stream);
  HIPCHK(hipFree(f));
  return DWF_OK;
}
int dwf_field_set_zero(dwf_fopr op, void* x) {
  CTX();
  HIPCHK(launch_set_zero(op->prec, x, op->nelem, g_ctx.stream));
  return DWF_OK;
}
int dwf_field_set(dwf_fopr op, void* x, double a) {
  CTX();
  if (!op || !x) return fail(DWF_ERR_INVALID_ARG, "null argument");
  HIPCHK(launch_set_value(op->prec, x, a, op->nelem, g_ctx.stream));
  return DWF_OK;
}
int dwf_field_copy(dwf_fopr op, void* y, const void* x) {
  CTX();
  HIPCHK(launch_copy(op->prec, y, x, op->nelem, g_ctx.stream));
  return DWF_OK;
}
int dwf_field_axpy(dwf_fopr op, void* y, double a, const void* x) {
  CTX();
  HIPCHK(launch_axpy(op->prec, y, a, x, op->nelem, g_ctx.stream));
  return DWF_OK;
}
int dwf_field_aypx(dwf_fopr op, double a, void* y, const void* x) {
  CTX();
  HIPCHK(launch_aypx(op->prec, a, y, x, op->nelem, g_ctx.stream));
  return DWF_OK;
}
int dwf_field_scal(dwf_fopr op, void* x, double a) {
  CTX();
  HIPCHK(launch_scal(op->prec, x, a, op->nelem, g_ctx.stream));
  return DWF_OK;
}
int dwf_field_norm2(dwf_fopr op, const void* x, double* r) {
  CTX();
  HIPCHK(launch_norm2(op->prec, x, op->nelem, g_ctx.red, 0, g_ctx.stream));
  return fetch(1, r);
}
int dwf_field_dot(dwf_fopr op, const void* x, const void* y, double* r) {
  CTX();
  HIPCHK(launch_dot(op->prec, x, y, op->nelem, g_ctx.red, 0, g_ctx.stream));
  return fetch(1, r);
}
int dwf_field_dotc(dwf_fopr op, const void* x, const void* y, double* re, double* im) {
  CTX();
  HIPCHK(launch_dotc_ls(op->prec, x, y, op->nelem, op->Ls, g_ctx.red, 0, g_ctx.stream));
  double v[2];
  RC(fetch(2, v));
  *re = v[0];
  *im = v[1];
  return DWF_OK;
}
int dwf_field_convert_precision(dwf_fopr op_to, void* v, dwf_fopr op_from, const void* w) {
  CTX();
  if (!op_to || !op_from || op_to->nelem != op_from->nelem)
    return fail(DWF_ERR_INVALID_ARG, "operators of different lattices");
  HIPCHK(launch_convert(op_to->prec, v, op_from->prec, w, op_to->nelem, g_ctx.stream));
  return DWF_OK;
}

// ---- solvers ----------------------------------------------------------------
// ---- learned polynomial preconditioner ---------------------------------------
// K = sum_{k=0}^{L} W_k H^k with H = 1 - D (D the even-odd Schur operator) and
// W_k per-chirality Ls x Ls real matrices fitted by least squares so that
// K D v ~ v for random v; K approximates D^-1, M = K K^+ approximates
// (D^+ D)^-1 and is symmetric positive semi-definite by construction.
//   K w:    acc = W_0 w, z = w; for k = 1..L: z = H z, acc += W_k z.
//   K^+ w:  t = W_L^T w; for k = L-1..0: t = W_k^T w + H^+ t.
// Every step is two fused hop kernels: the W_k terms ride in their epilogues
// (out2 accumulation for K, the chiral-matrix X term for K^+).  With Mp not
// spin-scalar the K recursion runs on Mp z, so that Mp leaves the hop
// prologue: the epilogue matrices are Mp (-E2) and W_k Mp^-1.

struct dwf_lpc_s {
  dwf_fopr op = nullptr;
  int L = 0;
  std::vector<ChiMat> W;
  ChiMat P, Pinv;          // Mp and its inverse (identity when Mp is spin-scalar)
  void* mats = nullptr;    // 2L + 3 matrices in op's compute type
  void* f[4] = {nullptr, nullptr, nullptr, nullptr};
  double* gpart = nullptr;
  double* gout = nullptr;
  double loss = -1.0;
};

// device matrix slots: 0 Q = P(-E2); 1+k A_k = W_k P^-1 (k = 0..L);
// L+2 Di^T W_L^T; L+3+k W_k^T (k = 0..L-1)
static const void* lpc_mat(const dwf_lpc_s* p, int slot) {
  return (const char*)p->mats + (size_t)slot * 2 * p->op->Ls * p->op->Ls * cbytes(p->op->prec);
}

static int lpc_upload(dwf_lpc_s* p) {
  const dwf_fopr op = p->op;
  const int L = p->L;
  std::vector<ChiMat> ms;
  ms.push_back(cm_mul(p->P, op->hmats[M_HE2]));
  ms.push_back(cm_mul(p->W[0], p->Pinv));   // applied to Mp w (the second output of step 0)
  for (int k = 1; k <= L; ++k) ms.push_back(cm_mul(p->W[k], p->Pinv));
  ms.push_back(cm_mul(op->hmats[M_DIT], cm_transpose(p->W[L])));
  for (int k = 0; k < L; ++k) ms.push_back(cm_transpose(p->W[k]));
  return upload_chimats(op->prec, p->mats, ms);
}

static int lpc_field(dwf_lpc_s* p, int i, void** out) {
  if (!p->f[i]) HIPCHK(hipMalloc(&p->f[i], p->op->nelem * prec_bytes(p->op->prec)));
  *out = p->f[i];
  return DWF_OK;
}

static int site_m(dwf_fopr op, void* out, const void* in, const void* E, void* out2 = nullptr,
                  const void* F = nullptr) {
  Epi e;
  e.E = E;
  e.out2 = out2;
  e.F = F;
  HIPCHK(launch_dwf(op->prec, K_SITE, out, in, op->links, op->nbr, 0, op->Vh, op->Ls, e,
                    g_ctx.stream));
  return DWF_OK;
}

// v = K w  (v != w)
static int lpc_apply_K(dwf_lpc_s* p, void* v, const void* w) {
  dwf_fopr op = p->op;
  const int L = p->L;
  void *zh, *t1, *tmp;
  RC(lpc_field(p, 0, &zh));
  RC(lpc_field(p, 1, &t1));
  RC(lpc_field(p, 2, &tmp));
  const void* zin = w;
  if (!op->mp_scalar) {
    RC(site_m(op, zh, w, mat(op, M_PRE), v, lpc_mat(p, 1)));   // zh = Mp w, v = W_0 w
    zin = zh;
  } else {
    RC(site_m(op, v, w, lpc_mat(p, 1)));   // P = 1: A_0 = W_0
  }
  for (int k = 1; k <= L; ++k) {
    RC(run_hop(op, false, 1, t1, zin, M_E1));
    Epi e;
    e.E = lpc_mat(p, 0);
    e.out2 = v;
    e.F = lpc_mat(p, 1 + k);
    e.X2 = v;
    HIPCHK(launch_dwf(op->prec, K_HOP, k < L ? zh : tmp, t1, op->links, op->nbr, 0, op->Vh,
                      op->Ls, e, g_ctx.stream));
    zin = zh;
  }
  return DWF_OK;
}

// v = K^+ w  (v != w)
static int lpc_apply_Kdag(dwf_lpc_s* p, void* v, const void* w) {
  dwf_fopr op = p->op;
  const int L = p->L;
  void *th, *t1, *tmp;
  RC(lpc_field(p, 0, &th));
  RC(lpc_field(p, 1, &t1));
  RC(lpc_field(p, 2, &tmp));
  RC(site_m(op, th, w, lpc_mat(p, L + 2)));   // th = Di^T W_L^T w
  for (int k = L - 1; k >= 0; --k) {
    RC(run_hop(op, true, 1, t1, th, M_G1));
    Epi e;
    e.E = mat(op, M_HG2);
    e.X = w;
    e.G = lpc_mat(p, L + 3 + k);
    if (k > 0) {
      e.out2 = th;
      e.F = mat(op, M_DIT);
    }
    HIPCHK(launch_dwf(op->prec, K_HOPDAG, k > 0 ? tmp : v, t1, op->links, op->nbr, 0, op->Vh,
                      op->Ls, e, g_ctx.stream));
  }
  return DWF_OK;
}

// v = K K^+ w  (v != w)
static int lpc_apply_M(dwf_lpc_s* p, void* v, const void* w) {
  void* t;
  RC(lpc_field(p, 3, &t));
  RC(lpc_apply_Kdag(p, t, w));
  return lpc_apply_K(p, v, t);
}

// H z = (1 - D) z = -E2 Hop_eo E1 Hop_oe (Mp) z
static int op_Hschur(dwf_fopr op, void* v, const void* w) {
  void *t0, *t1;
  RC(get_work(op, 0, &t0));
  RC(get_work(op, 1, &t1));
  const void* src = w;
  if (!op->mp_scalar) {
    RC(run_site(op, t0, w, M_PRE));
    src = t0;
  }
  RC(run_hop(op, false, 1, t1, src, M_E1));
  RC(run_hop(op, false, 0, v, t1, M_HE2));
  return DWF_OK;
}

// Cholesky solve of the symmetric positive definite n x n system A X = B (m columns)
static bool chol_solve(std::vector<long double>& A, int n, std::vector<long double>& B, int m) {
  for (int j = 0; j < n; ++j) {
    long double d = A[(size_t)j * n + j];
    for (int k = 0; k < j; ++k) d -= A[(size_t)j * n + k] * A[(size_t)j * n + k];
    if (!(d > 0)) return false;
    d = std::sqrt(d);
    A[(size_t)j * n + j] = d;
    for (int i = j + 1; i < n; ++i) {
      long double v = A[(size_t)i * n + j];
      for (int k = 0; k < j; ++k) v -= A[(size_t)i * n + k] * A[(size_t)j * n + k];
      A[(size_t)i * n + j] = v / d;
    }
  }
  for (int c = 0; c < m; ++c) {
    for (int i = 0; i < n; ++i) {
      long double v = B[(size_t)i * m + c];
      for (int k = 0; k < i; ++k) v -= A[(size_t)i * n + k] * B[(size_t)k * m + c];
      B[(size_t)i * m + c] = v / A[(size_t)i * n + i];
    }
    for (int i = n - 1; i >= 0; --i) {
      long double v = B[(size_t)i * m + c];
      for (int k = i + 1; k < n; ++k) v -= A[(size_t)k * n + i] * B[(size_t)k * m + c];
      B[(size_t)i * m + c] = v / A[(size_t)i * n + i];
    }
  }
  return true;
}

struct dwf_solver_s {
  int kind = 0;  // 0 CG, 1 Richardson(mixed), 2 flexible CG preconditioned by M = K K^+
  dwf_fopr op = nullptr;
  dwf_fopr op_f = nullptr;
  int niter = 0;
  double stop = 0;
  int guess = DWF_GUESS_RHS;
  int niter_inner = 0;
  double stop_inner = 0;
  void* f[5] = {nullptr, nullptr, nullptr, nullptr, nullptr};   // work in op precision
  void* g[2] = {nullptr, nullptr};   // work in op_f (kind 1) or the preconditioner's precision
  dwf_solver_s* inner = nullptr;
  int inner_iters = 0;
  dwf_lpc_s* lpc = nullptr;
  struct dwf_ptc_s* ptc = nullptr;   // flexible CG: the network preconditioner instead
  bool fast = false;   // inner CG: the sync-free iteration (cg_solve_fast)
};
static int ptc_apply_impl(struct dwf_ptc_s* n, void* out, const void* in);
static dwf_fopr ptc_op(struct dwf_ptc_s* n);

static int solver_field(dwf_fopr op, void** slot) {
  if (!*slot) HIPCHK(hipMalloc(slot, op->nelem * prec_bytes(op->prec)));
  return DWF_OK;
}

// ASolver_CG::solve (asolver_CG-tmpl.h:118-234)
static int cg_solve(dwf_solver_s* s, void* xq, const void* b, int* nconv_out, double* diff_out) {
  dwf_fopr op = s->op;
  for (int i = 0; i < 4; ++i) RC(solver_field(op, &s->f[i]));
  void *x = s->f[0], *r = s->f[1], *p = s->f[2], *ss = s->f[3];
  const int pr = op->prec;
  const size_t n = op->nelem;
  hipStream_t st = g_ctx.stream;
  double sr;
  HIPCHK(launch_norm2(pr, b, n, g_ctx.red, 0, st));
  RC(fetch(1, &sr));
  if (!(sr > 0)) {
    HIPCHK(launch_set_zero(pr, xq, n, st));
    *nconv_out = 0;
    *diff_out = 0;
    return DWF_OK;
  }
  const double snorm = 1.0 / sr;
  double rr, rrp;
  if (s->guess == DWF_GUESS_RHS) {
    HIPCHK(launch_copy(pr, r, b, n, st));
    HIPCHK(launch_copy(pr, x, b, n, st));
    RC(mult_impl(op, ss, x, op->mode, false));
    HIPCHK(launch_axpy(pr, r, -1.0, ss, n, st));
  } else if (s->guess == DWF_GUESS_ZERO) {
    HIPCHK(launch_copy(pr, r, b, n, st));
    HIPCHK(launch_set_zero(pr, x, n, st));
  } else {
    return fail(DWF_ERR_UNSUPPORTED, "InitialGuess::GIVEN is not supported by ASolver_CG");
  }
  HIPCHK(launch_copy(pr, p, r, n, st));
  HIPCHK(launch_norm2(pr, r, n, g_ctx.red, 0, st));
  RC(fetch(1, &rr));
  rrp = rr;
  int nconv = -1;
  for (int iter = 0; iter < s->niter; ++iter) {
    RC(mult_impl(op, ss, p, op->mode, false));
    double pap;
    HIPCHK(launch_dot(pr, ss, p, n, g_ctx.red, 0, st));
    RC(fetch(1, &pap));
    const double cr = rrp / pap;
    HIPCHK(launch_cg_xr(pr, x, r, p, ss, cr, n, g_ctx.red, 0, st));
    RC(fetch(1, &rr));
    const double bk = rr / rrp;
    HIPCHK(launch_aypx(pr, bk, p, r, n, st));
    rrp = rr;
    if (rr * snorm < s->stop) {
      nconv = iter + 1;
      break;
    }
  }
  HIPCHK(launch_copy(pr, xq, x, n, st));
  RC(mult_impl(op, ss, xq, op->mode, false));
  double diff2;
  HIPCHK(launch_sub_norm2(pr, ss, ss, b, n, g_ctx.red, 0, st));
  RC(fetch(1, &diff2));
  *nconv_out = nconv;
  *diff_out = std::sqrt(diff2 * snorm);
  s->inner_iters = nconv > 0 ? nconv : s->niter;
  return nconv < 0 ? fail(DWF_ERR_NOT_CONVERGED, "ASolver_CG: not converged") : DWF_OK;
}

// The sync-free CG iteration on op's DdagD (fields x, r, p, q, p2 in op's
// precision): <p, A p> = |D p|^2 comes from the epilogue of the last hop of D,
// the coefficients stay in device memory (ping-pong slots), and |r|^2 of
// iteration i is read by the host while iteration i + 1 already runs.
// Non-spin-scalar Mp: the update p = r + beta p runs as a site launch that
// also writes Mp p, the first factor of the next D p, into q (free between
// the x/r update and the next operator); p alternates between two fields.
// At Ls = 16 the r update moves into the last hop (fused matrix-core
// epilogue, DdagD p is never stored) and the x update into that site launch,
// or, with Mp folded into the hops, into one BLAS pass with the p update.
struct FastCG {
  dwf_fopr op = nullptr;
  bool pre = false, fused = false;
  void *x = nullptr, *r = nullptr, *p = nullptr, *q = nullptr, *p2 = nullptr;
  double* res = nullptr;
  int cur = 8, nxt = 9;
  static constexpr int PAP = 10;
  // |r|^2 of iteration i reaches the host in ring[i & 1]: written by the
  // finishing reduction kernel itself (fused path) or copied
  double* ring = nullptr;

  int init(dwf_solver_s* s) {
    op = s->op;
    pre = !op->mp_scalar;
    fused = op->Ls == 16 || op->Ls == 12;   // the matrix-core epilogue carries the updates
    for (int i = 0; i < (pre ? 5 : 4); ++i) RC(solver_field(op, &s->f[i]));
    x = s->f[0]; r = s->f[1]; p = s->f[2]; q = s->f[3]; p2 = s->f[4];
    res = g_ctx.red.result;
    cur = 8; nxt = 9;
    ring = fused ? g_ctx.host_ring : g_ctx.host_result + 32;
    return DWF_OK;
  }
  // r (|r|^2 in res[cur]) and p set: prepare the first step
  int start() {
    if (pre) RC(run_site(op, q, p, M_PRE));
    return DWF_OK;
  }
  // One iteration: r -= alpha A p (|r|^2 -> res[nxt], ring[iter & 1]),
  // x += alpha p, p = r + beta p.  Swaps cur and nxt.
  int step(int iter) {
    const int pr = op->prec;
    const size_t n = op->nelem;
    hipStream_t st = g_ctx.stream;
    if (fused) ring[iter & 1] = -1.0;
    if (fused && !pre) {
      RC(op_DdagD_cg(op, p, p, res + PAP, r, res + cur, res + nxt, ring + (iter & 1)));
      HIPCHK(launch_cg_px_dev(pr, x, p, r, res + cur, res + PAP, res + nxt, n, st));
    } else if (fused) {
      RC(op_DdagD_cg(op, p, q, res + PAP, r, res + cur, res + nxt, ring + (iter & 1)));
      Epi e;   // x += (|r|^2 / |D p|^2) p,  p2 = r + (|r_new|^2 / |r|^2) p,  q = Mp p2
      e.E = mat(op, M_ID);
      e.X = p;
      e.xnum = res + nxt;
      e.xden = res + cur;
      e.out2 = q;
      e.F = mat(op, M_PRE);
      e.upd = x;
      e.unum = res + cur;
      e.uden = res + PAP;
      HIPCHK(launch_dwf(op->prec, K_SITE, p2, r, op->links, op->nbr, 0, op->Vh, op->Ls, e, st));
      std::swap(p, p2);
    } else {
      RC(op_DdagD_nrm(op, q, p, res + PAP, pre ? q : nullptr));
      HIPCHK(launch_cg_xr_dev(pr, x, r, p, q, res + cur, res + PAP, n, g_ctx.red, nxt, st));
      if (pre) {
        Epi e;   // p2 = r + (|r_new|^2 / |r|^2) p,  q = Mp p2
        e.E = mat(op, M_ID);
        e.X = p;
        e.xnum = res + nxt;
        e.xden = res + cur;
        e.out2 = q;
        e.F = mat(op, M_PRE);
        HIPCHK(launch_dwf(op->prec, K_SITE, p2, r, op->links, op->nbr, 0, op->Vh, op->Ls, e,
                          st));
        std::swap(p, p2);
      } else {
        HIPCHK(launch_aypx_dev(pr, p, r, res + nxt, res + cur, n, st));
      }
    }
    if (!fused) {
      HIPCHK(hipMemcpyAsync(ring + (iter & 1), res + nxt, sizeof(double), hipMemcpyDeviceToHost,
                            st));
      HIPCHK(hipEventRecord(g_ctx.ring_ev[iter & 1], st));
    }
    std::swap(cur, nxt);
    return DWF_OK;
  }
  // wait for ring[slot]: the fused path polls its slot (set to -1 before the
  // iteration is queued, |r|^2 >= 0) and watches the stream for errors; the
  // other path waits for the event after the copy
  int wait(int slot) {
    if (!fused) {
      HIPCHK(hipEventSynchronize(g_ctx.ring_ev[slot]));
      return DWF_OK;
    }
    volatile double* v = ring + slot;
    for (unsigned spin = 1;; ++spin) {
      if (!(*v < 0)) return DWF_OK;
      if ((spin & 1023) == 0) {
        const hipError_t qs = hipStreamQuery(g_ctx.stream);
        if (qs == hipSuccess) {
          if (!(*v < 0)) return DWF_OK;
          return fail(DWF_ERR_HIP, "CG: residual norm not delivered");
        }
        if (qs != hipErrorNotReady) return fail(DWF_ERR_HIP, hipGetErrorString(qs));
      }
    }
  }
};

// CG on op's DdagD from a zero guess with the sync-free iteration.  The
// iteration that passes the test is followed by one more (never fewer
// iterations than the test asks for).  Used as the inner solver of the
// mixed-precision Richardson, where only the result matters.
// b_scratch: b may be overwritten (it then serves as the residual field).
static int cg_solve_fast(dwf_solver_s* s, void* xq, const void* b, int* nconv_out,
                         double* diff_out, bool b_scratch = false) {
  FastCG c;
  RC(c.init(s));
  dwf_fopr op = s->op;
  const int pr = op->prec;
  const size_t n = op->nelem;
  hipStream_t st = g_ctx.stream;
  double sr;
  HIPCHK(launch_norm2(pr, b, n, g_ctx.red, c.cur, st));   // = |r|^2 of the start
  RC(fetch(1, &sr, c.cur));
  if (!(sr > 0)) {
    HIPCHK(launch_set_zero(pr, xq, n, st));
    *nconv_out = 0;
    if (diff_out) *diff_out = 0;
    s->inner_iters = 0;
    return DWF_OK;
  }
  const double snorm = 1.0 / sr;
  if (b_scratch && !diff_out)
    c.r = const_cast<void*>(b);
  else
    HIPCHK(launch_copy(pr, c.r, b, n, st));
  HIPCHK(launch_copy(pr, c.p, b, n, st));
  HIPCHK(launch_set_zero(pr, c.x, n, st));
  RC(c.start());
  int nconv = -1, iter = 0;
  for (; iter < s->niter; ++iter) {
    RC(c.step(iter));
    if (iter >= 1) {
      RC(c.wait((iter - 1) & 1));
      if (c.ring[(iter - 1) & 1] * snorm < s->stop) {
        nconv = iter;   // the test passed after iteration iter - 1 (0-based)
        ++iter;         // iterations run: iter + 1
        break;
      }
    }
  }
  if (nconv < 0 && iter == s->niter && iter >= 1) {   // the last iteration's test
    RC(c.wait((iter - 1) & 1));
    if (c.ring[(iter - 1) & 1] * snorm < s->stop) nconv = iter;
  }
  HIPCHK(launch_copy(pr, xq, c.x, n, st));
  *nconv_out = nconv;
  s->inner_iters = iter;
  if (diff_out) {   // the true residual (the mixed solver's outer step does not need it)
    RC(mult_impl(op, c.q, xq, op->mode, false));
    double diff2;
    HIPCHK(launch_sub_norm2(pr, c.q, c.q, b, n, g_ctx.red, 0, st));
    RC(fetch(1, &diff2));
    *diff_out = std::sqrt(diff2 * snorm);
  }
  return nconv < 0 ? fail(DWF_ERR_NOT_CONVERGED, "CG: not converged") : DWF_OK;
}

// Flexible preconditioned CG on A = op (mode DdagD) with M = K K^+, zero
// initial guess:  beta = <z', r' - r> / <z, r> = -alpha <z', A p> / <z, r>.
// The preconditioner may run in another precision than op (conversions at
// its input and output).
static int lpc_apply_conv(dwf_solver_s* s, void* z, const void* r) {
  if (s->ptc) {   // the network runs on FP32 fields of its own operator
    dwf_fopr po = ptc_op(s->ptc);
    dwf_fopr op = s->op;
    if (po->prec == op->prec) return ptc_apply_impl(s->ptc, z, r);
    RC(solver_field(po, &s->g[0]));
    RC(solver_field(po, &s->g[1]));
    HIPCHK(launch_convert(po->prec, s->g[0], op->prec, r, op->nelem, g_ctx.stream));
    RC(ptc_apply_impl(s->ptc, s->g[1], s->g[0]));
    HIPCHK(launch_convert(op->prec, z, po->prec, s->g[1], op->nelem, g_ctx.stream));
    return DWF_OK;
  }
  dwf_lpc_s* P = s->lpc;
  dwf_fopr op = s->op;
  if (P->op->prec == op->prec) return lpc_apply_M(P, z, r);
  RC(solver_field(P->op, &s->g[0]));
  RC(solver_field(P->op, &s->g[1]));
  HIPCHK(launch_convert(P->op->prec, s->g[0], op->prec, r, op->nelem, g_ctx.stream));
  RC(lpc_apply_M(P, s->g[1], s->g[0]));
  HIPCHK(launch_convert(op->prec, z, P->op->prec, s->g[1], op->nelem, g_ctx.stream));
  return DWF_OK;
}

static int fcg_solve(dwf_solver_s* s, void* xq, const void* b, int* nconv_out, double* diff_out) {
  dwf_fopr op = s->op;
  for (int i = 0; i < 5; ++i) RC(solver_field(op, &s->f[i]));
  void *x = s->f[0], *r = s->f[1], *p = s->f[2], *q = s->f[3], *z = s->f[4];
  const int pr = op->prec;
  const size_t n = op->nelem;
  hipStream_t st = g_ctx.stream;
  double sr;
  HIPCHK(launch_norm2(pr, b, n, g_ctx.red, 0, st));
  RC(fetch(1, &sr));
  if (!(sr > 0)) {
    HIPCHK(launch_set_zero(pr, xq, n, st));
    *nconv_out = 0;
    *diff_out = 0;
    s->inner_iters = 0;
    return DWF_OK;
  }
  const double snorm = 1.0 / sr;
  HIPCHK(launch_copy(pr, r, b, n, st));
  HIPCHK(launch_set_zero(pr, x, n, st));
  RC(lpc_apply_conv(s, z, r));
  double rz;
  HIPCHK(launch_dot(pr, r, z, n, g_ctx.red, 0, st));
  RC(fetch(1, &rz));
  HIPCHK(launch_copy(pr, p, z, n, st));
  int nconv = -1;
  for (int iter = 0; iter < s->niter; ++iter) {
    RC(mult_impl(op, q, p, op->mode, false));
    double pq;
    HIPCHK(launch_dot(pr, p, q, n, g_ctx.red, 0, st));
    RC(fetch(1, &pq));
    const double alpha = rz / pq;
    double rr;
    HIPCHK(launch_cg_xr(pr, x, r, p, q, alpha, n, g_ctx.red, 0, st));
    RC(fetch(1, &rr));
    if (rr * snorm < s->stop) {
      nconv = iter + 1;
      break;
    }
    RC(lpc_apply_conv(s, z, r));
    double d[2];
    HIPCHK(launch_dot(pr, r, z, n, g_ctx.red, 0, st));
    HIPCHK(launch_dot(pr, z, q, n, g_ctx.red, 1, st));
    RC(fetch(2, d));
    const double beta = -alpha * d[1] / rz;
    rz = d[0];
    HIPCHK(launch_aypx(pr, beta, p, z, n, st));
  }
  HIPCHK(launch_copy(pr, xq, x, n, st));
  RC(mult_impl(op, q, xq, op->mode, false));
  double diff2;
  HIPCHK(launch_sub_norm2(pr, q, q, b, n, g_ctx.red, 0, st));
  RC(fetch(1, &diff2));
  *nconv_out = nconv;
  *diff_out = std::sqrt(diff2 * snorm);
  s->inner_iters = nconv > 0 ? nconv : s->niter;
  return nconv < 0 ? fail(DWF_ERR_NOT_CONVERGED, "flexible CG: not converged") : DWF_OK;
}

// ASolver_Richardson::solve with APrecond_Mixedprec
// (asolver_Richardson-tmpl.h:131-213, aprecond_Mixedprec-tmpl.h:55-108)
static int richardson_solve(dwf_solver_s* s, void* xq, const void* b, int* nconv_out,
                            double* diff_out) {
  dwf_fopr op = s->op, opf = s->op_f;
  RC(solver_field(op, &s->f[0]));
  RC(solver_field(op, &s->f[1]));
  RC(solver_field(opf, &s->g[0]));
  RC(solver_field(opf, &s->g[1]));
  void *r = s->f[0], *sv = s->f[1];
  void *rf = s->g[0], *sf = s->g[1];
  const int pr = op->prec;
  const size_t n = op->nelem;
  hipStream_t st = g_ctx.stream;
  s->inner_iters = 0;
  double bnorm2;
  HIPCHK(launch_norm2(pr, b, n, g_ctx.red, 0, st));
  RC(fetch(1, &bnorm2));
  if (!(bnorm2 > 0.0)) {
    HIPCHK(launch_set_zero(pr, xq, n, st));
    *nconv_out = 0;
    *diff_out = 0.0;
    return DWF_OK;
  }
  const double bnorm2_inv = 1.0 / bnorm2;
  HIPCHK(launch_copy(pr, r, b, n, st));
  HIPCHK(launch_scal(pr, r, -1.0, n, st));
  HIPCHK(launch_set_zero(pr, xq, n, st));
  int nconv = -1;
  double r2 = bnorm2;
  for (int iter = 0; iter < s->niter; ++iter) {
    if (r2 < bnorm2 * s->stop) {
      nconv = iter;
      break;
    }
    const double normalization = std::sqrt(r2);
    // r / |r| in the inner precision (one pass; the same values as scaling in
    // place and converting)
    if (pr == P64) {
      HIPCHK(launch_convert_scaled(opf->prec, rf, pr, r, 1.0 / normalization, n, st));
    } else {
      HIPCHK(launch_scal(pr, r, 1.0 / normalization, n, st));
      HIPCHK(launch_convert(opf->prec, rf, pr, r, n, st));
    }
    int inc = 0;
    double idiff = 0;
    // the fast inner CG may use rf as its residual field (rf is rebuilt each step)
    int rc = s->inner->kind == 2 ? fcg_solve(s->inner, sf, rf, &inc, &idiff)
             : (s->inner->fast && opf->mode == "DdagD")
                 ? cg_solve_fast(s->inner, sf, rf, &inc, nullptr, true)
                 : cg_solve(s->inner, sf, rf, &inc, &idiff);
    if (rc != DWF_OK && rc != DWF_ERR_NOT_CONVERGED) return rc;
    s->inner_iters += s->inner->inner_iters;
    if (pr == P64) {   // xq -= |r| s in one pass
      HIPCHK(launch_axpy_from(pr, xq, -normalization, opf->prec, sf, n, st));
    } else {
      HIPCHK(launch_convert(pr, sv, opf->prec, sf, n, st));
      HIPCHK(launch_axpy(pr, xq, -normalization, sv, n, st));
    }
    RC(mult_impl(op, r, xq, op->mode, false));
    HIPCHK(launch_sub_norm2(pr, r, r, b, n, g_ctx.red, 0, st));
    RC(fetch(1, &r2));
  }
  // r2 = |b - A xq|^2 of the final xq (computed by the last step, as the
  // separate residual evaluation after the loop would)
  *nconv_out = nconv;
  *diff_out = std::sqrt(r2 * bnorm2_inv);
  return nconv < 0 ? fail(DWF_ERR_NOT_CONVERGED, "ASolver_Richardson: not converged") : DWF_OK;
}

int dwf_solver_create_cg(dwf_fopr op, int max_iter, double stop_cond, int guess,
                         dwf_solver* out) {
  CTX();
  if (!op || !out || max_iter < 0) return fail(DWF_ERR_INVALID_ARG, "invalid argument");
  if (guess != DWF_GUESS_RHS && guess != DWF_GUESS_ZERO && guess != DWF_GUESS_GIVEN)
    return fail(DWF_ERR_INVALID_ARG, "unknown initial guess mode");
  dwf_solver_s* s = new dwf_solver_s;
  s->kind = 0;
  s->op = op;
  s->niter = max_iter;
  s->stop = stop_cond;
  s->guess = guess;
  *out = s;
  return DWF_OK;
}

int dwf_solver_create_richardson(dwf_fopr op_d, dwf_fopr op_f, int max_iter, double stop_cond,
                                 int max_iter_inner, double stop_cond_inner, dwf_solver* out) {
  CTX();
  if (!op_d || !op_f || !out) return fail(DWF_ERR_INVALID_ARG, "null argument");
  if (op_d->nelem != op_f->nelem) return fail(DWF_ERR_INVALID_ARG, "operators differ in size");
  dwf_solver_s* s = new dwf_solver_s;
  s->kind = 1;
  s->op = op_d;
  s->op_f = op_f;
  s->niter = max_iter;
  s->stop = stop_cond;
  s->niter_inner = max_iter_inner;
  s->stop_inner = stop_cond_inner;
  s->inner = new dwf_solver_s;
  s->inner->kind = 0;
  s->inner->op = op_f;
  s->inner->niter = max_iter_inner;
  s->inner->stop = stop_cond_inner;
  s->inner->guess = DWF_GUESS_ZERO;
  s->inner->fast = true;
  *out = s;
  return DWF_OK;
}

// Schedule autotuner of the mixed-precision solve: for every inner operator
// (storage type) and inner criterion, one untimed and one timed solve of the
// caller's system; the solver of the fastest candidate that reached stop_cond
// is returned, the others are destroyed.
int dwf_solver_tune_richardson(dwf_fopr op_d, int n_inner, const dwf_fopr* inner, int n_tol,
                               const double* tols, int max_iter, double stop_cond,
                               int max_iter_inner, const void* b, void* x, double* times,
                               int* best, dwf_solver* out) {
  CTX();
  if (!op_d || n_inner < 1 || !inner || n_tol < 1 || !tols || !b || !x || !times || !best ||
      !out)
    return fail(DWF_ERR_INVALID_ARG, "invalid argument");
  *out = nullptr;
  *best = -1;
  double tbest = 0.0;
  for (int i = 0; i < n_inner; ++i)
    for (int j = 0; j < n_tol; ++j) {
      const int c = i * n_tol + j;
      times[c] = -1.0;
      dwf_solver sv = nullptr;
      RC(dwf_solver_create_richardson(op_d, inner[i], max_iter, stop_cond, max_iter_inner,
                                      tols[j], &sv));
      int nconv = -1;
      double diff = 0.0;
      int rc = dwf_solver_solve(sv, x, b, &nconv, &diff);   // fields allocated here
      double dt = -1.0;
      if (rc == DWF_OK && nconv >= 0) {
        const auto t0 = std::chrono::steady_clock::now();
        rc = dwf_solver_solve(sv, x, b, &nconv, &diff);
        HIPCHK(hipStreamSynchronize(g_ctx.stream));
        dt = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
      }
      if (rc != DWF_OK && rc != DWF_ERR_NOT_CONVERGED) {
        dwf_solver_destroy(sv);
        if (*out) dwf_solver_destroy(*out);
        *out = nullptr;
        return rc;
      }
      if (rc == DWF_OK && nconv >= 0 && dt >= 0.0) {
        times[c] = dt;
        if (*best < 0 || dt < tbest) {
          if (*out) dwf_solver_destroy(*out);
          *out = sv;
          *best = c;
          tbest = dt;
          continue;
        }
      }
      dwf_solver_destroy(sv);
    }
  if (*best < 0) return fail(DWF_ERR_NOT_CONVERGED, "no candidate reached the stopping condition");
  return DWF_OK;
}

int dwf_solver_create_fcg(dwf_fopr op, dwf_lpc prec, int max_iter, double stop_cond,
                          dwf_solver* out) {
  CTX();
  if (!op || !prec || !out || max_iter < 0) return fail(DWF_ERR_INVALID_ARG, "invalid argument");
  if (op->nelem != prec->op->nelem) return fail(DWF_ERR_INVALID_ARG, "operators differ in size");
  dwf_solver_s* s = new dwf_solver_s;
  s->kind = 2;
  s->op = op;
  s->niter = max_iter;
  s->stop = stop_cond;
  s->guess = DWF_GUESS_ZERO;
  s->lpc = prec;
  *out = s;
  return DWF_OK;
}

int dwf_solver_create_richardson_lpc(dwf_fopr op_d, dwf_fopr op_f, dwf_lpc prec, int max_iter,
                                     double stop_cond, int max_iter_inner,
                                     double stop_cond_inner, dwf_solver* out) {
  CTX();
  if (!op_d || !op_f || !prec || !out) return fail(DWF_ERR_INVALID_ARG, "null argument");
  if (op_d->nelem != op_f->nelem || op_f->nelem != prec->op->nelem)
    return fail(DWF_ERR_INVALID_ARG, "operators differ in size");
  RC(dwf_solver_create_richardson(op_d, op_f, max_iter, stop_cond, max_iter_inner,
                                  stop_cond_inner, out));
  (*out)->inner->kind = 2;
  (*out)->inner->lpc = prec;
  return DWF_OK;
}

int dwf_solver_solve(dwf_solver s, void* x, const void* b, int* nconv, double* diff) {
  CTX();
  if (!s || !x || !b || !nconv || !diff) return fail(DWF_ERR_INVALID_ARG, "null argument");
  if (!s->op->have_config || (s->op_f && !s->op_f->have_config))
    return fail(DWF_ERR_NO_CONFIG, "set_config has not been called");
  if (s->kind == 1) {
    if (s->op_f->mode != s->op->mode) s->op_f->mode = s->op->mode;
    return richardson_solve(s, x, b, nconv, diff);
  }
  if (s->kind == 2) {
    if (s->op->mode != "DdagD")
      return fail(DWF_ERR_UNSUPPORTED, "the learned preconditioner solves mode DdagD");
    return fcg_solve(s, x, b, nconv, diff);
  }
  return cg_solve(s, x, b, nconv, diff);
}

int dwf_solver_inner_iterations(dwf_solver s) { return s ? s->inner_iters : 0; }

int dwf_solver_destroy(dwf_solver s) {
  if (!s) return DWF_OK;
  if (g_ctx.stream) (void)hipStreamSynchronize(g_ctx.stream);
  for (auto& f : s->f)
    if (f) (void)hipFree(f);
  for (auto& g : s->g)
    if (g) (void)hipFree(g);
  if (s->inner) dwf_solver_destroy(s->inner);
  delete s;
  return DWF_OK;
}

int dwf_lpc_create(dwf_fopr op, int degree, dwf_lpc* out) {
  CTX();
  if (!op || !out) return fail(DWF_ERR_INVALID_ARG, "null argument");
  if (degree < 1 || degree > 4) return fail(DWF_ERR_INVALID_ARG, "degree must be 1..4");
  if (op->Ls > 16) return fail(DWF_ERR_UNSUPPORTED, "the learned preconditioner needs Ls <= 16");
  dwf_lpc_s* p = new dwf_lpc_s;
  p->op = op;
  p->L = degree;
  const int Ls = op->Ls;
  // start from the truncated Neumann series: W_k = 1
  p->W.assign(degree + 1, cm_identity(Ls));
  if (op->mp_scalar) {
    p->P = cm_identity(Ls);
    p->Pinv = cm_identity(Ls);
  } else {
    p->P = op->hmats[M_PRE];
    if (!cm_inverse(p->P, p->Pinv)) {
      delete p;
      return fail(DWF_ERR_UNSUPPORTED, "Mp is singular");
    }
  }
  const size_t nm = 2 * (size_t)degree + 3;
  int rc = DWF_OK;
  if (hipMalloc(&p->mats, nm * 2 * Ls * Ls * cbytes(op->prec)) != hipSuccess)
    rc = fail(DWF_ERR_HIP, "device allocation failed");
  if (rc == DWF_OK) rc = lpc_upload(p);
  if (rc != DWF_OK) {
    if (p->mats) (void)hipFree(p->mats);
    delete p;
    return rc;
  }
  *out = p;
  return DWF_OK;
}

int dwf_lpc_destroy(dwf_lpc p) {
  if (!p) return DWF_OK;
  if (g_ctx.stream) (void)hipStreamSynchronize(g_ctx.stream);
  for (auto& f : p->f)
    if (f) (void)hipFree(f);
  if (p->mats) (void)hipFree(p->mats);
  if (p->gpart) (void)hipFree(p->gpart);
  if (p->gout) (void)hipFree(p->gout);
  delete p;
  return DWF_OK;
}

int dwf_lpc_degree(dwf_lpc p) { return p ? p->L : -1; }

int dwf_lpc_set_weights(dwf_lpc p, const double* W) {
  CTX();
  if (!p || !W) return fail(DWF_ERR_INVALID_ARG, "null argument");
  const int Ls = p->op->Ls;
  for (int k = 0; k <= p->L; ++k)
    for (int c = 0; c < 2; ++c)
      for (int i = 0; i < Ls * Ls; ++i)
        p->W[k].m[c][i] = W[((size_t)k * 2 + c) * Ls * Ls + i];
  return lpc_upload(p);
}

int dwf_lpc_get_weights(dwf_lpc p, double* W) {
  if (!p || !W) return fail(DWF_ERR_INVALID_ARG, "null argument");
  const int Ls = p->op->Ls;
  for (int k = 0; k <= p->L; ++k)
    for (int c = 0; c < 2; ++c)
      for (int i = 0; i < Ls * Ls; ++i)
        W[((size_t)k * 2 + c) * Ls * Ls + i] = (double)p->W[k].m[c][i];
  return DWF_OK;
}

int dwf_lpc_apply(dwf_lpc p, const char* which, void* v, const void* w) {
  CTX();
  if (!p || !which || !v || !w) return fail(DWF_ERR_INVALID_ARG, "null argument");
  if (v == w) return fail(DWF_ERR_INVALID_ARG, "the preconditioner does not work in place");
  if (!p->op->have_config) return fail(DWF_ERR_NO_CONFIG, "set_config has not been called");
  const std::string m(which);
  if (m == "K") return lpc_apply_K(p, v, w);
  if (m == "Kdag") return lpc_apply_Kdag(p, v, w);
  if (m == "M") return lpc_apply_M(p, v, w);
  return fail(DWF_ERR_UNKNOWN_MODE, "unknown preconditioner part: " + m);
}

int dwf_lpc_train(dwf_lpc p, dwf_fopr opt, int nvec, unsigned long long seed, double ridge,
                  double* loss) {
  CTX();
  if (!p || !opt) return fail(DWF_ERR_INVALID_ARG, "null argument");
  if (nvec < 1 || !(ridge >= 0)) return fail(DWF_ERR_INVALID_ARG, "invalid argument");
  dwf_fopr op = p->op;
  for (int mu = 0; mu < 4; ++mu)
    if (opt->N[mu] != op->N[mu]) return fail(DWF_ERR_INVALID_ARG, "lattice sizes differ");
  if (opt->Ls != op->Ls) return fail(DWF_ERR_INVALID_ARG, "Ls differs");
  if (!opt->have_config) return fail(DWF_ERR_NO_CONFIG, "set_config has not been called");
  const int L = p->L, Ls = op->Ls, F = L + 2;
  const int NP = F * (F + 1) / 2;
  const size_t per = 2 * (size_t)Ls * Ls;
  hipStream_t st = g_ctx.stream;
  if (!p->gpart) HIPCHK(hipMalloc(&p->gpart, gram_partial_doubles() * sizeof(double)));
  if (!p->gout) HIPCHK(hipMalloc(&p->gout, 21 * per * sizeof(double)));
  std::vector<void*> fl(F, nullptr);
  auto release = [&]() {
    (void)hipStreamSynchronize(st);
    for (auto& f : fl)
      if (f) (void)hipFree(f);
  };
  for (int i = 0; i < F; ++i)
    if (hipMalloc(&fl[i], opt->nelem * prec_bytes(opt->prec)) != hipSuccess) {
      release();
      return fail(DWF_ERR_HIP, "device allocation failed");
    }
  std::vector<long double> acc(NP * per, 0.0L);
  std::vector<double> h(NP * per);
  int rc = DWF_OK;
  for (int j = 0; j < nvec && rc == DWF_OK; ++j) {
    // features z_k = H^k D v, target v
    if (launch_fill_uniform(opt->prec, fl[0], opt->nelem, seed + (unsigned long long)j *
                                0xD1B54A32D192ED03ull, st) != hipSuccess) {
      rc = fail(DWF_ERR_HIP, "fill kernel failed");
      break;
    }
    rc = mult_impl(opt, fl[1], fl[0], "D", false);
    for (int k = 1; k <= L && rc == DWF_OK; ++k) rc = op_Hschur(opt, fl[1 + k], fl[k]);
    if (rc != DWF_OK) break;
    if (launch_chiral_gram(opt->prec, fl.data(), F, opt->Vh, Ls, p->gpart, p->gout, st) !=
            hipSuccess ||
        hipMemcpyAsync(h.data(), p->gout, NP * per * sizeof(double), hipMemcpyDeviceToHost,
                       st) != hipSuccess ||
        hipStreamSynchronize(st) != hipSuccess) {
      rc = fail(DWF_ERR_HIP, "Gram kernel failed");
      break;
    }
    for (size_t i = 0; i < h.size(); ++i) acc[i] += h[i];
  }
  release();
  if (rc != DWF_OK) return rc;
  // pair index of fields (i <= j)
  auto pid = [&](int i, int j) { return i * F - i * (i - 1) / 2 + (j - i); };
  auto G = [&](int i, int j, int c, int a, int b) -> long double {  // sum p(f_i)(a) p(f_j)(b)
    if (i <= j) return acc[pid(i, j) * per + (size_t)c * Ls * Ls + (size_t)a * Ls + b];
    return acc[pid(j, i) * per + (size_t)c * Ls * Ls + (size_t)b * Ls + a];
  };
  const int d = (L + 1) * Ls;
  long double lsum = 0, vsum = 0;
  std::vector<ChiMat> Wn(L + 1, ChiMat(Ls));
  for (int c = 0; c < 2; ++c) {
    std::vector<long double> A((size_t)d * d), B((size_t)d * Ls);
    long double tr = 0;
    for (int ka = 0; ka <= L; ++ka)
      for (int a = 0; a < Ls; ++a) {
        for (int kb = 0; kb <= L; ++kb)
          for (int b = 0; b < Ls; ++b)
            A[(size_t)(ka * Ls + a) * d + kb * Ls + b] = G(1 + ka, 1 + kb, c, a, b);
        tr += A[(size_t)(ka * Ls + a) * d + ka * Ls + a];
        for (int sidx = 0; sidx < Ls; ++sidx)   // B = Bbig^T: [(k, t)][s] = sum v(s) z_k(t)
          B[(size_t)(ka * Ls + a) * Ls + sidx] = G(0, 1 + ka, c, sidx, a);
      }
    std::vector<long double> A0(A), B0(B);
    const long double reg = ridge * tr / d;
    for (int i = 0; i < d; ++i) A[(size_t)i * d + i] += reg;
    if (!chol_solve(A, d, B, Ls)) return fail(DWF_ERR_NOT_CONVERGED, "training system is singular");
    // B now holds W^T: [(k, t)][s] = W_k[s][t]
    long double vv = 0;
    for (int sidx = 0; sidx < Ls; ++sidx) vv += G(0, 0, c, sidx, sidx);
    // |W Z - V|^2 = vv - 2 tr(W Bbig^T) + tr(W G W^T)
    long double wb = 0, wgw = 0;
    for (int i = 0; i < d; ++i)
      for (int sidx = 0; sidx < Ls; ++sidx) {
        wb += B[(size_t)i * Ls + sidx] * B0[(size_t)i * Ls + sidx];
        long double gw = 0;
        for (int j2 = 0; j2 < d; ++j2) gw += A0[(size_t)i * d + j2] * B[(size_t)j2 * Ls + sidx];
        wgw += B[(size_t)i * Ls + sidx] * gw;
      }
    lsum += 0.5L * (vv - 2 * wb + wgw);
    vsum += 0.5L * vv;
    for (int k = 0; k <= L; ++k)
      for (int sidx = 0; sidx < Ls; ++sidx)
        for (int t = 0; t < Ls; ++t) Wn[k].at(c, sidx, t) = B[(size_t)(k * Ls + t) * Ls + sidx];
  }
  p->W = Wn;
  p->loss = (double)(lsum / vsum);
  if (loss) *loss = p->loss;
  return lpc_upload(p);
}

int dwf_lpc_chiral_gram(dwf_fopr op, int nfields, const void* const* fields, double* out_host) {
  CTX();
  if (!op || !fields || !out_host || nfields < 1 || nfields > 6)
    return fail(DWF_ERR_INVALID_ARG, "invalid argument");
  if (op->Ls > 16) return fail(DWF_ERR_UNSUPPORTED, "Ls > 16");
  const size_t per = 2 * (size_t)op->Ls * op->Ls;
  const int NP = nfields * (nfields + 1) / 2;
  double *part = nullptr, *o = nullptr;
  HIPCHK(hipMalloc(&part, gram_partial_doubles() * sizeof(double)));
  int rc = DWF_OK;
  if (hipMalloc(&o, NP * per * sizeof(double)) != hipSuccess)
    rc = fail(DWF_ERR_HIP, "device allocation failed");
  if (rc == DWF_OK &&
      (launch_chiral_gram(op->prec, fields, nfields, op->Vh, op->Ls, part, o, g_ctx.stream) !=
           hipSuccess ||
       hipMemcpyAsync(out_host, o, NP * per * sizeof(double), hipMemcpyDeviceToHost,
                      g_ctx.stream) != hipSuccess ||
       hipStreamSynchronize(g_ctx.stream) != hipSuccess))
    rc = fail(DWF_ERR_HIP, "Gram kernel failed");
  (void)hipStreamSynchronize(g_ctx.stream);
  (void)hipFree(part);
  if (o) (void)hipFree(o);
  return rc;
}

int dwf_field_fill_uniform(dwf_fopr op, void* x, unsigned long long seed) {
  CTX();
  if (!op || !x) return fail(DWF_ERR_INVALID_ARG, "null argument");
  HIPCHK(launch_fill_uniform(op->prec, x, op->nelem, seed, g_ctx.stream));
  return DWF_OK;
}

// AShiftsolver_CG::solve (ashiftsolver_CG-tmpl.h:89-329), FP64
// The iteration of AShiftsolver_CG (ashiftsolver_CG-tmpl.h:89-286): the
// unshifted CG (shift 0) runs as in the reference; the other shifts' vectors
// are not touched every iteration.  With the residuals r_1 .. r_m of a block of
// m <= kFoldMax iterations kept in a ring, x_j and p_j after the block are
//   p_j = a p_j' + sum_l c_l r_l,   x_j = x_j' + b p_j' + sum_l d_l r_l
// (primes: at the block start), where a, b, c, d follow from the reference's
// scalar recurrences (x_j -= beta_s p_j; p_j = alpha_s p_j + zeta r); one fold
// pass per block applies them (k_ms_fold), reading each r_l once for all
// shifts: 1 + 4 n_shift / m field passes per iteration instead of 4 n_shift + 1.
// The truncation rule needs |p_j|^2 every iteration: inside a block it follows
// |p_j|^2 <- alpha_s^2 |p_j|^2 + zeta^2 |r|^2 (r is orthogonal to the previous
// p_j in CG), and every fold replaces it by the computed norm.
int dwf_shiftsolver_cg_solve(dwf_fopr op, int max_iter, double stop_cond, int nshift,
                             const double* sigma, void** xq, const void* b, int* nconv_out,
                             double* diff_out) {
  CTX();
  if (!op || !sigma || !xq || !b || nshift < 1) return fail(DWF_ERR_INVALID_ARG, "invalid argument");
  if (op->prec != P64) return fail(DWF_ERR_UNSUPPORTED, "multishift CG runs on the FP64 operator");
  if (!op->have_config) return fail(DWF_ERR_NO_CONFIG, "set_config has not been called");
  RC(ensure_shift_scratch(nshift));
  const size_t n = op->nelem;
  const size_t bytes = n * sizeof(double);
  const int K = kFoldMax;
  hipStream_t st = g_ctx.stream;
  std::vector<void*> p(nshift, nullptr), R(K + 1, nullptr);
  void *s = nullptr, *coef = nullptr, *rptr = nullptr;
  auto release = [&]() {
    (void)hipStreamSynchronize(st);
    for (auto& q : p)
      if (q) (void)hipFree(q);
    for (auto& q : R)
      if (q) (void)hipFree(q);
    if (s) (void)hipFree(s);
    if (coef) (void)hipFree(coef);
    if (rptr) (void)hipFree(rptr);
  };
  bool ok = hipMalloc(&s, bytes) == hipSuccess &&
            hipMalloc(&coef, sizeof(double) * kFoldCoef * nshift) == hipSuccess &&
            hipMalloc(&rptr, sizeof(void*) * K) == hipSuccess;
  for (int i = 0; ok && i < nshift; ++i) ok = hipMalloc(&p[i], bytes) == hipSuccess;
  for (int i = 0; ok && i <= K; ++i) ok = hipMalloc(&R[i], bytes) == hipSuccess;
  if (!ok) {
    release();
    return fail(DWF_ERR_HIP, "device allocation failed");
  }
  int rc = DWF_OK;
  auto body = [&]() -> int {
    double bn;
    HIPCHK(launch_norm2(P64, b, n, g_ctx.red, 0, st));
    RC(fetch(1, &bn));
    if (!(bn > 0)) return fail(DWF_ERR_INVALID_ARG, "zero source");
    const double snorm = 1.0 / bn;
    std::vector<double> zeta1(nshift, 1.0), zeta2(nshift, 1.0), csh2(nshift), pp(nshift, 0.0);
    for (int i = 0; i < nshift; ++i) csh2[i] = sigma[i] - sigma[0];
    const double sigma0 = sigma[0];
    int Nshift2 = nshift;
    HIPCHK(launch_copy(P64, R[0], b, n, st));
    for (int i = 0; i < nshift; ++i) {
      HIPCHK(launch_copy(P64, p[i], b, n, st));
      HIPCHK(launch_set_zero(P64, xq[i], n, st));
    }
    void** hp = (void**)(g_ctx.host_coef + 3 * nshift);
    for (int i = 0; i < nshift; ++i) {
      hp[i] = xq[i];
      hp[nshift + i] = p[i];
    }
    HIPCHK(hipMemcpyAsync(g_ctx.sh_ptrs, hp, sizeof(void*) * 2 * nshift, hipMemcpyHostToDevice, st));
    HIPCHK(hipStreamSynchronize(st));
    double rr = bn;
    // block coefficients per shift: [a, b, c_1..c_K, d_1..d_K]; |p_j|^2 estimates
    std::vector<double> cf((size_t)kFoldCoef * nshift, 0.0), ppe(nshift, bn);
    auto reset_coef = [&]() {
      std::fill(cf.begin(), cf.end(), 0.0);
      for (int j = 0; j < nshift; ++j) cf[(size_t)j * kFoldCoef] = 1.0;
    };
    reset_coef();
    int m = 0;            // iterations in the current block (R[1..m] hold their residuals)
    int nfold = nshift;   // shifts 1..nfold-1 were active at the block start
    auto fold = [&]() -> int {
      if (nfold > 1 && m > 0) {
        HIPCHK(hipMemcpyAsync(coef, cf.data(), sizeof(double) * kFoldCoef * nfold,
                              hipMemcpyHostToDevice, st));
        HIPCHK(hipMemcpyAsync(rptr, R.data() + 1, sizeof(void*) * m, hipMemcpyHostToDevice, st));
        // the fold kernel takes up to 31 shifts (local 1..31) per pass
        for (int j0 = 0; j0 + 1 < nfold; j0 += 31) {
          const int nloc = std::min(32, nfold - j0);
          HIPCHK(launch_ms_fold(g_ctx.sh_ptrs + j0, g_ctx.sh_ptrs + nshift + j0,
                                (const void* const*)rptr, (const double*)coef + (size_t)j0 * kFoldCoef,
                                nloc, m, n, g_ctx.sh_partial, g_ctx.sh_pp + j0, st));
        }
        double* hpp = g_ctx.host_coef + 5 * nshift;
        HIPCHK(hipMemcpyAsync(hpp, g_ctx.sh_pp, sizeof(double) * nfold, hipMemcpyDeviceToHost, st));
        HIPCHK(hipStreamSynchronize(st));
        for (int j = 1; j < nfold; ++j) ppe[j] = hpp[j];
      }
      std::swap(R[0], R[m]);   // the block's last residual starts the next block
      reset_coef();
      m = 0;
      nfold = Nshift2;
      return DWF_OK;
    };
    double alpha_p = 0.0, beta_p = 1.0;
    int Nconv2 = -1;
    bool converged = false;
    // DdagD: <p0, (DdagD + sigma0) p0> = |D p0|^2 + sigma0 |p0|^2, the first norm
    // fused into the last hop of D and the second into the update of p0, and
    // sigma0 p0 added in the residual update: no separate pass over s and p0.
    // Other modes: s += sigma0 p0 and <p0, s> in one pass, as the reference.
    const bool fused = op->mode == "DdagD";
    double pp0 = bn;   // |p0|^2 (p0 = b at the start)
    for (int iter = 0; iter < max_iter; ++iter) {
      double pa_p;
      if (fused) {
        RC(op_DdagD_nrm(op, s, p[0], g_ctx.red.result));
        double v[2];
        RC(fetch(iter == 0 ? 1 : 2, v));
        if (iter > 0) pp0 = v[1];
        pa_p = v[0] + sigma0 * pp0;
      } else {
        RC(mult_impl(op, s, p[0], op->mode, false));
        HIPCHK(launch_axpy_dot(P64, s, sigma0, p[0], n, g_ctx.red, 0, st));
        RC(fetch(1, &pa_p));
      }
      const double rr_p = rr;
      const double beta = -rr_p / pa_p;
      ++m;
      HIPCHK(launch_cg_xr_to(P64, xq[0], R[m], R[m - 1], p[0], s, -beta, fused ? sigma0 : 0.0, n,
                             g_ctx.red, 0, st));
      RC(fetch(1, &rr));
      const double alpha = rr / rr_p;
      if (fused) HIPCHK(launch_aypx_norm(P64, alpha, p[0], R[m], n, g_ctx.red, 1, st));
      else HIPCHK(launch_aypx(P64, alpha, p[0], R[m], n, st));
      pp[0] = rr;
      const double alpha_h = 1.0 + alpha_p * beta / beta_p;
      for (int ish = 1; ish < Nshift2; ++ish) {
        double zeta = (alpha_h - csh2[ish] * beta) / zeta1[ish] + (1.0 - alpha_h) / zeta2[ish];
        zeta = 1.0 / zeta;
        const double zr = zeta / zeta1[ish];
        const double beta_s = beta * zr;
        const double alpha_s = alpha * zr * zr;
        double* c = cf.data() + (size_t)ish * kFoldCoef;
        // x_j -= beta_s p_j (p_j before its update), then p_j = alpha_s p_j + zeta r_m
        c[1] -= beta_s * c[0];
        for (int l = 0; l < m - 1; ++l) c[2 + K + l] -= beta_s * c[2 + l];
        c[0] *= alpha_s;
        for (int l = 0; l < m - 1; ++l) c[2 + l] *= alpha_s;
        c[2 + m - 1] = zeta;
        ppe[ish] = alpha_s * alpha_s * ppe[ish] + zeta * zeta * rr;
        pp[ish] = ppe[ish] * snorm;
        zeta2[ish] = zeta1[ish];
        zeta1[ish] = zeta;
      }
      int ish1 = Nshift2;
      for (int ish = Nshift2 - 1; ish >= 0; --ish)
        if (pp[ish] > stop_cond) {
          ish1 = ish + 1;
          break;
        }
      Nshift2 = ish1;
      alpha_p = alpha;
      beta_p = beta;
      Nconv2 += 1;
      converged = rr * snorm < stop_cond;
      if (m == K || converged) RC(fold());
      if (converged) break;
    }
    if (!converged && m > 0) RC(fold());
    double dmax = -1.0;
    for (int i = 0; i < nshift; ++i) {
      RC(mult_impl(op, s, xq[i], op->mode, false));
      HIPCHK(launch_axpy(P64, s, sigma[i], xq[i], n, st));
      double d2;
      HIPCHK(launch_sub_norm2(P64, s, s, b, n, g_ctx.red, 0, st));
      RC(fetch(1, &d2));
      const double d1 = std::sqrt(d2 * snorm);
      if (d1 > dmax) dmax = d1;
    }
    *nconv_out = Nconv2;
    *diff_out = dmax;
    return converged ? DWF_OK : fail(DWF_ERR_NOT_CONVERGED, "AShiftsolver_CG: not converged");
  };
  rc = body();
  release();
  return rc;
}

// ---- batched multishift CG ---------------------------------------------------
// nsys systems (right-hand sides), each solved for the same nshift shifts with
// the recurrences of AShiftsolver_CG (ashiftsolver_CG-tmpl.h:89-286), all
// advanced together.  Systems on the same operator form one member: its nr
// right-hand sides (padded to a power of two with zero slots) are contiguous
// fields that every hop of D^+D processes in one launch, reading the links of
// a site once for all of them.  Members on different operators (lattices of
// different volume, same Ls and precision) run as one grouped launch per hop.
// The scalars stay on the device (k_ms_ctrl); the host polls a status word in
// coherent memory one iteration behind, or replays the iteration from a
// captured HIP graph.
struct MsMember {
  dwf_fopr_s* op = nullptr;
  int nrs = 0, npad = 1;          // 2^nrs fields per buffer, of which the first nr are systems
  std::vector<int> sys;           // system index of each used slot
  void* buf[7] = {};              // P0 S (unused) T0 T1 U U2, each npad fields
};
enum { MB_P0, MB_S, MB_R, MB_T0, MB_T1, MB_U, MB_U2 };

struct dwf_msbatch_s {
  int nsys = 0, nshift = 0, prec = 0, Ls = 0, flags = 0;
  bool mp_scalar = false;
  std::vector<dwf_fopr_s*> ops;            // per system
  std::vector<MsMember> mem;
  std::vector<std::pair<int, int>> slot;   // system -> (member, slot)
  std::vector<std::vector<void*>> p;       // shift search directions p_kj, j >= 1
  std::vector<std::vector<void*>> ring;    // residual ring per system (kFoldMax + 1 fields)
  // operator stages: per stage and member the launch, or one grouped launch
  std::vector<std::vector<GroupMember>> stages;
  std::vector<int> stage_op;
  std::vector<GroupLaunch> glaunch;
  // device tables and state
  MsLaunch L;
  void* dtab = nullptr;                    // all pointer tables and sizes
  int* dcounter = nullptr;
  double* hstatus = nullptr;               // coherent host memory
  hipGraph_t graph = nullptr;
  hipGraphExec_t gexec = nullptr;
  int iters_total = 0;
  // mixed precision: low-precision sources / solutions and CG work fields
  std::vector<void*> bl, xl;
  void* lw[4] = {};
  int refine_iters = 0;
};

// the D^+D chain of op_D (with v2) + op_Ddag (with pre) on the member buffers
static void ms_build_stages(dwf_msbatch_s* m) {
  auto add = [&](int kop, auto fill) {
    std::vector<GroupMember> st(m->mem.size());
    for (size_t g = 0; g < m->mem.size(); ++g) {
      MsMember& M = m->mem[g];
      GroupMember& G = st[g];
      G.links = M.op->links;
      G.nbr = M.op->nbr;
      G.Vh = M.op->Vh;
      G.e.nrs = M.nrs;
      G.e.rstride = M.op->nelem;
      fill(M, G);
    }
    m->stages.push_back(st);
    m->stage_op.push_back(kop);
  };
  if (!m->mp_scalar)
    add(K_SITE, [](MsMember& M, GroupMember& G) {
      G.out = M.buf[MB_T0];
      G.in = M.buf[MB_P0];
      G.e.E = mat(M.op, M_PRE);
    });
  const bool pre = !m->mp_scalar;
  add(K_HOP, [pre](MsMember& M, GroupMember& G) {
    G.parity = 1;
    G.out = M.buf[MB_T1];
    G.in = pre ? M.buf[MB_T0] : M.buf[MB_P0];
    G.e.E = mat(M.op, M_E1);
  });
  add(K_HOP, [](MsMember& M, GroupMember& G) {
    G.parity = 0;
    G.out = M.buf[MB_U];
    G.in = M.buf[MB_T1];
    G.e.E = mat(M.op, M_E2);
    G.e.X = M.buf[MB_P0];
    G.e.xcoef = 1.0;
    G.e.out2 = M.buf[MB_U2];
    G.e.F = mat(M.op, M_DIT);
  });
  add(K_HOPDAG, [](MsMember& M, GroupMember& G) {
    G.parity = 1;
    G.out = M.buf[MB_T1];
    G.in = M.buf[MB_U2];
    G.e.E = mat(M.op, M_G1);
  });
  add(K_HOPDAG, [](MsMember& M, GroupMember& G) {
    G.parity = 0;
    G.out = M.buf[MB_S];
    G.in = M.buf[MB_T1];
    G.e.E = mat(M.op, M_G2);
    G.e.X = M.buf[MB_U];
    G.e.xcoef = 1.0;
  });
}

static int ms_run_operator(dwf_msbatch_s* m) {
  for (size_t i = 0; i < m->stages.size(); ++i) {
    if (m->mem.size() == 1) {
      const GroupMember& G = m->stages[i][0];
      HIPCHK(launch_dwf(m->prec, m->stage_op[i], G.out, G.in, G.links, G.nbr, G.parity, G.Vh,
                        m->Ls, G.e, g_ctx.stream));
    } else {
      HIPCHK(launch_dwf_group(m->glaunch[i], g_ctx.stream));
    }
  }
  return DWF_OK;
}

static int ms_run_iteration(dwf_msbatch_s* m) {
  RC(ms_run_operator(m));
  HIPCHK(launch_ms_blas(m->prec, m->L, 0, g_ctx.stream));
  HIPCHK(launch_ms_blas(m->prec, m->L, 1, g_ctx.stream));
  HIPCHK(launch_ms_ctrl(m->L, g_ctx.stream));
  HIPCHK(launch_ms_blas(m->prec, m->L, 2, g_ctx.stream));
  return DWF_OK;
}

static void ms_release(dwf_msbatch_s* m) {
  if (g_ctx.stream) (void)hipStreamSynchronize(g_ctx.stream);
  if (m->gexec) (void)hipGraphExecDestroy(m->gexec);
  if (m->graph) (void)hipGraphDestroy(m->graph);
  for (auto& gl : m->glaunch) free_dwf_group(&gl);
  for (auto& M : m->mem)
    for (auto& b : M.buf)
      if (b) (void)hipFree(b);
  for (auto& v : m->p)
    for (auto& q : v)
      if (q) (void)hipFree(q);
  for (auto& v : m->ring)
    for (auto& q : v)
      if (q) (void)hipFree(q);
  if (m->dtab) (void)hipFree(m->dtab);
  if (m->dcounter) (void)hipFree(m->dcounter);
  if (m->L.state) (void)hipFree(m->L.state);
  if (m->L.partial) (void)hipFree(m->L.partial);
  if (m->L.pp_res) (void)hipFree(m->L.pp_res);
  if (m->hstatus) (void)hipHostFree(m->hstatus);
  for (auto& q : m->bl)
    if (q) (void)hipFree(q);
  for (auto& q : m->xl)
    if (q) (void)hipFree(q);
  for (auto& q : m->lw)
    if (q) (void)hipFree(q);
}

extern "C" {
int dwf_msbatch_destroy(dwf_msbatch m);

int dwf_msbatch_create(int nsys, const dwf_fopr* ops, int nshift, int flags, dwf_msbatch* out) {
  CTX();
  if (!out || !ops || nsys < 1 || nsys > 64 || nshift < 1 || nshift > kMsShifts)
    return fail(DWF_ERR_INVALID_ARG, "batched multishift: 1 <= nsys <= 64, 1 <= nshift <= 32");
  *out = nullptr;
  auto* m = new dwf_msbatch_s;
  m->nsys = nsys;
  m->nshift = nshift;
  m->flags = flags;
  auto bad = [&](int code, const char* msg) {
    ms_release(m);
    delete m;
    return fail(code, msg);
  };
  for (int k = 0; k < nsys; ++k) {
    dwf_fopr_s* op = ops[k];
    if (!op) return bad(DWF_ERR_INVALID_ARG, "null operator");
    if (!op->have_config) return bad(DWF_ERR_NO_CONFIG, "set_config has not been called");
    if (op->prec != P64 && op->prec != P32)
      return bad(DWF_ERR_UNSUPPORTED, "batched multishift runs on FP64 or FP32 operators");
    if (k == 0) {
      m->prec = op->prec;
      m->Ls = op->Ls;
      m->mp_scalar = op->mp_scalar;
    } else if (op->prec != m->prec || op->Ls != m->Ls || op->mp_scalar != m->mp_scalar) {
      return bad(DWF_ERR_UNSUPPORTED,
                 "batched multishift: all operators need the same precision, Ls and b, c");
    }
    m->ops.push_back(op);
    int g = -1;
    for (size_t i = 0; i < m->mem.size(); ++i)
      if (m->mem[i].op == op) g = (int)i;
    if (g < 0) {
      m->mem.emplace_back();
      m->mem.back().op = op;
      g = (int)m->mem.size() - 1;
    }
    m->slot.push_back({g, (int)m->mem[g].sys.size()});
    m->mem[g].sys.push_back(k);
  }
  if (m->mem.size() > (size_t)kGroupMax)
    return bad(DWF_ERR_UNSUPPORTED, "batched multishift: at most 4 different operators");
  if (m->mem.size() > 1 && m->Ls != 12 && m->Ls != 16)
    return bad(DWF_ERR_UNSUPPORTED, "several lattices in one launch need Ls = 12 or 16");
  // member buffers
  for (auto& M : m->mem) {
    while ((1 << M.nrs) < (int)M.sys.size()) ++M.nrs;
    M.npad = 1 << M.nrs;
    const size_t fb = M.op->nelem * prec_bytes(M.op->prec) * M.npad;
    for (int b = 0; b < 7; ++b) {
      if ((b == MB_T0 && m->mp_scalar) || b == MB_R) continue;
      if (hipMalloc(&M.buf[b], fb) != hipSuccess) return bad(DWF_ERR_HIP, "device allocation failed");
      if (hipMemset(M.buf[b], 0, fb) != hipSuccess) return bad(DWF_ERR_HIP, "hipMemset failed");
    }
  }
  // shift directions
  m->p.assign(nsys, std::vector<void*>(nshift, nullptr));
  for (int k = 0; k < nsys; ++k)
    for (int j = 1; j < nshift; ++j)
      if (hipMalloc(&m->p[k][j], m->ops[k]->nelem * prec_bytes(m->prec)) != hipSuccess)
        return bad(DWF_ERR_HIP, "device allocation failed");
  m->ring.assign(nsys, std::vector<void*>(kFoldMax + 1, nullptr));
  for (int k = 0; k < nsys; ++k)
    for (int q = 0; q <= kFoldMax; ++q)
      if (hipMalloc(&m->ring[k][q], m->ops[k]->nelem * prec_bytes(m->prec)) != hipSuccess)
        return bad(DWF_ERR_HIP, "device allocation failed");
  ms_build_stages(m);
  if (m->mem.size() > 1) {
    m->glaunch.resize(m->stages.size());
    for (size_t i = 0; i < m->stages.size(); ++i)
      if (prepare_dwf_group(m->prec, m->stage_op[i], m->Ls, (int)m->mem.size(),
                            m->stages[i].data(), &m->glaunch[i]) != hipSuccess)
        return bad(DWF_ERR_UNSUPPORTED, "grouped launch could not be prepared");
  }
  // device tables: s p0 x0 [nsys], ring [nsys][kFoldMax + 1], x p [nsys][32], u [nsys],
  // n [nsys]
  constexpr int NR = kFoldMax + 1;
  const size_t ntab = 3 * nsys + NR * nsys + 2 * nsys * kMsShifts + nsys;
  const size_t ptab = sizeof(void*) * ntab;
  if (hipMalloc(&m->dtab, ptab + sizeof(size_t) * nsys) != hipSuccess)
    return bad(DWF_ERR_HIP, "device allocation failed");
  size_t maxn = 0;
  for (auto* op : m->ops) maxn = std::max(maxn, op->nelem);
  m->L.nsys = nsys;
  m->L.nblk = (int)std::min<size_t>(kRedBlocks, (maxn / 2 + 255) / 256);
  if (m->L.nblk < 1) m->L.nblk = 1;
  void** t = (void**)m->dtab;
  m->L.s = t;
  m->L.p0 = t + nsys;
  m->L.x0 = t + 2 * nsys;
  m->L.ring = t + 3 * nsys;
  m->L.x = t + 3 * nsys + NR * nsys;
  m->L.p = t + 3 * nsys + NR * nsys + nsys * kMsShifts;
  m->L.u = t + 3 * nsys + NR * nsys + 2 * nsys * kMsShifts;
  m->L.n = (const size_t*)((char*)m->dtab + ptab);
  if (hipMalloc(&m->L.state, sizeof(MsState) * nsys) != hipSuccess ||
      hipMalloc(&m->L.partial, sizeof(double) * nsys * kMsShifts * m->L.nblk) != hipSuccess ||
      hipMalloc(&m->L.pp_res, sizeof(double) * nsys * kMsShifts) != hipSuccess ||
      hipMalloc(&m->dcounter, sizeof(int)) != hipSuccess)
    return bad(DWF_ERR_HIP, "device allocation failed");
  if (hipMemset(m->L.partial, 0, sizeof(double) * nsys * kMsShifts * m->L.nblk) != hipSuccess ||
      hipMemset(m->L.pp_res, 0, sizeof(double) * nsys * kMsShifts) != hipSuccess ||
      hipMemset(m->dcounter, 0, sizeof(int)) != hipSuccess)
    return bad(DWF_ERR_HIP, "hipMemset failed");
  if (hipHostMalloc(&m->hstatus, sizeof(double) * 4, hipHostMallocCoherent) != hipSuccess)
    return bad(DWF_ERR_HIP, "pinned allocation failed");
  m->L.status = m->hstatus;
  m->L.counter = m->dcounter;
  // fixed part of the pointer tables (the member buffers)
  std::vector<void*> h(ntab, nullptr);
  std::vector<size_t> hn(nsys);
  for (int k = 0; k < nsys; ++k) {
    const MsMember& M = m->mem[m->slot[k].first];
    const size_t off = (size_t)m->slot[k].second * M.op->nelem * prec_bytes(m->prec);
    h[k] = (char*)M.buf[MB_S] + off;
    h[nsys + k] = (char*)M.buf[MB_P0] + off;
    h[3 * nsys + NR * nsys + 2 * nsys * kMsShifts + k] = (char*)M.buf[MB_U] + off;   // D p0
    for (int q = 0; q < NR; ++q) h[3 * nsys + k * NR + q] = m->ring[k][q];
    for (int j = 1; j < nshift; ++j)
      h[3 * nsys + NR * nsys + nsys * kMsShifts + k * kMsShifts + j] = m->p[k][j];
    hn[k] = M.op->nelem;
  }
  if (hipMemcpy(m->dtab, h.data(), ptab, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy((char*)m->dtab + ptab, hn.data(), sizeof(size_t) * nsys,
                hipMemcpyHostToDevice) != hipSuccess)
    return bad(DWF_ERR_HIP, "table upload failed");
  *out = m;
  return DWF_OK;
}

int dwf_msbatch_destroy(dwf_msbatch m) {
  if (!m) return DWF_OK;
  ms_release(m);
  delete m;
  return DWF_OK;
}

int dwf_msbatch_iterations(dwf_msbatch m) { return m ? m->iters_total : -1; }

}  // extern "C"

// CG on (A + sigma) e = r of op (mode DdagD, any precision) from e = 0 to
// |r|^2/|b|^2 < tol (ASolver_CG's iteration with the shift added to the
// operator); t1, t2, t3: work fields of op.  Returns the iterations done.
static int shifted_cg(dwf_fopr_s* op, double sigma, void* e, const void* b, double tol,
                      int max_iter, void* r, void* p, void* s, int* iters) {
  hipStream_t st = g_ctx.stream;
  const size_t n = op->nelem;
  double bb;
  HIPCHK(launch_norm2(op->prec, b, n, g_ctx.red, 0, st));
  RC(fetch(1, &bb));
  HIPCHK(launch_set_zero(op->prec, e, n, st));
  HIPCHK(launch_copy(op->prec, r, b, n, st));
  HIPCHK(launch_copy(op->prec, p, b, n, st));
  double rr = bb;
  int it = 0;
  for (; it < max_iter && rr > tol * bb; ++it) {
    RC(mult_impl(op, s, p, "DdagD", false));
    double pa;
    HIPCHK(launch_axpy_dot(op->prec, s, sigma, p, n, g_ctx.red, 0, st));
    RC(fetch(1, &pa));
    const double a = rr / pa;
    double rr_new;
    HIPCHK(launch_cg_xr(op->prec, e, r, p, s, a, n, g_ctx.red, 0, st));
    RC(fetch(1, &rr_new));
    HIPCHK(launch_aypx(op->prec, rr_new / rr, p, r, n, st));
    rr = rr_new;
  }
  *iters = it;
  return DWF_OK;
}

extern "C" {

// Mixed precision: the batched multishift on the low-precision operators of m
// to stop_cond_low, then every (system, shift) refined in FP64 by the
// Richardson / Mixedprec scheme of asolver_Richardson-tmpl.h:131-213 and
// aprecond_Mixedprec-tmpl.h:55-108 on the shifted operator: r = b - (A + s) x
// in FP64, e = (A + s)^-1 (r / |r|) by CG on the low-precision operator to
// stop_cond_inner, x += |r| e, until |r|^2 / |b|^2 < stop_cond.
int dwf_msbatch_solve_mixed(dwf_msbatch m, const dwf_fopr* ops64, const double* sigma, void** x,
                            const void* const* b, int max_iter, double stop_cond,
                            double stop_cond_low, double stop_cond_inner, int* nconv,
                            double* diff) {
  CTX();
  if (!m || !ops64 || !sigma || !x || !b || !nconv || !diff)
    return fail(DWF_ERR_INVALID_ARG, "invalid argument");
  if (m->prec != P32) return fail(DWF_ERR_UNSUPPORTED, "the batched solver must run in FP32");
  const int nsys = m->nsys, nshift = m->nshift;
  hipStream_t st = g_ctx.stream;
  for (int k = 0; k < nsys; ++k) {
    dwf_fopr_s* o = ops64[k];
    if (!o || o->prec != P64 || o->nelem != m->ops[k]->nelem || o->mode != "DdagD")
      return fail(DWF_ERR_INVALID_ARG, "ops64[k]: the FP64 operator (mode DdagD) of system k");
  }
  // low-precision sources and solutions, kept by the solver object
  if (m->xl.empty()) {
    m->bl.assign(nsys, nullptr);
    m->xl.assign((size_t)nsys * nshift, nullptr);
    for (int k = 0; k < nsys; ++k) {
      const size_t fb = m->ops[k]->nelem * prec_bytes(P32);
      HIPCHK(hipMalloc(&m->bl[k], fb));
      for (int j = 0; j < nshift; ++j) HIPCHK(hipMalloc(&m->xl[k * nshift + j], fb));
    }
  }
  for (int k = 0; k < nsys; ++k)
    HIPCHK(launch_convert(P32, m->bl[k], P64, b[k], m->ops[k]->nelem, st));
  std::vector<int> nc(nsys);
  std::vector<double> dl(nsys);
  const int rc = dwf_msbatch_solve(m, sigma, m->xl.data(), (const void* const*)m->bl.data(),
                                   max_iter, stop_cond_low, nc.data(), dl.data());
  if (rc != DWF_OK && rc != DWF_ERR_NOT_CONVERGED) return rc;
  m->refine_iters = 0;
  bool all = true;
  for (int k = 0; k < nsys; ++k) {
    dwf_fopr_s* o = ops64[k];
    dwf_fopr_s* ol = m->ops[k];
    const size_t n = o->nelem;
    void *t64, *rl, *el, *w1, *w2, *w3;
    RC(get_work(o, 4, &t64));
    RC(get_work(ol, 4, &rl));
    if (!m->lw[0]) {
      size_t mx = 0;
      for (auto* q : m->ops) mx = std::max(mx, q->nelem);
      for (auto& q : m->lw) HIPCHK(hipMalloc(&q, mx * prec_bytes(P32)));
    }
    el = m->lw[0];
    w1 = m->lw[1];
    w2 = m->lw[2];
    w3 = m->lw[3];
    double bn;
    HIPCHK(launch_norm2(P64, b[k], n, g_ctx.red, 0, st));
    RC(fetch(1, &bn));
    double dmax = 0.0;
    for (int j = 0; j < nshift; ++j) {
      void* xj = x[k * nshift + j];
      HIPCHK(launch_convert(P64, xj, P32, m->xl[k * nshift + j], n, st));
      double rr = 0;
      int outer = 0;
      for (;; ++outer) {
        // r = b - (A + sigma_j) x  (as (A + s) x - b, norm, then scaled by -1/|r|)
        RC(mult_impl(o, t64, xj, "DdagD", false));
        HIPCHK(launch_axpy(P64, t64, sigma[j], xj, n, st));
        HIPCHK(launch_sub_norm2(P64, t64, t64, b[k], n, g_ctx.red, 0, st));
        RC(fetch(1, &rr));
        if (rr < stop_cond * bn || outer >= 100) break;
        const double rn = std::sqrt(rr);
        HIPCHK(launch_convert_scaled(P32, rl, P64, t64, -1.0 / rn, n, st));
        int it = 0;
        RC(shifted_cg(ol, sigma[j], el, rl, stop_cond_inner, 100000, w1, w2, w3, &it));
        m->refine_iters += it;
        HIPCHK(launch_axpy_from(P64, xj, rn, P32, el, n, st));
      }
      const double d1 = std::sqrt(rr / bn);
      if (!(rr < stop_cond * bn)) all = false;
      dmax = std::max(dmax, d1);
    }
    diff[k] = dmax;
    nconv[k] = nc[k];
  }
  return all ? DWF_OK : fail(DWF_ERR_NOT_CONVERGED, "mixed batched multishift: not converged");
}

int dwf_msbatch_refine_iterations(dwf_msbatch m) { return m ? m->refine_iters : -1; }

int dwf_msbatch_solve(dwf_msbatch m, const double* sigma, void** x, const void* const* b,
                      int max_iter, double stop_cond, int* nconv, double* diff) {
  CTX();
  if (!m || !sigma || !x || !b || !nconv || !diff || max_iter < 1)
    return fail(DWF_ERR_INVALID_ARG, "invalid argument");
  for (auto* op : m->ops)
    if (op->mode != "DdagD")
      return fail(DWF_ERR_UNSUPPORTED, "batched multishift solves DdagD (set_mode(\"DdagD\"))");
  hipStream_t st = g_ctx.stream;
  const int nsys = m->nsys, nshift = m->nshift;
  const size_t eb = prec_bytes(m->prec);
  // |b|^2 and the start: r = p0 = p_j = b, x_j = 0
  std::vector<MsState> hs(nsys);
  for (int k = 0; k < nsys; ++k) {
    dwf_fopr_s* op = m->ops[k];
    double bn;
    HIPCHK(launch_norm2(m->prec, b[k], op->nelem, g_ctx.red, 0, st));
    RC(fetch(1, &bn));
    if (!(bn > 0)) return fail(DWF_ERR_INVALID_ARG, "zero source");
    const MsMember& M = m->mem[m->slot[k].first];
    const size_t off = (size_t)m->slot[k].second * op->nelem * eb;
    HIPCHK(launch_copy(m->prec, m->ring[k][0], b[k], op->nelem, st));
    HIPCHK(launch_copy(m->prec, (char*)M.buf[MB_P0] + off, b[k], op->nelem, st));
    for (int j = 0; j < nshift; ++j) {
      if (j > 0) HIPCHK(launch_copy(m->prec, m->p[k][j], b[k], op->nelem, st));
      HIPCHK(launch_set_zero(m->prec, x[k * nshift + j], op->nelem, st));
    }
    MsState& S = hs[k];
    std::memset(&S, 0, sizeof(S));
    S.rr = bn;
    S.alpha_p = 0.0;
    S.beta_p = 1.0;
    S.snorm = 1.0 / bn;
    S.sigma0 = sigma[0];
    S.stop = stop_cond;
    S.active = 1;
    S.nshift2 = nshift;
    S.nconv = -1;
    S.max_iter = max_iter;
    S.nfold = nshift;   // q0 = m = 0: the block starts at ring slot 0 (r = b)
    for (int j = 0; j < kMsShifts; ++j) {
      S.zeta1[j] = S.zeta2[j] = 1.0;
      S.csh2[j] = j < nshift ? sigma[j] - sigma[0] : 0.0;
      S.ppe[j] = bn;
      S.fc[0][j][0] = S.fc[1][j][0] = 1.0;
    }
  }
  // solution pointers of this solve
  std::vector<void*> hx(2 * nsys + nsys * kMsShifts, nullptr);   // x0 [nsys], x [nsys][32]
  for (int k = 0; k < nsys; ++k) {
    hx[k] = x[k * nshift];
    for (int j = 1; j < nshift; ++j) hx[nsys + k * kMsShifts + j] = x[k * nshift + j];
  }
  HIPCHK(hipMemcpyAsync(m->L.x0, hx.data(), sizeof(void*) * nsys, hipMemcpyHostToDevice, st));
  HIPCHK(hipMemcpyAsync(m->L.x, hx.data() + nsys, sizeof(void*) * nsys * kMsShifts,
                        hipMemcpyHostToDevice, st));
  HIPCHK(hipMemcpyAsync(m->L.state, hs.data(), sizeof(MsState) * nsys, hipMemcpyHostToDevice, st));
  // |p0|^2 of the first iteration's pa: p0 = b (slot 0 of pp_res; the update of
  // p0 writes it afterwards)
  std::vector<double> hpp((size_t)nsys * kMsShifts, 0.0);
  for (int k = 0; k < nsys; ++k) hpp[(size_t)k * kMsShifts] = hs[k].rr;
  HIPCHK(hipMemcpyAsync(m->L.pp_res, hpp.data(), sizeof(double) * nsys * kMsShifts,
                        hipMemcpyHostToDevice, st));
  HIPCHK(hipStreamSynchronize(st));
  int cnt0 = 0;
  HIPCHK(hipMemcpy(&cnt0, m->dcounter, sizeof(int), hipMemcpyDeviceToHost));
  m->hstatus[0] = (double)cnt0 - 1.0;
  m->hstatus[1] = nsys;
  const bool use_graph = (m->flags & DWF_MSBATCH_GRAPH) != 0;
  if (use_graph && !m->gexec) {
    HIPCHK(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
    const int rc = ms_run_iteration(m);
    hipGraph_t g = nullptr;
    const hipError_t e = hipStreamEndCapture(st, &g);
    if (rc != DWF_OK) {
      if (g) (void)hipGraphDestroy(g);
      return rc;
    }
    HIPCHK(e);
    m->graph = g;
    HIPCHK(hipGraphInstantiate(&m->gexec, m->graph, nullptr, nullptr, 0));
  }
  // iterations: the host enqueues iteration i + 1 before it waits for the
  // status of iteration i (kernels of inactive systems return at once)
  volatile double* status = m->hstatus;
  int it = 0;
  for (; it < max_iter; ++it) {
    if (use_graph) HIPCHK(hipGraphLaunch(m->gexec, st));
    else RC(ms_run_iteration(m));
    if (it >= 1) {
      const double want = (double)(cnt0 + it - 1);
      for (long spin = 0; status[0] < want; ++spin)
        if ((spin & 1023) == 1023 && hipStreamQuery(st) == hipSuccess) break;   // stream idle
      if (status[1] < 0.5) break;
    }
  }
  HIPCHK(hipStreamSynchronize(st));
  HIPCHK(hipMemcpy(hs.data(), m->L.state, sizeof(MsState) * nsys, hipMemcpyDeviceToHost));
  m->iters_total = 0;
  bool all = true;
  for (int k = 0; k < nsys; ++k) {
    nconv[k] = hs[k].conv ? hs[k].nconv : -1;
    all = all && hs[k].conv;
    m->iters_total = std::max(m->iters_total, hs[k].nconv + 1);
  }
  // diff (the reference's check after the iteration): max over the shifts of
  // |(A + sigma_j) x_j - b| / |b|, with the operator itself
  for (int k = 0; k < nsys; ++k) {
    dwf_fopr_s* op = m->ops[k];
    void* s;
    RC(get_work(op, 4, &s));
    double dmax = -1.0;
    for (int j = 0; j < nshift; ++j) {
      RC(mult_impl(op, s, x[k * nshift + j], op->mode, false));
      HIPCHK(launch_axpy(op->prec, s, sigma[j], x[k * nshift + j], op->nelem, st));
      double d2;
      HIPCHK(launch_sub_norm2(op->prec, s, s, b[k], op->nelem, g_ctx.red, 0, st));
      RC(fetch(1, &d2));
      const double d1 = std::sqrt(d2 * hs[k].snorm);
      if (d1 > dmax) dmax = d1;
    }
    diff[k] = dmax;
  }
  return all ? DWF_OK : fail(DWF_ERR_NOT_CONVERGED, "batched multishift: not converged");
}
}  // extern "C"

// ---- DWF fermion force (FP64) -----------------------------------------------
// AForce_F_Domainwall_5din_eo (aforce_F_Domainwall_5din_eo-tmpl.h:299-451) with
// the lexical tensor product of aforce_F_Domainwall_5din-tmpl.h:338-933 applied
// to the two parity fields.
struct dwf_force_s {
  dwf_fopr op = nullptr;
  std::string mode = "D";
  void* w[8] = {};
};

static int force_field(dwf_force_s* f, int i, void** out) {
  if (!f->w[i]) HIPCHK(hipMalloc(&f->w[i], f->op->nelem * sizeof(double)));
  *out = f->w[i];
  return DWF_OK;
}

// F (+)= scale * udiv1_D (dag = false) or udiv1_Ddag, where udiv1 = -(sum of the
// tensor products)
static int force_udiv1_eo(dwf_force_s* f, double* F, const void* zeta, const void* eta, bool dag,
                          double scale, bool accumulate) {
  dwf_fopr op = f->op;
  void *a_o, *b_e, *b_o, *t, *y_e, *y_o;
  RC(force_field(f, 0, &a_o));
  RC(force_field(f, 1, &b_e));
  RC(force_field(f, 2, &b_o));
  RC(force_field(f, 3, &t));
  RC(force_field(f, 4, &y_e));
  RC(force_field(f, 5, &y_o));
  const void* plain = dag ? zeta : eta;   // (x, Doo^-1 Doe x)
  const void* dside = dag ? eta : zeta;   // (Dee^-+ x, Doo^-+ Deo^+ Dee^-+ x)
  RC(mult_impl(op, t, plain, "Doe", false));
  RC(mult_impl(op, a_o, t, "Doo_inv", false));
  RC(mult_impl(op, b_e, dside, "Dee_inv", true));
  RC(mult_impl(op, t, b_e, "Deo", true));
  RC(mult_impl(op, b_o, t, "Doo_inv", true));
  RC(run_site(op, y_e, plain, M_M5F));
  RC(run_site(op, y_o, a_o, M_M5F));
  const int* N = op->N;
  if (!dag) {  // W = (1 - g_mu) y(x + mu), Z = dagger side
    HIPCHK(launch_force_tensor(F, (const double*)b_e, (const double*)y_o, op->nbr, 0, N[0], N[1],
                               N[2], N[3], op->bc, op->Ls, -1.0, -scale, accumulate, g_ctx.stream));
    HIPCHK(launch_force_tensor(F, (const double*)b_o, (const double*)y_e, op->nbr, 1, N[0], N[1],
                               N[2], N[3], op->bc, op->Ls, -1.0, -scale, accumulate, g_ctx.stream));
  } else {     // W = (1 + g_mu) eta'(x + mu), Z = M5-weighted zeta
    HIPCHK(launch_force_tensor(F, (const double*)y_e, (const double*)b_o, op->nbr, 0, N[0], N[1],
                               N[2], N[3], op->bc, op->Ls, +1.0, -scale, accumulate, g_ctx.stream));
    HIPCHK(launch_force_tensor(F, (const double*)y_o, (const double*)b_e, op->nbr, 1, N[0], N[1],
                               N[2], N[3], op->bc, op->Ls, +1.0, -scale, accumulate, g_ctx.stream));
  }
  return DWF_OK;
}

static int force_udiv1_mode(dwf_force_s* f, const std::string& m, double* F, const void* zeta,
                            const void* eta, double scale, bool accumulate) {
  if (m == "D") return force_udiv1_eo(f, F, zeta, eta, false, scale, accumulate);
  if (m == "Ddag") return force_udiv1_eo(f, F, zeta, eta, true, scale, accumulate);
  void* g;
  RC(force_field(f, 6, &g));
  if (m == "H") {
    RC(run_site(f->op, g, zeta, M_G5R));
    return force_udiv1_eo(f, F, g, eta, false, scale, accumulate);
  }
  if (m == "Hdag") {
    RC(run_site(f->op, g, eta, M_G5R));
    return force_udiv1_eo(f, F, zeta, g, true, scale, accumulate);
  }
  return fail(DWF_ERR_UNKNOWN_MODE, "irrelevant force mode: " + m);
}

// force_udiv (aforce_F_Domainwall_5din_eo-tmpl.h:299-315):
// F (+)= scale [udiv1_D(D eta, eta) + udiv1_Ddag(eta, D eta)]
static int force_udiv_impl(dwf_force_s* f, double* F, const void* eta, double scale,
                           bool accumulate) {
  void* z;
  RC(force_field(f, 7, &z));
  RC(mult_impl(f->op, z, eta, "D", false));
  RC(force_udiv1_eo(f, F, z, eta, false, scale, accumulate));
  return force_udiv1_eo(f, F, eta, z, true, scale, true);
}

extern "C" {

int dwf_device_alloc(size_t bytes, void** p) {
  CTX();
  if (!p) return fail(DWF_ERR_INVALID_ARG, "null argument");
  HIPCHK(hipMalloc(p, bytes));
  return DWF_OK;
}
int dwf_device_free(void* p) {
  if (!p) return DWF_OK;
  if (g_ctx.stream) (void)hipStreamSynchronize(g_ctx.stream);
  HIPCHK(hipFree(p));
  return DWF_OK;
}
int dwf_memcpy_to_device(void* dst, const void* src, size_t bytes) {
  CTX();
  HIPCHK(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, g_ctx.stream));
  HIPCHK(hipStreamSynchronize(g_ctx.stream));
  return DWF_OK;
}
int dwf_memcpy_to_host(void* dst, const void* src, size_t bytes) {
  CTX();
  HIPCHK(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, g_ctx.stream));
  HIPCHK(hipStreamSynchronize(g_ctx.stream));
  return DWF_OK;
}

int dwf_force_create(dwf_fopr op, dwf_force* out) {
  CTX();
  if (!op || !out) return fail(DWF_ERR_INVALID_ARG, "null argument");
  if (op->prec != P64) return fail(DWF_ERR_UNSUPPORTED, "the force runs on the FP64 operator");
  if (op->alpha != 1.0) return fail(DWF_ERR_UNSUPPORTED, "the force needs parameter_alpha = 1");
  // mult_generator recovers U from the packed forward links by their +-1 boundary factor
  for (int mu = 0; mu < 4; ++mu)
    if (op->bc[mu] != 1 && op->bc[mu] != -1)
      return fail(DWF_ERR_UNSUPPORTED, "the force needs boundary conditions of +-1");
  dwf_force_s* f = new dwf_force_s;
  f->op = op;
  *out = f;
  return DWF_OK;
}
int dwf_force_destroy(dwf_force f) {
  if (!f) return DWF_OK;
  if (g_ctx.stream) (void)hipStreamSynchronize(g_ctx.stream);
  for (auto& w : f->w)
    if (w) (void)hipFree(w);
  delete f;
  return DWF_OK;
}
int dwf_force_set_mode(dwf_force f, const char* mode) {
  if (!f || !mode) return fail(DWF_ERR_INVALID_ARG, "null argument");
  std::string m(mode);
  if (m != "D" && m != "Ddag" && m != "H" && m != "Hdag")
    return fail(DWF_ERR_UNKNOWN_MODE, "irrelevant force mode: " + m);
  f->mode = m;
  return DWF_OK;
}
int dwf_force_udiv1(dwf_force f, double* F, const void* zeta, const void* eta) {
  CTX();
  if (!f || !F || !zeta || !eta) return fail(DWF_ERR_INVALID_ARG, "null argument");
  if (!f->op->have_config) return fail(DWF_ERR_NO_CONFIG, "set_config has not been called");
  return force_udiv1_mode(f, f->mode, F, zeta, eta, 1.0, false);
}
int dwf_force_udiv(dwf_force f, double* F, const void* eta) {
  CTX();
  if (!f || !F || !eta) return fail(DWF_ERR_INVALID_ARG, "null argument");
  if (!f->op->have_config) return fail(DWF_ERR_NO_CONFIG, "set_config has not been called");
  return force_udiv_impl(f, F, eta, 1.0, false);
}
int dwf_force_mult_generator(dwf_force f, double* F) {
  CTX();
  if (!f || !F) return fail(DWF_ERR_INVALID_ARG, "null argument");
  const int* N = f->op->N;
  HIPCHK(launch_mult_generator(F, (const double*)f->op->links, N[0], N[1], N[2], N[3], f->op->bc,
                               g_ctx.stream));
  return DWF_OK;
}
int dwf_force_core(dwf_force f, double* F, const void* eta) {
  RC(dwf_force_udiv(f, F, eta));
  return dwf_force_mult_generator(f, F);
}
int dwf_force_core1(dwf_force f, double* F, const void* zeta, const void* eta) {
  RC(dwf_force_udiv1(f, F, zeta, eta));
  return dwf_force_mult_generator(f, F);
}

}  // extern "C"

// ---- rational functions of A^dag A and AForce_F_Rational_Ratio (FP64) ----------

static void force_release(dwf_force_s* f) {
  for (auto& w : f->w)
    if (w) (void)hipFree(w);
  for (auto& w : f->w) w = nullptr;
}

static int check_rational(const dwf_rational_params* r) {
  if (!r || r->np < 1 || !r->bl || !r->cl || r->max_iter < 1 || !(r->stop_cond > 0))
    return fail(DWF_ERR_INVALID_ARG, "invalid rational parameters");
  return DWF_OK;
}

// v = a0 b + sum_i bl_i x_i in one pass (afopr_Rational-tmpl.h:211-215)
static int rational_combine(dwf_fopr op, void* v, const void* b, double a0,
                            const std::vector<void*>& x, const double* bl) {
  const int m = (int)x.size();
  RC(ensure_shift_scratch(m));
  hipStream_t st = g_ctx.stream;
  double* hc = g_ctx.host_coef;
  void** hp = (void**)(g_ctx.host_coef + 3 * g_ctx.sh_cap);
  HIPCHK(hipStreamSynchronize(st));   // the pinned staging area is free again
  for (int i = 0; i < m; ++i) {
    hc[i] = bl[i];
    hp[i] = x[i];
  }
  HIPCHK(hipMemcpyAsync(g_ctx.sh_coef, hc, sizeof(double) * m, hipMemcpyHostToDevice, st));
  HIPCHK(hipMemcpyAsync(g_ctx.sh_ptrs, hp, sizeof(void*) * m, hipMemcpyHostToDevice, st));
  HIPCHK(launch_lincomb(v, b, a0, g_ctx.sh_ptrs, g_ctx.sh_coef, m, op->nelem, st));
  HIPCHK(hipStreamSynchronize(st));
  return DWF_OK;
}

static int alloc_fields(dwf_fopr op, std::vector<void*>& v, int n) {
  v.assign(n, nullptr);
  for (int i = 0; i < n; ++i) HIPCHK(hipMalloc(&v[i], op->nelem * sizeof(double)));
  return DWF_OK;
}
static void free_fields(std::vector<void*>& v) {
  for (auto& p : v)
    if (p) (void)hipFree(p);
  v.clear();
}

// shift solve of AFopr_Rational / AForce_F_Rational_Ratio on mode DdagD; a
// solver that does not converge is recorded in *status and the caller goes on,
// as the reference does (it logs Nconv and diff and continues)
static int rational_solve(dwf_fopr op, const dwf_rational_params& r, std::vector<void*>& x,
                          const void* b, int* nconv, double* diff, int* status) {
  op->mode = "DdagD";
  const int rc = dwf_shiftsolver_cg_solve(op, r.max_iter, r.stop_cond, r.np, r.cl, x.data(), b,
                                          nconv, diff);
  if (rc == DWF_ERR_NOT_CONVERGED) {
    *status = rc;
    return DWF_OK;
  }
  return rc;
}

struct dwf_ratio_force_s {
  dwf_fopr op1 = nullptr, op2 = nullptr;
  dwf_force_s f1, f2;
  dwf_rational_params r1{}, r2{};
  std::vector<double> bl1, cl1, bl2, cl2;
  std::vector<void*> eta1, zeta1, eta2;
  void* v1 = nullptr;
  void* v2 = nullptr;
};

static void ratio_release(dwf_ratio_force_s* rf) {
  if (g_ctx.stream) (void)hipStreamSynchronize(g_ctx.stream);
  free_fields(rf->eta1);
  free_fields(rf->zeta1);
  free_fields(rf->eta2);
  if (rf->v1) (void)hipFree(rf->v1);
  if (rf->v2) (void)hipFree(rf->v2);
  rf->v1 = rf->v2 = nullptr;
  force_release(&rf->f1);
  force_release(&rf->f2);
}

// aforce_F_Rational_Ratio-tmpl.h:345-447
static int ratio_force_udiv(dwf_ratio_force_s* rf, double* F, const void* eta) {
  dwf_fopr op1 = rf->op1;
  int status = DWF_OK, nconv;
  double diff;
  // m_eta1 = (A^dag A + cl1)^-1 eta;  m_v1 = r1(A^dag A) eta
  RC(rational_solve(op1, rf->r1, rf->eta1, eta, &nconv, &diff, &status));
  RC(rational_combine(op1, rf->v1, eta, rf->r1.a0, rf->eta1, rf->r1.bl));
  // m_eta2 = (B^dag B + cl2)^-1 m_v1;  m_v2 = r2(B^dag B) m_v1
  RC(rational_solve(rf->op2, rf->r2, rf->eta2, rf->v1, &nconv, &diff, &status));
  RC(rational_combine(op1, rf->v2, rf->v1, rf->r2.a0, rf->eta2, rf->r2.bl));
  // m_zeta1 = (A^dag A + cl1)^-1 m_v2
  RC(rational_solve(op1, rf->r1, rf->zeta1, rf->v2, &nconv, &diff, &status));
  // denominator: sum_i bl2_i force_udiv^B(eta2_i)
  for (int i = 0; i < rf->r2.np; ++i)
    RC(force_udiv_impl(&rf->f2, F, rf->eta2[i], rf->r2.bl[i], i > 0));
  // numerator: the four H / Hdag terms per pole
  op1->mode = "H";
  for (int i = 0; i < rf->r1.np; ++i) {
    const double b = rf->r1.bl[i];
    RC(mult_impl(op1, rf->v1, rf->eta1[i], "H", false));
    RC(mult_impl(op1, rf->v2, rf->zeta1[i], "H", false));
    RC(force_udiv1_mode(&rf->f1, "H", F, rf->v1, rf->zeta1[i], b, true));
    RC(force_udiv1_mode(&rf->f1, "H", F, rf->v2, rf->eta1[i], b, true));
    RC(force_udiv1_mode(&rf->f1, "Hdag", F, rf->eta1[i], rf->v2, b, true));
    RC(force_udiv1_mode(&rf->f1, "Hdag", F, rf->zeta1[i], rf->v1, b, true));
  }
  rf->f1.mode = "Hdag";
  if (status != DWF_OK) return fail(status, "AForce_F_Rational_Ratio: a shift solver did not converge");
  return DWF_OK;
}

extern "C" {

int dwf_rational_mult(dwf_fopr op, const dwf_rational_params* r, void* v, const void* b,
                      int* nconv, double* diff) {
  CTX();
  if (!op || !v || !b || !nconv || !diff) return fail(DWF_ERR_INVALID_ARG, "null argument");
  RC(check_rational(r));
  if (op->prec != P64) return fail(DWF_ERR_UNSUPPORTED, "rational functions run on the FP64 operator");
  if (!op->have_config) return fail(DWF_ERR_NO_CONFIG, "set_config has not been called");
  std::vector<void*> x;
  int rc = alloc_fields(op, x, r->np);
  int status = DWF_OK;
  if (rc == DWF_OK) rc = rational_solve(op, *r, x, b, nconv, diff, &status);
  if (rc == DWF_OK) rc = rational_combine(op, v, b, r->a0, x, r->bl);
  if (g_ctx.stream) (void)hipStreamSynchronize(g_ctx.stream);
  free_fields(x);
  if (rc != DWF_OK) return rc;
  if (status != DWF_OK) return fail(status, "AFopr_Rational: the shift solver did not converge");
  return DWF_OK;
}

int dwf_force_rational_ratio_create(dwf_fopr fopr1, dwf_fopr fopr2,
                                    const dwf_rational_params* numerator,
                                    const dwf_rational_params* denominator,
                                    dwf_ratio_force* out) {
  CTX();
  if (!fopr1 || !fopr2 || !out) return fail(DWF_ERR_INVALID_ARG, "null argument");
  RC(check_rational(numerator));
  RC(check_rational(denominator));
  for (dwf_fopr op : {fopr1, fopr2}) {
    if (op->prec != P64) return fail(DWF_ERR_UNSUPPORTED, "the force runs on FP64 operators");
    if (op->alpha != 1.0) return fail(DWF_ERR_UNSUPPORTED, "the force needs parameter_alpha = 1");
    for (int mu = 0; mu < 4; ++mu)
      if (op->bc[mu] != 1 && op->bc[mu] != -1)
        return fail(DWF_ERR_UNSUPPORTED, "the force needs boundary conditions of +-1");
  }
  for (int mu = 0; mu < 4; ++mu)
    if (fopr1->N[mu] != fopr2->N[mu])
      return fail(DWF_ERR_INVALID_ARG, "numerator and denominator lattices differ");
  if (fopr1->Ls != fopr2->Ls || fopr1->nelem != fopr2->nelem)
    return fail(DWF_ERR_INVALID_ARG, "numerator and denominator fields differ");
  dwf_ratio_force_s* rf = new dwf_ratio_force_s;
  rf->op1 = fopr1;
  rf->op2 = fopr2;
  rf->f1.op = fopr1;
  rf->f2.op = fopr2;
  rf->bl1.assign(numerator->bl, numerator->bl + numerator->np);
  rf->cl1.assign(numerator->cl, numerator->cl + numerator->np);
  rf->bl2.assign(denominator->bl, denominator->bl + denominator->np);
  rf->cl2.assign(denominator->cl, denominator->cl + denominator->np);
  rf->r1 = *numerator;
  rf->r1.bl = rf->bl1.data();
  rf->r1.cl = rf->cl1.data();
  rf->r2 = *denominator;
  rf->r2.bl = rf->bl2.data();
  rf->r2.cl = rf->cl2.data();
  int rc = alloc_fields(fopr1, rf->eta1, numerator->np);
  if (rc == DWF_OK) rc = alloc_fields(fopr1, rf->zeta1, numerator->np);
  if (rc == DWF_OK) rc = alloc_fields(fopr1, rf->eta2, denominator->np);
  if (rc == DWF_OK && hipMalloc(&rf->v1, fopr1->nelem * sizeof(double)) != hipSuccess)
    rc = fail(DWF_ERR_HIP, "device allocation failed");
  if (rc == DWF_OK && hipMalloc(&rf->v2, fopr1->nelem * sizeof(double)) != hipSuccess)
    rc = fail(DWF_ERR_HIP, "device allocation failed");
  if (rc != DWF_OK) {
    ratio_release(rf);
    delete rf;
    return rc;
  }
  *out = rf;
  return DWF_OK;
}

int dwf_force_rational_ratio_destroy(dwf_ratio_force rf) {
  if (!rf) return DWF_OK;
  ratio_release(rf);
  delete rf;
  return DWF_OK;
}

int dwf_force_rational_ratio_udiv(dwf_ratio_force rf, double* F, const void* eta) {
  CTX();
  if (!rf || !F || !eta) return fail(DWF_ERR_INVALID_ARG, "null argument");
  if (!rf->op1->have_config || !rf->op2->have_config)
    return fail(DWF_ERR_NO_CONFIG, "set_config has not been called");
  return ratio_force_udiv(rf, F, eta);
}

int dwf_force_rational_ratio_core(dwf_ratio_force rf, double* F, const void* eta) {
  const int rc = dwf_force_rational_ratio_udiv(rf, F, eta);
  if (rc != DWF_OK && rc != DWF_ERR_NOT_CONVERGED) return rc;
  const int* N = rf->op1->N;
  HIPCHK(launch_mult_generator(F, (const double*)rf->op1->links, N[0], N[1], N[2], N[3],
                               rf->op1->bc, g_ctx.stream));
  return rc;
}

}  // extern "C"

// ---- smearing ---------------------------------------------------------------
struct dwf_smear_s {
  int N[4];
  double rho;
  int nsmear;
  size_t V;
  std::vector<double*> levels;
  bool have_config = false;
  double* itheta = nullptr;   // scratch of the force recursion (lazy)
};

int dwf_smear_create(const int dims[4], double rho, int nsmear, dwf_smear* out) {
  CTX();
  if (!dims || !out || nsmear < 0) return fail(DWF_ERR_INVALID_ARG, "invalid argument");
  dwf_smear_s* sm = new dwf_smear_s;
  sm->V = 1;
  for (int mu = 0; mu < 4; ++mu) {
    sm->N[mu] = dims[mu];
    sm->V *= dims[mu];
  }
  sm->rho = rho;
  sm->nsmear = nsmear;
  sm->levels.assign(nsmear + 1, nullptr);
  for (auto& l : sm->levels)
    if (hipMalloc(&l, sm->V * 4 * 18 * sizeof(double)) != hipSuccess) {
      dwf_smear_destroy(sm);
      return fail(DWF_ERR_HIP, "device allocation failed");
    }
  *out = sm;
  return DWF_OK;
}
int dwf_smear_destroy(dwf_smear sm) {
  if (!sm) return DWF_OK;
  if (g_ctx.stream) (void)hipStreamSynchronize(g_ctx.stream);
  for (auto l : sm->levels)
    if (l) (void)hipFree(l);
  if (sm->itheta) (void)hipFree(sm->itheta);
  delete sm;
  return DWF_OK;
}
static int smear_levels(dwf_smear sm) {
  for (int i = 0; i < sm->nsmear; ++i)
    HIPCHK(launch_stout(sm->levels[i + 1], sm->levels[i], sm->N[0], sm->N[1], sm->N[2], sm->N[3],
                        sm->rho, g_ctx.stream));
  sm->have_config = true;
  return DWF_OK;
}
int dwf_smear_set_config_device(dwf_smear sm, const double* U) {
  CTX();
  if (!sm || !U) return fail(DWF_ERR_INVALID_ARG, "null argument");
  HIPCHK(hipMemcpyAsync(sm->levels[0], U, sm->V * 4 * 18 * sizeof(double),
                        hipMemcpyDeviceToDevice, g_ctx.stream));
  return smear_levels(sm);
}
int dwf_smear_set_config(dwf_smear sm, const double* U_host) {
  CTX();
  if (!sm || !U_host) return fail(DWF_ERR_INVALID_ARG, "null argument");
  HIPCHK(hipMemcpyAsync(sm->levels[0], U_host, sm->V * 4 * 18 * sizeof(double),
                        hipMemcpyHostToDevice, g_ctx.stream));
  RC(smear_levels(sm));
  HIPCHK(hipStreamSynchronize(g_ctx.stream));
  return DWF_OK;
}
int dwf_smear_get_config(dwf_smear sm, int level, double* U_host) {
  CTX();
  if (!sm || !U_host || level < 0 || level > sm->nsmear)
    return fail(DWF_ERR_INVALID_ARG, "invalid argument");
  HIPCHK(hipMemcpyAsync(U_host, sm->levels[level], sm->V * 4 * 18 * sizeof(double),
                        hipMemcpyDeviceToHost, g_ctx.stream));
  HIPCHK(hipStreamSynchronize(g_ctx.stream));
  return DWF_OK;
}
const double* dwf_smear_config_device(dwf_smear sm, int level) {
  if (!sm || level < 0 || level > sm->nsmear) return nullptr;
  return sm->levels[level];
}
// Director_alt_Smear::force_udiv for one level (director_alt_Smear-tmpl.h:253-270 ->
// AForceSmear_APE::force_udiv); Sigma may alias Sigmap (each link is read and
// written by one thread only)
int dwf_smear_force_udiv(dwf_smear sm, int level, double* Sigma, const double* Sigmap) {
  CTX();
  if (!sm || !Sigma || !Sigmap) return fail(DWF_ERR_INVALID_ARG, "null argument");
  if (level < 0 || level >= sm->nsmear) return fail(DWF_ERR_INVALID_ARG, "no such smearing level");
  if (!sm->have_config) return fail(DWF_ERR_NO_CONFIG, "set_config has not been called");
  if (!sm->itheta) HIPCHK(hipMalloc(&sm->itheta, sm->V * 4 * 18 * sizeof(double)));
  HIPCHK(launch_stout_force(Sigma, sm->itheta, Sigmap, sm->levels[level], sm->N[0], sm->N[1],
                            sm->N[2], sm->N[3], sm->rho, g_ctx.stream));
  return DWF_OK;
}
// AForce_F_Smeared_alt::mult_jacobian (aforce_F_Smeared_alt-tmpl.h:170-191), in place
int dwf_smear_mult_jacobian(dwf_smear sm, double* F) {
  CTX();
  if (!sm || !F) return fail(DWF_ERR_INVALID_ARG, "null argument");
  for (int ismear = sm->nsmear - 1; ismear >= 0; --ismear)
    RC(dwf_smear_force_udiv(sm, ismear, F, F));
  return DWF_OK;
}
// AForce_F::mult_generator with the thin links (m_Ucp of AForce_F_Smeared_alt,
// aforce_F_Smeared_alt-tmpl.h:72-80; lib_alt_QXS/Force/Fermion/aforce_F.cpp:97-110)
int dwf_smear_mult_generator(dwf_smear sm, double* F) {
  CTX();
  if (!sm || !F) return fail(DWF_ERR_INVALID_ARG, "null argument");
  if (!sm->have_config) return fail(DWF_ERR_NO_CONFIG, "set_config has not been called");
  HIPCHK(launch_mult_generator_gauge(F, sm->levels[0], sm->V * 4, g_ctx.stream));
  return DWF_OK;
}

int dwf_fopr_set_config_smeared(dwf_fopr op, dwf_smear sm, const double* U_host) {
  CTX();
  if (!op || !sm) return fail(DWF_ERR_INVALID_ARG, "null argument");
  for (int mu = 0; mu < 4; ++mu)
    if (sm->N[mu] != op->N[mu]) return fail(DWF_ERR_INVALID_ARG, "lattice sizes differ");
  if (U_host) RC(dwf_smear_set_config(sm, U_host));
  RC(dwf_fopr_set_config_device(op, sm->levels[sm->nsmear]));
  HIPCHK(hipStreamSynchronize(g_ctx.stream));
  return DWF_OK;
}

// ---- parallel-transport convolution network (learned preconditioner) ----------
// Layers l = 0 .. L-1 map ch[l] to ch[l + 1] channels; the input is the even
// field (channel 0 of the even parity, odd parity zero), the output the even
// part of channel 0 of the last layer.  Weights: per layer [q][c_out][c_in]
// x 16 complex spin entries (re, im), bfloat16 on the device.
struct dwf_ptc_s {
  dwf_fopr op = nullptr;                 // FP32 operator: links, neighbours, Ls, Vh
  std::vector<int> ch;                   // channels, L + 1 entries
  std::vector<size_t> woff;              // float offset of each layer's weights
  size_t nw = 0;
  unsigned short* W = nullptr;           // device bf16
  std::vector<float*> act;               // L + 1 activations (2 parities each)
  float* g[2] = {nullptr, nullptr};      // backward ping-pong
  double* partial = nullptr;             // weight-gradient block partials
  double* G = nullptr;                   // one layer's gradient (device)
  int nblk = 0;
};

static dwf_fopr ptc_op(dwf_ptc_s* n) { return n->op; }

static unsigned short f32_to_bf16(float f) {   // round to nearest even
  uint32_t u;
  std::memcpy(&u, &f, 4);
  const uint32_t r = ((u >> 16) & 1u) + 0x7fffu;
  return (unsigned short)((u + r) >> 16);
}
static float bf16_to_f32(unsigned short h) {
  const uint32_t u = (uint32_t)h << 16;
  float f;
  std::memcpy(&f, &u, 4);
  return f;
}

static int ptc_forward(dwf_ptc_s* n, const void* in) {
  dwf_fopr op = n->op;
  hipStream_t st = g_ctx.stream;
  HIPCHK(launch_ptc_embed(n->act[0], (const float*)in, op->nelem, st));
  for (size_t l = 0; l + 1 < n->ch.size(); ++l)
    HIPCHK(launch_ptc_layer(false, n->act[l + 1], n->act[l], n->ch[l], n->ch[l + 1],
                            n->W + n->woff[l], n->ch[l], n->ch[l + 1], (const float*)op->links,
                            op->nbr, op->Vh, op->Ls, st));
  return DWF_OK;
}

static int ptc_apply_impl(dwf_ptc_s* n, void* out, const void* in) {
  RC(ptc_forward(n, in));
  HIPCHK(hipMemcpyAsync(out, n->act.back(), n->op->nelem * sizeof(float),
                        hipMemcpyDeviceToDevice, g_ctx.stream));
  return DWF_OK;
}

extern "C" {

int dwf_ptc_destroy(dwf_ptc n) {
  if (!n) return DWF_OK;
  if (g_ctx.stream) (void)hipStreamSynchronize(g_ctx.stream);
  if (n->W) (void)hipFree(n->W);
  for (auto* a : n->act)
    if (a) (void)hipFree(a);
  for (auto* a : n->g)
    if (a) (void)hipFree(a);
  if (n->partial) (void)hipFree(n->partial);
  if (n->G) (void)hipFree(n->G);
  delete n;
  return DWF_OK;
}

int dwf_ptc_create(dwf_fopr op, int nlayers, const int* channels, dwf_ptc* out) {
  CTX();
  if (!op || !channels || !out || nlayers < 1 || nlayers > 8)
    return fail(DWF_ERR_INVALID_ARG, "invalid argument");
  if (op->prec != P32) return fail(DWF_ERR_UNSUPPORTED, "the network runs on an FP32 operator");
  if (channels[0] != 1 || channels[nlayers] != 1)
    return fail(DWF_ERR_INVALID_ARG, "first and last channel counts must be 1");
  auto* n = new dwf_ptc_s;
  n->op = op;
  int cmax = 1;
  for (int l = 0; l <= nlayers; ++l) {
    if (channels[l] < 1 || channels[l] > 16) {
      delete n;
      return fail(DWF_ERR_INVALID_ARG, "1 <= channels <= 16");
    }
    n->ch.push_back(channels[l]);
    cmax = std::max(cmax, channels[l]);
  }
  for (int l = 0; l < nlayers; ++l) {
    n->woff.push_back(n->nw);
    n->nw += (size_t)11 * channels[l] * channels[l + 1] * 32;
  }
  const size_t fb = op->nelem * sizeof(float);
  auto bad = [&](const char* m) {
    dwf_ptc_destroy(n);
    return fail(DWF_ERR_HIP, m);
  };
  if (hipMalloc(&n->W, n->nw * sizeof(unsigned short)) != hipSuccess ||
      hipMemset(n->W, 0, n->nw * sizeof(unsigned short)) != hipSuccess)
    return bad("device allocation failed");
  for (int l = 0; l <= nlayers; ++l) {
    float* a = nullptr;
    if (hipMalloc(&a, 2 * fb * channels[l]) != hipSuccess) return bad("device allocation failed");
    n->act.push_back(a);
  }
  for (auto& g : n->g)
    if (hipMalloc(&g, 2 * fb * cmax) != hipSuccess) return bad("device allocation failed");
  n->nblk = 64;
  if (hipMalloc(&n->partial, sizeof(double) * 32 * 11 * cmax * cmax * n->nblk) != hipSuccess ||
      hipMalloc(&n->G, sizeof(double) * 32 * 11 * cmax * cmax) != hipSuccess)
    return bad("device allocation failed");
  *out = n;
  return DWF_OK;
}

size_t dwf_ptc_num_weights(dwf_ptc n) { return n ? n->nw : 0; }

int dwf_ptc_set_weights(dwf_ptc n, const float* W) {
  CTX();
  if (!n || !W) return fail(DWF_ERR_INVALID_ARG, "null argument");
  std::vector<unsigned short> h(n->nw);
  for (size_t i = 0; i < n->nw; ++i) h[i] = f32_to_bf16(W[i]);
  HIPCHK(hipMemcpy(n->W, h.data(), n->nw * sizeof(unsigned short), hipMemcpyHostToDevice));
  return DWF_OK;
}

int dwf_ptc_get_weights(dwf_ptc n, float* W) {
  CTX();
  if (!n || !W) return fail(DWF_ERR_INVALID_ARG, "null argument");
  std::vector<unsigned short> h(n->nw);
  HIPCHK(hipStreamSynchronize(g_ctx.stream));
  HIPCHK(hipMemcpy(h.data(), n->W, n->nw * sizeof(unsigned short), hipMemcpyDeviceToHost));
  for (size_t i = 0; i < n->nw; ++i) W[i] = bf16_to_f32(h[i]);
  return DWF_OK;
}

int dwf_ptc_apply(dwf_ptc n, void* out, const void* in) {
  CTX();
  if (!n || !out || !in) return fail(DWF_ERR_INVALID_ARG, "null argument");
  if (!n->op->have_config) return fail(DWF_ERR_NO_CONFIG, "set_config has not been called");
  return ptc_apply_impl(n, out, in);
}

int dwf_ptc_loss_grad(dwf_ptc n, int nbatch, const void* const* b, const void* const* x,
                      double* loss, double* grad) {
  CTX();
  if (!n || nbatch < 1 || !b || !x || !loss || !grad)
    return fail(DWF_ERR_INVALID_ARG, "invalid argument");
  if (!n->op->have_config) return fail(DWF_ERR_NO_CONFIG, "set_config has not been called");
  dwf_fopr op = n->op;
  hipStream_t st = g_ctx.stream;
  const int L = (int)n->ch.size() - 1;
  std::fill(grad, grad + n->nw, 0.0);
  *loss = 0;
  std::vector<double> hG;
  for (int k = 0; k < nbatch; ++k) {
    RC(ptc_forward(n, b[k]));
    unsigned nb = 0;
    HIPCHK(launch_ptc_residual(n->g[0], n->act[L], (const float*)x[k], op->nelem,
                               g_ctx.red.partial, &nb, st));
    HIPCHK(launch_finish_partials(1, g_ctx.red.partial, nb, g_ctx.red.result, st));
    double l2;
    RC(fetch(1, &l2));
    *loss += l2;
    int cur = 0;
    for (int l = L - 1; l >= 0; --l) {
      const int cin = n->ch[l], cout = n->ch[l + 1];
      // block count: about 1024 lattice points per block, at most nblk
      const size_t pts = (size_t)2 * op->Vh * op->Ls;
      const int nbk = (int)std::max<size_t>(1, std::min<size_t>(n->nblk, pts / 1024));
      HIPCHK(launch_ptc_wgrad(n->G, n->partial, nbk, n->g[cur], n->act[l], cin, cout,
                              (const float*)op->links, op->nbr, op->Vh, op->Ls, st));
      const size_t ng = (size_t)11 * cin * cout * 32;
      hG.resize(ng);
      HIPCHK(hipMemcpyAsync(hG.data(), n->G, ng * sizeof(double), hipMemcpyDeviceToHost, st));
      if (l > 0)
        HIPCHK(launch_ptc_layer(true, n->g[1 - cur], n->g[cur], cout, cin, n->W + n->woff[l], cin,
                                cout, (const float*)op->links, op->nbr, op->Vh, op->Ls, st));
      HIPCHK(hipStreamSynchronize(st));
      // [q][b][a][32] (block y = q cin + b, z = a) -> [q][a][b][32]; with
      // g = dL/dRe out + i dL/dIm out, dL/dRe W + i dL/dIm W = G = sum g conj(T in)
      double* gl = grad + n->woff[l];
      for (int q = 0; q < 11; ++q)
        for (int bb = 0; bb < cin; ++bb)
          for (int a = 0; a < cout; ++a)
            for (int e = 0; e < 16; ++e) {
              const double* src = &hG[(((size_t)q * cin + bb) * cout + a) * 32];
              double* dst = gl + (((size_t)q * cout + a) * cin + bb) * 32;
              dst[2 * e] += src[2 * e];
              dst[2 * e + 1] += src[2 * e + 1];
            }
      cur = 1 - cur;
    }
  }
  return DWF_OK;
}

int dwf_solver_create_fcg_ptc(dwf_fopr op, dwf_ptc prec, int max_iter, double stop_cond,
                              dwf_solver* out) {
  CTX();
  if (!op || !prec || !out || max_iter < 0) return fail(DWF_ERR_INVALID_ARG, "invalid argument");
  if (op->nelem != prec->op->nelem) return fail(DWF_ERR_INVALID_ARG, "operators differ in size");
  dwf_solver_s* s = new dwf_solver_s;
  s->kind = 2;
  s->op = op;
  s->niter = max_iter;
  s->stop = stop_cond;
  s->guess = DWF_GUESS_ZERO;
  s->ptc = prec;
  *out = s;
  return DWF_OK;
}

int dwf_solver_create_richardson_ptc(dwf_fopr op_d, dwf_fopr op_f, dwf_ptc prec, int max_iter,
                                     double stop_cond, int max_iter_inner,
                                     double stop_cond_inner, dwf_solver* out) {
  CTX();
  if (!op_d || !op_f || !prec || !out) return fail(DWF_ERR_INVALID_ARG, "null argument");
  if (op_d->nelem != op_f->nelem || op_f->nelem != prec->op->nelem)
    return fail(DWF_ERR_INVALID_ARG, "operators differ in size");
  RC(dwf_solver_create_richardson(op_d, op_f, max_iter, stop_cond, max_iter_inner,
                                  stop_cond_inner, out));
  (*out)->inner->kind = 2;
  (*out)->inner->ptc = prec;
  return DWF_OK;
}

}  // extern "C"

// ---- HMC gauge part -----------------------------------------------------------
static size_t lat_volume(const int dims[4]) {
  return (size_t)dims[0] * dims[1] * dims[2] * dims[3];
}

int dwf_hmc_set_iP(const int dims[4], double* iP, const double* gauss_host, double* H_kin) {
  CTX();
  if (!dims || !iP || !gauss_host || !H_kin) return fail(DWF_ERR_INVALID_ARG, "null argument");
  const size_t nl = 4 * lat_volume(dims);
  double* h = nullptr;
  HIPCHK(hipMalloc(&h, nl * 8 * sizeof(double)));
  int rc = DWF_OK;
  unsigned nb = 0;
  if (hipMemcpyAsync(h, gauss_host, nl * 8 * sizeof(double), hipMemcpyHostToDevice,
                     g_ctx.stream) != hipSuccess ||
      launch_set_iP(iP, h, nl, g_ctx.red.partial, &nb, g_ctx.stream) != hipSuccess)
    rc = fail(DWF_ERR_HIP, "set_iP kernel failed");
  if (rc == DWF_OK) {
    HIPCHK(launch_finish_partials(1, g_ctx.red.partial, nb, g_ctx.red.result, g_ctx.stream));
    double n2;
    rc = fetch(1, &n2);
    if (rc == DWF_OK) *H_kin = 0.5 * n2;
  }
  (void)hipStreamSynchronize(g_ctx.stream);
  (void)hipFree(h);
  return rc;
}

int dwf_hmc_exp_update(const int dims[4], double* U, double alpha, const double* iP, int nprec) {
  CTX();
  if (!dims || !U || !iP || nprec < 1) return fail(DWF_ERR_INVALID_ARG, "invalid argument");
  HIPCHK(launch_exp_update(U, alpha, iP, U, 4 * lat_volume(dims), nprec, g_ctx.stream));
  return DWF_OK;
}

int dwf_hmc_rect_action(const int dims[4], const double* U, double beta, double c_plaq,
                        double c_rect, double* H, double* plaq) {
  CTX();
  if (!dims || !U || !H) return fail(DWF_ERR_INVALID_ARG, "null argument");
  unsigned nb = 0;
  HIPCHK(launch_rect_action(U, dims, g_ctx.red.partial, &nb, g_ctx.stream));
  HIPCHK(launch_finish_partials(2, g_ctx.red.partial, nb, g_ctx.red.result, g_ctx.stream));
  double v[2];
  RC(fetch(2, v));
  const double V = (double)lat_volume(dims);
  const double Ndim2 = 6.0, Nc = 3.0;
  // Action_G_Rectangle::calcH (action_G_Rectangle.cpp:183-186)
  *H = beta * (c_plaq * Ndim2 * V - c_plaq * v[0] / Nc + c_rect * Ndim2 * V * 2 -
               c_rect * v[1] / Nc);
  if (plaq) *plaq = v[0] / Nc / V / Ndim2;
  return DWF_OK;
}

int dwf_hmc_rect_force(const int dims[4], double* F, const double* U, double beta, double c_plaq,
                       double c_rect) {
  CTX();
  if (!dims || !F || !U) return fail(DWF_ERR_INVALID_ARG, "null argument");
  HIPCHK(launch_rect_force(F, U, dims, beta, c_plaq, c_rect, g_ctx.stream));
  return DWF_OK;
}

int dwf_device_axpy(double* y, double a, const double* x, size_t n) {
  CTX();
  if (!y || !x) return fail(DWF_ERR_INVALID_ARG, "null argument");
  HIPCHK(launch_axpy(P64, y, a, x, n, g_ctx.stream));
  return DWF_OK;
}

int dwf_device_copy(double* y, const double* x, size_t n) {
  CTX();
  if (!y || !x) return fail(DWF_ERR_INVALID_ARG, "null argument");
  HIPCHK(hipMemcpyAsync(y, x, n * sizeof(double), hipMemcpyDeviceToDevice, g_ctx.stream));
  return DWF_OK;
}

int dwf_device_norm2(const double* x, size_t n, double* result) {
  CTX();
  if (!x || !result) return fail(DWF_ERR_INVALID_ARG, "null argument");
  HIPCHK(launch_norm2(P64, x, n, g_ctx.red, 0, g_ctx.stream));
  return fetch(1, result);
}

int dwf_plaquette_device(const int dims[4], const double* U, double* result) {
  CTX();
  if (!dims || !U || !result) return fail(DWF_ERR_INVALID_ARG, "null argument");
  HIPCHK(launch_plaquette(U, dims[0], dims[1], dims[2], dims[3], g_ctx.red, 0, g_ctx.stream));
  double sum = 0;
  RC(fetch(1, &sum));
  *result = sum / (3.0 * 6.0 * (double)lat_volume(dims));
  return DWF_OK;
}

int dwf_plaquette(const int dims[4], const double* U_host, double* result) {
  CTX();
  if (!dims || !U_host || !result) return fail(DWF_ERR_INVALID_ARG, "null argument");
  const size_t V = (size_t)dims[0] * dims[1] * dims[2] * dims[3];
  double* d = nullptr;
  HIPCHK(hipMalloc(&d, V * 4 * 18 * sizeof(double)));
  int rc = DWF_OK;
  if (hipMemcpyAsync(d, U_host, V * 4 * 18 * sizeof(double), hipMemcpyHostToDevice,
                     g_ctx.stream) != hipSuccess ||
      launch_plaquette(d, dims[0], dims[1], dims[2], dims[3], g_ctx.red, 0, g_ctx.stream) !=
          hipSuccess)
    rc = fail(DWF_ERR_HIP, "plaquette kernel failed");
  double sum = 0;
  if (rc == DWF_OK) rc = fetch(1, &sum);
  (void)hipFree(d);
  if (rc == DWF_OK) *result = sum / (3.0 * 6.0 * (double)V);
  return rc;
}

}  // extern "C"
