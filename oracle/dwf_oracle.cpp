/*
 * dwf_oracle.cpp -- CPU oracle (test infrastructure only; see dwf_oracle.h).
 *
 * Scalar restatement of the reference algorithm.  The operator follows the
 * corelib even-odd Moebius domain-wall operator step by step, in the same
 * order of operations, so that its rounding is close to the reference's:
 *   src/lib/Fopr/afopr_Domainwall_eo-tmpl.h   (5th-dimension structure)
 *   src/lib/Fopr/Imp/fopr_Wilson_eo_impl.cpp  (4D even-odd Wilson hop)
 *   src/lib/Fopr/Imp/fopr_Wilson_impl_common-inc.h (Dirac-basis projectors)
 * Solvers follow src/lib_alt/Solver/asolver_CG-tmpl.h,
 * asolver_Richardson-tmpl.h, aprecond_Mixedprec-tmpl.h and
 * src/lib/Solver/ashiftsolver_CG-tmpl.h; smearing follows
 * src/lib_alt_QXS/Smear/asmear_APE-tmpl.h, aprojection_Stout_SU3-tmpl.h and
 * src/lib_alt_QXS/Measurements/Gauge/astaple_lex-tmpl.h.
 */
#include "dwf_oracle.h"

#include <cmath>
#include <complex>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>
#include <omp.h>

namespace {

typedef std::complex<double> cd;

struct Geom {
  int N[4];
  int Nx2, V, Vh;
};

inline int lex(const Geom& g, const int x[4]) {
  return x[0] + g.N[0] * (x[1] + g.N[1] * (x[2] + g.N[2] * x[3]));
}

// ---------------------------------------------------------------------------
// Operator state, templated on the arithmetic type real_t (double or float).
// ---------------------------------------------------------------------------
template <typename R>
struct Dwf {
  Geom g;
  int Ls;
  R mq, M0, alpha;
  std::vector<R> b, c;
  std::vector<R> dp, dm, e, f;
  R gg;
  int bc[4];
  double kappa;            // hopping parameter of the kernel, 1/(8-2 M0)
  std::vector<R> U;        // lexical gauge field (18, V, 4) in real_t
  bool have_U = false;
  // neighbour tables for the even-odd hop: for output parity ieo and half
  // site s2: nbr[ieo][mu][fwd/bwd][s2] = half-site index of the neighbour,
  // lnk = lexical index of the link site, bcf = boundary factor.
  std::vector<int> nbr[2][4][2];
  std::vector<int> lnk[2][4][2];
  std::vector<signed char> bcf[2][4][2];
  // work 4D slices
  std::vector<R> w4, v4, y4, t4;
  std::vector<R> v1, v2, w1;
};

template <typename R>
void set_precond(Dwf<R>& d) {
  // afopr_Domainwall_eo-tmpl.h:269-306 (set_precond_parameters)
  const int Ns = d.Ls;
  d.dp.assign(Ns, 0);
  d.dm.assign(Ns, 0);
  d.e.assign(Ns - 1, 0);
  d.f.assign(Ns - 1, 0);
  for (int is = 0; is < Ns; ++is) {
    d.dp[is] = d.alpha * (1.0 + d.b[is] * (4.0 - d.M0));
    d.dm[is] = d.alpha * (1.0 - d.c[is] * (4.0 - d.M0));
  }
  d.e[0] = d.mq * d.dm[Ns - 1] / d.dp[0];
  d.f[0] = d.mq * d.dm[0] / d.alpha;
  for (int is = 1; is < Ns - 1; ++is) {
    d.e[is] = d.e[is - 1] * d.dm[is - 1] / d.dp[is];
    d.f[is] = d.f[is - 1] * d.dm[is] / d.dp[is - 1];
  }
  d.gg = d.e[Ns - 2] * d.dm[Ns - 2];
}

template <typename R>
void build_tables(Dwf<R>& d) {
  const Geom& g = d.g;
  for (int ieo = 0; ieo < 2; ++ieo)
    for (int mu = 0; mu < 4; ++mu)
      for (int fb = 0; fb < 2; ++fb) {
        d.nbr[ieo][mu][fb].resize(g.Vh);
        d.lnk[ieo][mu][fb].resize(g.Vh);
        d.bcf[ieo][mu][fb].resize(g.Vh);
      }
  for (int ieo = 0; ieo < 2; ++ieo) {
    for (int s2 = 0; s2 < g.Vh; ++s2) {
      int x2 = s2 % g.Nx2;
      int r = s2 / g.Nx2;
      int y = r % g.N[1];
      r /= g.N[1];
      int z = r % g.N[2];
      int t = r / g.N[2];
      int keo = (y + z + t + ieo) % 2;  // index_eo.cpp convertField
      int x[4] = {2 * x2 + keo, y, z, t};
      for (int mu = 0; mu < 4; ++mu) {
        // forward neighbour x + mu, link U_mu(x)
        int xf[4] = {x[0], x[1], x[2], x[3]};
        int bf = 1;
        xf[mu] += 1;
        if (xf[mu] == g.N[mu]) { xf[mu] = 0; bf = d.bc[mu]; }
        // backward neighbour x - mu, link U_mu(x - mu)^dagger
        int xb[4] = {x[0], x[1], x[2], x[3]};
        int bb = 1;
        xb[mu] -= 1;
        if (xb[mu] < 0) { xb[mu] = g.N[mu] - 1; bb = d.bc[mu]; }
        d.nbr[ieo][mu][0][s2] = xf[0] / 2 + g.Nx2 * (xf[1] + g.N[1] * (xf[2] + g.N[2] * xf[3]));
        d.nbr[ieo][mu][1][s2] = xb[0] / 2 + g.Nx2 * (xb[1] + g.N[1] * (xb[2] + g.N[2] * xb[3]));
        d.lnk[ieo][mu][0][s2] = lex(g, x);
        d.lnk[ieo][mu][1][s2] = lex(g, xb);
        d.bcf[ieo][mu][0][s2] = (signed char)bf;
        d.bcf[ieo][mu][1][s2] = (signed char)bb;
      }
    }
  }
}

// ---- 4D slice helpers (Field(24, Vh, 1)) -----------------------------------
template <typename R>
inline void gm5_4d(const Geom& g, R* v, const R* w) {
  // gamma_5 in the Dirac basis swaps upper and lower spinor components
  // (fopr_Wilson_impl_common-inc.h:14-22).  v may alias w.
#pragma omp parallel for schedule(static)
  for (int s = 0; s < g.Vh; ++s) {
    R* vp = v + 24 * s;
    const R* wp = w + 24 * s;
    R tmp[24];
    for (int i = 0; i < 12; ++i) { tmp[i] = wp[i + 12]; tmp[i + 12] = wp[i]; }
    for (int i = 0; i < 24; ++i) vp[i] = tmp[i];
  }
}
template <typename R>
inline void copy4(const Geom& g, R* v, const R* w) {
  std::memcpy(v, w, sizeof(R) * 24 * (size_t)g.Vh);
}
template <typename R>
inline void scal4(const Geom& g, R* v, R a) {
#pragma omp parallel for schedule(static)
  for (long i = 0; i < 24L * g.Vh; ++i) v[i] *= a;
}
template <typename R>
inline void axpy4(const Geom& g, R* y, R a, const R* x) {
#pragma omp parallel for schedule(static)
  for (long i = 0; i < 24L * g.Vh; ++i) y[i] += a * x[i];
}
template <typename R>
inline void set4(const Geom& g, R* v, R a) {
#pragma omp parallel for schedule(static)
  for (long i = 0; i < 24L * g.Vh; ++i) v[i] = a;
}
template <typename R>
inline R* sl(R* v, const Geom& g, int s) { return v + 24L * g.Vh * s; }
template <typename R>
inline const R* sl(const R* v, const Geom& g, int s) { return v + 24L * g.Vh * s; }

// ---- SU(3) times colour vector (fopr_Wilson_impl_SU3-inc.h:30-60) --------
template <typename R>
inline void uv(R& re, R& im, const R* u /*row*/, const R* w) {
  re = u[0] * w[0] - u[1] * w[1] + u[2] * w[2] - u[3] * w[3] + u[4] * w[4] - u[5] * w[5];
  im = u[0] * w[1] + u[1] * w[0] + u[2] * w[3] + u[3] * w[2] + u[4] * w[5] + u[5] * w[4];
}
template <typename R>
inline void udagv(R& re, R& im, const R* u /*column*/, const R* w) {
  re = u[0] * w[0] + u[1] * w[1] + u[6] * w[2] + u[7] * w[3] + u[12] * w[4] + u[13] * w[5];
  im = u[0] * w[1] - u[1] * w[0] + u[6] * w[3] - u[7] * w[2] + u[12] * w[5] - u[13] * w[4];
}

// ---- the 4D even-odd Wilson hop: v = -kappa sum_mu [...] ------------------
// Dirac-basis spin projection.  For each direction the two-component
// projection h1, h2 of (1 -/+ gamma_mu) psi, the link multiply and the
// reconstruction of the four spin components follow the reference's
// set_sp2_*/set_sp4_* pairs (fopr_Wilson_impl_common-inc.h:24-234).  The
// forward term uses U_mu(x), the backward term U_mu(x-mu)^dagger; the boundary
// factor multiplies the half spinor (forward) or the product (backward), as in
// fopr_Wilson_eo_impl.cpp mult_xp/mult_xm (:714-890) and mult_tp/tm_dirac
// (:1256-1430).
template <typename R>
void wilson_meo(const Dwf<R>& d, R* vout, const R* win, int ieo) {
  const Geom& g = d.g;
  const R* U = d.U.data();
  const long V = g.V;
  const R mk = (R)(-d.kappa);
#pragma omp parallel for schedule(static)
  for (int s2 = 0; s2 < g.Vh; ++s2) {
    R v[24];
    for (int i = 0; i < 24; ++i) v[i] = 0;
    for (int mu = 0; mu < 4; ++mu) {
      for (int fb = 0; fb < 2; ++fb) {
        const int nb = d.nbr[ieo][mu][fb][s2];
        const R* w = win + 24L * nb;
        const R* u = U + 18L * (d.lnk[ieo][mu][fb][s2] + V * mu);
        const R bcv = (R)d.bcf[ieo][mu][fb][s2];
        R h1[6], h2[6];
        // projection
        for (int ic = 0; ic < 3; ++ic) {
          const int r = 2 * ic, i = 2 * ic + 1;
          const R* w0 = w;  const R* w1 = w + 6;
          const R* w2 = w + 12; const R* w3 = w + 18;
          if (mu == 0 && fb == 0) {        // (1 - g1): h1 = w0 + i w3, h2 = w1 + i w2
            h1[r] = w0[r] - w3[i]; h1[i] = w0[i] + w3[r];
            h2[r] = w1[r] - w2[i]; h2[i] = w1[i] + w2[r];
          } else if (mu == 0) {             // (1 + g1): h1 = w0 - i w3, h2 = w1 - i w2
            h1[r] = w0[r] + w3[i]; h1[i] = w0[i] - w3[r];
            h2[r] = w1[r] + w2[i]; h2[i] = w1[i] - w2[r];
          } else if (mu == 1 && fb == 0) { // (1 - g2): h1 = w0 + w3, h2 = w1 - w2
            h1[r] = w0[r] + w3[r]; h1[i] = w0[i] + w3[i];
            h2[r] = w1[r] - w2[r]; h2[i] = w1[i] - w2[i];
          } else if (mu == 1) {             // (1 + g2): h1 = w0 - w3, h2 = w1 + w2
            h1[r] = w0[r] - w3[r]; h1[i] = w0[i] - w3[i];
            h2[r] = w1[r] + w2[r]; h2[i] = w1[i] + w2[i];
          } else if (mu == 2 && fb == 0) { // (1 - g3): h1 = w0 + i w2, h2 = w1 - i w3
            h1[r] = w0[r] - w2[i]; h1[i] = w0[i] + w2[r];
            h2[r] = w1[r] + w3[i]; h2[i] = w1[i] - w3[r];
          } else if (mu == 2) {             // (1 + g3): h1 = w0 - i w2, h2 = w1 + i w3
            h1[r] = w0[r] + w2[i]; h1[i] = w0[i] - w2[r];
            h2[r] = w1[r] - w3[i]; h2[i] = w1[i] + w3[r];
          } else if (fb == 0) {             // (1 - g4): h1 = 2 w2, h2 = 2 w3
            h1[r] = 2 * w2[r]; h1[i] = 2 * w2[i];
            h2[r] = 2 * w3[r]; h2[i] = 2 * w3[i];
          } else {                          // (1 + g4): h1 = 2 w0, h2 = 2 w1
            h1[r] = 2 * w0[r]; h1[i] = 2 * w0[i];
            h2[r] = 2 * w1[r]; h2[i] = 2 * w1[i];
          }
        }
        if (fb == 0 && bcv != (R)1) {
          for (int k = 0; k < 6; ++k) { h1[k] *= bcv; h2[k] *= bcv; }
        }
        for (int ic = 0; ic < 3; ++ic) {
          R a1r, a1i, a2r, a2i;
          if (fb == 0) {
            uv(a1r, a1i, u + 6 * ic, h1);
            uv(a2r, a2i, u + 6 * ic, h2);
          } else {
            udagv(a1r, a1i, u + 2 * ic, h1);
            udagv(a2r, a2i, u + 2 * ic, h2);
            if (bcv != (R)1) { a1r *= bcv; a1i *= bcv; a2r *= bcv; a2i *= bcv; }
          }
          R* o0 = v + 2 * ic; R* o1 = v + 6 + 2 * ic;
          R* o2 = v + 12 + 2 * ic; R* o3 = v + 18 + 2 * ic;
          if (mu == 0 && fb == 0) {        // o2 += -i a2, o3 += -i a1
            o0[0] += a1r; o0[1] += a1i; o1[0] += a2r; o1[1] += a2i;
            o2[0] += a2i; o2[1] -= a2r; o3[0] += a1i; o3[1] -= a1r;
          } else if (mu == 0) {             // o2 += i a2, o3 += i a1
            o0[0] += a1r; o0[1] += a1i; o1[0] += a2r; o1[1] += a2i;
            o2[0] -= a2i; o2[1] += a2r; o3[0] -= a1i; o3[1] += a1r;
          } else if (mu == 1 && fb == 0) { // o2 -= a2, o3 += a1
            o0[0] += a1r; o0[1] += a1i; o1[0] += a2r; o1[1] += a2i;
            o2[0] -= a2r; o2[1] -= a2i; o3[0] += a1r; o3[1] += a1i;
          } else if (mu == 1) {             // o2 += a2, o3 -= a1
            o0[0] += a1r; o0[1] += a1i; o1[0] += a2r; o1[1] += a2i;
            o2[0] += a2r; o2[1] += a2i; o3[0] -= a1r; o3[1] -= a1i;
          } else if (mu == 2 && fb == 0) { // o2 += -i a1, o3 += i a2
            o0[0] += a1r; o0[1] += a1i; o1[0] += a2r; o1[1] += a2i;
            o2[0] += a1i; o2[1] -= a1r; o3[0] -= a2i; o3[1] += a2r;
          } else if (mu == 2) {             // o2 += i a1, o3 += -i a2
            o0[0] += a1r; o0[1] += a1i; o1[0] += a2r; o1[1] += a2i;
            o2[0] -= a1i; o2[1] += a1r; o3[0] += a2i; o3[1] -= a2r;
          } else if (fb == 0) {             // lower components only
            o2[0] += a1r; o2[1] += a1i; o3[0] += a2r; o3[1] += a2i;
          } else {                          // upper components only
            o0[0] += a1r; o0[1] += a1i; o1[0] += a2r; o1[1] += a2i;
          }
        }
      }
    }
    R* vo = vout + 24L * s2;
    for (int k = 0; k < 24; ++k) vo[k] = v[k] * mk;
  }
}

// ---- 5D even-odd operator pieces (afopr_Domainwall_eo-tmpl.h) --------------

// multiply by (fac1 + fac2 gamma_5) when fac2 != 0 (the alpha end factors)
template <typename R>
inline void alpha_end(Dwf<R>& d, R* w4, R fac1, R fac2) {
  if (fac2 != (R)0) {
    gm5_4d(d.g, d.t4.data(), w4);
    scal4(d.g, w4, fac1);
    axpy4(d.g, w4, fac2, d.t4.data());
  }
}
template <typename R>
inline void alpha_site(Dwf<R>& d, R* w4, int is) {
  if (is == 0) {
    alpha_end(d, w4, (R)(0.5 * (1.0 + d.alpha)), (R)(0.5 * (-1.0 + d.alpha)));
  } else if (is == d.Ls - 1) {
    alpha_end(d, w4, (R)(0.5 * (1.0 + d.alpha)), (R)(0.5 * (1.0 - d.alpha)));
  } else {
    scal4(d.g, w4, d.alpha);
  }
}

// m_v4 = Fup (1 - g5) w(s+1) + Fdn (1 + g5) w(s-1), walls -mq/2
template <typename R>
void m5_neighbours(Dwf<R>& d, R* v4, const R* w, int is) {
  const Geom& g = d.g;
  set4(g, v4, (R)0);
  int is_up = (is + 1) % d.Ls;
  R Fup = (R)(0.5 * d.alpha);
  if (is == d.Ls - 1) Fup = (R)(-0.5 * d.mq);
  copy4(g, d.y4.data(), sl(w, g, is_up));
  gm5_4d(g, d.t4.data(), d.y4.data());
  axpy4(g, d.y4.data(), (R)-1, d.t4.data());
  axpy4(g, v4, Fup, d.y4.data());
  int is_dn = (is - 1 + d.Ls) % d.Ls;
  R Fdn = (R)(0.5 * d.alpha);
  if (is == 0) Fdn = (R)(-0.5 * d.mq);
  copy4(g, d.y4.data(), sl(w, g, is_dn));
  gm5_4d(g, d.t4.data(), d.y4.data());
  axpy4(g, d.y4.data(), (R)1, d.t4.data());
  axpy4(g, v4, Fdn, d.y4.data());
}

// D_eo: afopr_Domainwall_eo-tmpl.h:765-828
template <typename R>
void D_eo(Dwf<R>& d, R* v, const R* w, int ieo) {
  const Geom& g = d.g;
  for (int is = 0; is < d.Ls; ++is) {
    m5_neighbours(d, d.v4.data(), w, is);
    copy4(g, d.w4.data(), sl(w, g, is));
    alpha_site(d, d.w4.data(), is);
    scal4(g, d.w4.data(), d.b[is]);
    axpy4(g, d.w4.data(), d.c[is], d.v4.data());
    wilson_meo(d, d.v4.data(), d.w4.data(), ieo);
    scal4(g, d.v4.data(), (R)(4.0 - d.M0));
    copy4(g, sl(v, g, is), d.v4.data());
  }
}

// Ddag_eo: afopr_Domainwall_eo-tmpl.h:830-898
template <typename R>
void Ddag_eo(Dwf<R>& d, R* v, const R* w, int ieo) {
  const Geom& g = d.g;
  const long n = 24L * g.Vh * d.Ls;
  std::memset(v, 0, sizeof(R) * n);
  R* w4 = d.w4.data(); R* v4 = d.v4.data(); R* y4 = d.y4.data();
  for (int is = 0; is < d.Ls; ++is) {
    copy4(g, w4, sl(w, g, is));
    gm5_4d(g, v4, w4);
    wilson_meo(d, y4, v4, ieo);
    gm5_4d(g, v4, y4);
    scal4(g, v4, (R)(4.0 - d.M0));
    copy4(g, w4, v4);
    alpha_site(d, v4, is);
    axpy4(g, sl(v, g, is), d.b[is], v4);
    scal4(g, w4, d.c[is]);
    int is_up = (is + 1) % d.Ls;
    R Fup = (R)(0.5 * d.alpha);
    if (is_up == 0) Fup = (R)(-0.5 * d.mq);
    gm5_4d(g, y4, w4);
    axpy4(g, y4, (R)-1, w4);
    axpy4(g, sl(v, g, is_up), (R)(-Fup), y4);
    int is_dn = (is - 1 + d.Ls) % d.Ls;
    R Fdn = (R)(0.5 * d.alpha);
    if (is_dn == d.Ls - 1) Fdn = (R)(-0.5 * d.mq);
    gm5_4d(g, y4, w4);
    axpy4(g, y4, (R)1, w4);
    axpy4(g, sl(v, g, is_dn), Fdn, y4);
  }
}

// D_ee: afopr_Domainwall_eo-tmpl.h:900-959
template <typename R>
void D_ee(Dwf<R>& d, R* v, const R* w) {
  const Geom& g = d.g;
  for (int is = 0; is < d.Ls; ++is) {
    m5_neighbours(d, d.v4.data(), w, is);
    copy4(g, d.w4.data(), sl(w, g, is));
    alpha_site(d, d.w4.data(), is);
    R F1 = (R)(d.b[is] * (4.0 - d.M0) + 1.0);
    scal4(g, d.w4.data(), F1);
    R F2 = (R)(d.c[is] * (4.0 - d.M0) - 1.0);
    axpy4(g, d.w4.data(), F2, d.v4.data());
    copy4(g, sl(v, g, is), d.w4.data());
  }
}

// Ddag_ee: afopr_Domainwall_eo-tmpl.h:962-1019
template <typename R>
void Ddag_ee(Dwf<R>& d, R* v, const R* w) {
  const Geom& g = d.g;
  std::memset(v, 0, sizeof(R) * 24L * g.Vh * d.Ls);
  R* w4 = d.w4.data(); R* y4 = d.y4.data(); R* t4 = d.t4.data();
  for (int is = 0; is < d.Ls; ++is) {
    copy4(g, w4, sl(w, g, is));
    R F1 = (R)(d.b[is] * (4.0 - d.M0) + 1.0);
    scal4(g, w4, F1);
    alpha_site(d, w4, is);
    axpy4(g, sl(v, g, is), (R)1, w4);
    copy4(g, y4, sl(w, g, is));
    R F2 = (R)(d.c[is] * (4.0 - d.M0) - 1.0);
    scal4(g, y4, F2);
    int is_up = (is + 1) % d.Ls;
    R Fup = (R)(0.5 * d.alpha);
    if (is_up == 0) Fup = (R)(-0.5 * d.mq);
    gm5_4d(g, t4, y4);
    axpy4(g, t4, (R)-1, y4);
    axpy4(g, sl(v, g, is_up), (R)(-Fup), t4);
    int is_dn = (is - 1 + d.Ls) % d.Ls;
    R Fdn = (R)(0.5 * d.alpha);
    if (is_dn == d.Ls - 1) Fdn = (R)(-0.5 * d.mq);
    gm5_4d(g, t4, y4);
    axpy4(g, t4, (R)1, y4);
    axpy4(g, sl(v, g, is_dn), Fdn, t4);
  }
}

// L_inv: afopr_Domainwall_eo-tmpl.h:1022-1065
template <typename R>
void L_inv(Dwf<R>& d, R* v, const R* w) {
  const Geom& g = d.g;
  const int Ns = d.Ls;
  R* v4 = d.v4.data(); R* w4 = d.w4.data(); R* y4 = d.y4.data();
  copy4(g, sl(v, g, 0), sl(w, g, 0));
  copy4(g, y4, sl(w, g, 0));
  scal4(g, y4, d.e[0]);
  for (int is = 1; is < Ns - 1; ++is) {
    copy4(g, sl(v, g, is), sl(w, g, is));
    copy4(g, v4, sl(v, g, is - 1));
    gm5_4d(g, w4, v4);
    axpy4(g, v4, (R)1, w4);
    scal4(g, v4, (R)((R)0.5 * d.dm[is] / d.dp[is - 1]));
    axpy4(g, sl(v, g, is), (R)1, v4);
    copy4(g, w4, sl(v, g, is));
    axpy4(g, y4, d.e[is], w4);
  }
  int is = Ns - 1;
  copy4(g, sl(v, g, is), sl(w, g, is));
  copy4(g, v4, sl(v, g, is - 1));
  gm5_4d(g, w4, v4);
  axpy4(g, v4, (R)1, w4);
  scal4(g, v4, (R)((R)0.5 * d.dm[is] / d.dp[is - 1]));
  axpy4(g, sl(v, g, is), (R)1, v4);
  gm5_4d(g, w4, y4);
  axpy4(g, y4, (R)-1, w4);
  scal4(g, y4, (R)-0.5);
  axpy4(g, sl(v, g, is), (R)1, y4);
}

// U_inv: afopr_Domainwall_eo-tmpl.h:1068-1117
template <typename R>
void U_inv(Dwf<R>& d, R* v, const R* w) {
  const Geom& g = d.g;
  const int Ns = d.Ls;
  R* v4 = d.v4.data(); R* w4 = d.w4.data(); R* y4 = d.y4.data(); R* t4 = d.t4.data();
  int is = Ns - 1;
  copy4(g, y4, sl(w, g, is));
  alpha_end(d, y4, (R)(0.5 * (1.0 + d.alpha)), (R)(0.5 * (-1.0 + d.alpha)));
  scal4(g, y4, (R)((R)1 / (d.dp[is] + d.gg)));
  copy4(g, sl(v, g, is), y4);
  gm5_4d(g, w4, y4);
  axpy4(g, y4, (R)1, w4);
  scal4(g, y4, (R)0.5);
  for (int js = Ns - 2; js >= 0; --js) {
    copy4(g, v4, sl(w, g, js));
    copy4(g, w4, sl(v, g, js + 1));
    gm5_4d(g, t4, w4);
    axpy4(g, w4, (R)-1, t4);
    axpy4(g, v4, (R)((R)0.5 * d.dm[js]), w4);
    axpy4(g, v4, (R)(-d.f[js]), y4);
    scal4(g, v4, (R)((R)1 / d.dp[js]));
    if (js == 0) {
      alpha_end(d, v4, (R)(0.5 * (1.0 + d.alpha)), (R)(0.5 * (1.0 - d.alpha)));
    }
    copy4(g, sl(v, g, js), v4);
  }
}

// Udag_inv: afopr_Domainwall_eo-tmpl.h:1122-1188
template <typename R>
void Udag_inv(Dwf<R>& d, R* v, const R* w) {
  const Geom& g = d.g;
  const int Ns = d.Ls;
  R* v4 = d.v4.data(); R* w4 = d.w4.data(); R* y4 = d.y4.data(); R* t4 = d.t4.data();
  copy4(g, v4, sl(w, g, 0));
  alpha_end(d, v4, (R)(0.5 * (1.0 + d.alpha)), (R)(0.5 * (1.0 - d.alpha)));
  scal4(g, v4, (R)((R)1 / d.dp[0]));
  copy4(g, sl(v, g, 0), v4);
  copy4(g, y4, v4);
  scal4(g, y4, d.f[0]);
  for (int is = 1; is < Ns - 1; ++is) {
    copy4(g, t4, sl(w, g, is));
    copy4(g, v4, sl(v, g, is - 1));
    gm5_4d(g, w4, v4);
    axpy4(g, v4, (R)-1, w4);
    axpy4(g, t4, (R)((R)0.5 * d.dm[is - 1]), v4);
    scal4(g, t4, (R)((R)1 / d.dp[is]));
    copy4(g, sl(v, g, is), t4);
    axpy4(g, y4, d.f[is], t4);
  }
  int is = Ns - 1;
  copy4(g, t4, sl(w, g, is));
  copy4(g, v4, sl(v, g, is - 1));
  gm5_4d(g, w4, v4);
  axpy4(g, v4, (R)-1, w4);
  axpy4(g, t4, (R)((R)0.5 * d.dm[is - 1]), v4);
  gm5_4d(g, w4, y4);
  axpy4(g, y4, (R)1, w4);
  scal4(g, y4, (R)0.5);
  axpy4(g, t4, (R)-1, y4);
  scal4(g, t4, (R)((R)1 / (d.dp[is] + d.gg)));
  // multiply (alpha P+ + P-)
  {
    R fac1 = (R)(0.5 * (1.0 + d.alpha));
    R fac2 = (R)(0.5 * (-1.0 + d.alpha));
    if (fac2 != (R)0) {
      gm5_4d(g, y4, t4);
      scal4(g, t4, fac1);
      axpy4(g, t4, fac2, y4);
    }
  }
  copy4(g, sl(v, g, is), t4);
}

// Ldag_inv: afopr_Domainwall_eo-tmpl.h:1191-1220
template <typename R>
void Ldag_inv(Dwf<R>& d, R* v, const R* w) {
  const Geom& g = d.g;
  const int Ns = d.Ls;
  R* v4 = d.v4.data(); R* y4 = d.y4.data(); R* t4 = d.t4.data();
  int is = Ns - 1;
  copy4(g, sl(v, g, is), sl(w, g, is));
  copy4(g, y4, sl(w, g, is));
  gm5_4d(g, t4, y4);
  axpy4(g, y4, (R)-1, t4);
  scal4(g, y4, (R)0.5);
  for (int js = Ns - 2; js >= 0; --js) {
    copy4(g, sl(v, g, js), sl(w, g, js));
    copy4(g, v4, sl(v, g, js + 1));
    gm5_4d(g, t4, v4);
    axpy4(g, v4, (R)1, t4);
    scal4(g, v4, (R)((R)0.5 * d.dm[js + 1] / d.dp[js]));
    axpy4(g, v4, (R)(-d.e[js]), y4);
    axpy4(g, sl(v, g, js), (R)1, v4);
  }
}

template <typename R>
inline long n5(const Dwf<R>& d) { return 24L * d.g.Vh * d.Ls; }

template <typename R>
void copy5(const Dwf<R>& d, R* v, const R* w) { std::memcpy(v, w, sizeof(R) * n5(d)); }
template <typename R>
void axpy5(const Dwf<R>& d, R* y, R a, const R* x) {
  const long n = n5(d);
#pragma omp parallel for schedule(static)
  for (long i = 0; i < n; ++i) y[i] += a * x[i];
}

template <typename R>
void gm5R(Dwf<R>& d, R* v, const R* w) {
  // mult_gm5R: afopr_Domainwall_eo-tmpl.h:744-760
  for (int is = 0; is < d.Ls; ++is)
    gm5_4d(d.g, sl(v, d.g, d.Ls - 1 - is), sl(w, d.g, is));
}
template <typename R>
void gm5_5d(Dwf<R>& d, R* v, const R* w) {
  for (int is = 0; is < d.Ls; ++is) gm5_4d(d.g, sl(v, d.g, is), sl(w, d.g, is));
}

// D: afopr_Domainwall_eo-tmpl.h:611-627
template <typename R>
void Dop(Dwf<R>& d, R* v, const R* w) {
  R* v1 = d.v1.data(); R* v2 = d.v2.data();
  D_eo(d, v1, w, 1);
  L_inv(d, v2, v1);
  U_inv(d, v1, v2);
  D_eo(d, v2, v1, 0);
  L_inv(d, v1, v2);
  U_inv(d, v2, v1);
  copy5(d, v, w);
  axpy5(d, v, (R)-1, v2);
}
// Ddag: afopr_Domainwall_eo-tmpl.h:652-669
template <typename R>
void Ddagop(Dwf<R>& d, R* v, const R* w) {
  R* v1 = d.v1.data(); R* v2 = d.v2.data();
  Udag_inv(d, v1, w);
  Ldag_inv(d, v2, v1);
  Ddag_eo(d, v1, v2, 1);
  Udag_inv(d, v2, v1);
  Ldag_inv(d, v1, v2);
  Ddag_eo(d, v2, v1, 0);
  copy5(d, v, w);
  axpy5(d, v, (R)-1, v2);
}
// DdagD: afopr_Domainwall_eo-tmpl.h:584-608 (D then the daggered chain)
template <typename R>
void DdagDop(Dwf<R>& d, R* v, const R* w) {
  std::vector<R> tmp(n5(d));
  Dop(d, tmp.data(), w);
  Ddagop(d, v, tmp.data());
}
template <typename R>
void Hop(Dwf<R>& d, R* v, const R* w) {
  std::vector<R> tmp(n5(d));
  Dop(d, tmp.data(), w);
  gm5R(d, v, tmp.data());
}
template <typename R>
void Hdagop(Dwf<R>& d, R* v, const R* w) {
  std::vector<R> tmp(n5(d));
  gm5R(d, tmp.data(), w);
  Ddagop(d, v, tmp.data());
}

template <typename R>
int mult_mode(Dwf<R>& d, const std::string& mode, R* v, const R* w) {
  // afopr_Domainwall_eo-tmpl.h:384-431 (mult) and :475-498 (mult with mode)
  if (mode == "D") Dop(d, v, w);
  else if (mode == "Ddag") Ddagop(d, v, w);
  else if (mode == "DdagD") DdagDop(d, v, w);
  else if (mode == "DDdag") { std::vector<R> t(n5(d)); Ddagop(d, t.data(), w); Dop(d, v, t.data()); }
  else if (mode == "H") Hop(d, v, w);
  else if (mode == "Hdag") Hdagop(d, v, w);
  else if (mode == "Deo") D_eo(d, v, w, 0);
  else if (mode == "Doe") D_eo(d, v, w, 1);
  else if (mode == "Dee" || mode == "Doo") D_ee(d, v, w);
  else if (mode == "Dee_inv" || mode == "Doo_inv") { L_inv(d, d.v1.data(), w); U_inv(d, v, d.v1.data()); }
  else return -1;
  return 0;
}
template <typename R>
int mult_dag_mode(Dwf<R>& d, const std::string& mode, R* v, const R* w) {
  // afopr_Domainwall_eo-tmpl.h:435-470 and :501-525
  if (mode == "D") Ddagop(d, v, w);
  else if (mode == "Ddag") Dop(d, v, w);
  else if (mode == "DdagD") DdagDop(d, v, w);
  else if (mode == "DDdag") { std::vector<R> t(n5(d)); Ddagop(d, t.data(), w); Dop(d, v, t.data()); }
  else if (mode == "H") Hdagop(d, v, w);
  else if (mode == "Hdag") Hop(d, v, w);
  else if (mode == "Deo") Ddag_eo(d, v, w, 1);
  else if (mode == "Doe") Ddag_eo(d, v, w, 0);
  else if (mode == "Dee" || mode == "Doo") Ddag_ee(d, v, w);
  else if (mode == "Dee_inv" || mode == "Doo_inv") { Udag_inv(d, d.v1.data(), w); Ldag_inv(d, v, d.v1.data()); }
  else return -1;
  return 0;
}

// ---- BLAS on real_t vectors, reductions accumulated in double ------------
template <typename R>
double norm2v(const R* x, long n) {
  double s = 0;
#pragma omp parallel for reduction(+ : s) schedule(static)
  for (long i = 0; i < n; ++i) s += (double)x[i] * (double)x[i];
  return s;
}
template <typename R>
double dotv(const R* x, const R* y, long n) {
  double s = 0;
#pragma omp parallel for reduction(+ : s) schedule(static)
  for (long i = 0; i < n; ++i) s += (double)x[i] * (double)y[i];
  return s;
}
template <typename R>
void axpyv(R* y, R a, const R* x, long n) {
#pragma omp parallel for schedule(static)
  for (long i = 0; i < n; ++i) y[i] += a * x[i];
}
template <typename R>
void aypxv(R a, R* y, const R* x, long n) {  // y = a y + x
#pragma omp parallel for schedule(static)
  for (long i = 0; i < n; ++i) y[i] = a * y[i] + x[i];
}
template <typename R>
void scalv(R* y, R a, long n) {
#pragma omp parallel for schedule(static)
  for (long i = 0; i < n; ++i) y[i] *= a;
}

// ---- CG (asolver_CG-tmpl.h:118-234) --------------------------------------
template <typename R>
int cg_solve(Dwf<R>& d, const std::string& mode, R* xq, const R* b, int niter,
             double stop_cond, int init_mode, int* nconv_out, double* diff_out) {
  const long n = n5(d);
  std::vector<R> x(n), r(n), p(n), s(n);
  copy5(d, s.data(), b);
  R sr = (R)norm2v(s.data(), n);
  R snorm = (R)1 / sr;
  R rr, rrp;
  if (init_mode == 0) {  // InitialGuess::RHS
    copy5(d, r.data(), s.data());
    copy5(d, x.data(), s.data());
    mult_mode(d, mode, s.data(), x.data());
    axpyv(r.data(), (R)-1, s.data(), n);
    copy5(d, p.data(), r.data());
  } else {  // InitialGuess::ZERO
    copy5(d, r.data(), s.data());
    std::fill(s.begin(), s.end(), (R)0);
    std::fill(x.begin(), x.end(), (R)0);
    copy5(d, p.data(), r.data());
  }
  rr = (R)norm2v(r.data(), n);
  rrp = rr;
  int nconv = -1;
  for (int iter = 0; iter < niter; ++iter) {
    mult_mode(d, mode, s.data(), p.data());
    R pap = (R)dotv(s.data(), p.data(), n);
    R cr = rrp / pap;
    axpyv(x.data(), cr, p.data(), n);
    axpyv(r.data(), (R)(-cr), s.data(), n);
    rr = (R)norm2v(r.data(), n);
    R bk = rr / rrp;
    aypxv(bk, p.data(), r.data(), n);
    rrp = rr;
    if (rr * snorm < (R)stop_cond) { nconv = iter + 1; break; }
  }
  copy5(d, xq, x.data());
  mult_mode(d, mode, s.data(), xq);
  axpyv(s.data(), (R)-1, b, n);
  R diff2 = (R)norm2v(s.data(), n) * snorm;
  *nconv_out = nconv;
  *diff_out = std::sqrt((double)diff2);
  return nconv < 0 ? 1 : 0;
}

// ---- multishift CG (ashiftsolver_CG-tmpl.h:89-329) ------------------------
int shiftcg_solve(Dwf<double>& d, const std::string& mode, int Nshift,
                  const double* sigma, double** xq, const double* b,
                  int niter, double stop_cond, int* nconv_out, double* diff_out) {
  const long n = n5(d);
  std::vector<std::vector<double>> p(Nshift, std::vector<double>(n)),
      x(Nshift, std::vector<double>(n, 0.0));
  std::vector<double> zeta1(Nshift, 1.0), zeta2(Nshift, 1.0), csh2(Nshift), pp(Nshift, 0.0);
  for (int i = 0; i < Nshift; ++i) csh2[i] = sigma[i] - sigma[0];
  const double sigma0 = sigma[0];
  int Nshift2 = Nshift;
  const double snorm = 1.0 / norm2v(b, n);
  std::vector<double> s(b, b + n), r(b, b + n);
  for (int i = 0; i < Nshift; ++i) copy5(d, p[i].data(), s.data());
  double rr = norm2v(r.data(), n);
  double alpha_p = 0.0, beta_p = 1.0;
  int Nconv2 = -1;
  bool conv = false;
  for (int iter = 0; iter < niter; ++iter) {
    // solve_step
    mult_mode(d, mode, s.data(), p[0].data());
    axpyv(s.data(), sigma0, p[0].data(), n);
    double rr_p = rr;
    double pa_p = dotv(s.data(), p[0].data(), n);
    double beta = -rr_p / pa_p;
    axpyv(x[0].data(), -beta, p[0].data(), n);
    axpyv(r.data(), beta, s.data(), n);
    rr = norm2v(r.data(), n);
    double alpha = rr / rr_p;
    aypxv(alpha, p[0].data(), r.data(), n);
    pp[0] = rr;
    double alpha_h = 1.0 + alpha_p * beta / beta_p;
    for (int ish = 1; ish < Nshift2; ++ish) {
      double zeta = (alpha_h - csh2[ish] * beta) / zeta1[ish] + (1.0 - alpha_h) / zeta2[ish];
      zeta = 1.0 / zeta;
      double zr = zeta / zeta1[ish];
      double beta_s = beta * zr;
      double alpha_s = alpha * zr * zr;
      axpyv(x[ish].data(), -beta_s, p[ish].data(), n);
      scalv(p[ish].data(), alpha_s, n);
      axpyv(p[ish].data(), zeta, r.data(), n);
      double ppr = norm2v(p[ish].data(), n);
      pp[ish] = ppr * snorm;
      zeta2[ish] = zeta1[ish];
      zeta1[ish] = zeta;
    }
    int ish1 = Nshift2;
    for (int ish = Nshift2 - 1; ish >= 0; --ish) {
      if (pp[ish] > stop_cond) { ish1 = ish + 1; break; }
    }
    Nshift2 = ish1;
    alpha_p = alpha;
    beta_p = beta;
    Nconv2 += 1;
    if (rr * snorm < stop_cond) { conv = true; break; }
  }
  double dmax = -1.0;
  for (int i = 0; i < Nshift; ++i) {
    mult_mode(d, mode, s.data(), x[i].data());
    axpyv(s.data(), sigma[i], x[i].data(), n);
    axpyv(s.data(), -1.0, b, n);
    double d1 = std::sqrt(norm2v(s.data(), n) * snorm);
    if (d1 > dmax) dmax = d1;
    copy5(d, xq[i], x[i].data());
  }
  *nconv_out = Nconv2;
  *diff_out = dmax;
  return conv ? 0 : 1;
}

// ---- SU(3) helpers for smearing ------------------------------------------
struct M3 { cd a[3][3]; };
inline M3 load3(const double* p) {
  M3 m;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) m.a[i][j] = cd(p[2 * (3 * i + j)], p[2 * (3 * i + j) + 1]);
  return m;
}
inline void store3(double* p, const M3& m) {
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) { p[2 * (3 * i + j)] = m.a[i][j].real(); p[2 * (3 * i + j) + 1] = m.a[i][j].imag(); }
}
inline M3 mul(const M3& x, const M3& y) {
  M3 r;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      cd s = 0;
      for (int k = 0; k < 3; ++k) s += x.a[i][k] * y.a[k][j];
      r.a[i][j] = s;
    }
  return r;
}
inline M3 dag(const M3& x) {
  M3 r;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) r.a[i][j] = std::conj(x.a[j][i]);
  return r;
}
inline M3 zero3() { M3 r; for (int i = 0; i < 3; ++i) for (int j = 0; j < 3; ++j) r.a[i][j] = 0; return r; }
inline M3 unit3() { M3 r = zero3(); for (int i = 0; i < 3; ++i) r.a[i][i] = 1; return r; }

// anti-hermitian traceless part: (X - X^dag)/2 - tr(X - X^dag)/6
inline M3 at3(const M3& x) {
  M3 r;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) r.a[i][j] = 0.5 * (x.a[i][j] - std::conj(x.a[j][i]));
  cd tr = (r.a[0][0] + r.a[1][1] + r.a[2][2]) / 3.0;
  for (int i = 0; i < 3; ++i) r.a[i][i] -= tr;
  return r;
}

// exp(iQ) by the Cayley-Hamilton form of Morningstar and Peardon,
// Phys. Rev. D 69 (2004) 054501, eqs. (19)-(34); follows
// aprojection_Stout_SU3-tmpl.h:897-977 (set_fj, set_uw, func_xi0).
M3 exp_iQ(const M3& iQ) {
  M3 iQ2 = mul(iQ, iQ);
  M3 iQ3 = mul(iQ, iQ2);
  double n2 = 0;
  for (int i = 0; i < 3; ++i) for (int j = 0; j < 3; ++j) n2 += std::norm(iQ.a[i][j]);
  if (!(n2 > 1.0e-10)) return unit3();
  // Q = -i iQ ; c0 = det Q = tr(Q^3)/3 = -Im tr(iQ^3)/3 ; c1 = tr(Q^2)/2 = -Re tr(iQ^2)/2
  double c0 = -(iQ3.a[0][0].imag() + iQ3.a[1][1].imag() + iQ3.a[2][2].imag()) / 3.0;
  double c1 = -0.5 * (iQ2.a[0][0].real() + iQ2.a[1][1].real() + iQ2.a[2][2].real());
  double c13r = std::sqrt(c1 / 3.0);
  double c0max = 2.0 * c13r * c13r * c13r;
  double ratio = c0 / c0max;
  if (ratio > 1.0) ratio = 1.0;
  if (ratio < -1.0) ratio = -1.0;
  double theta = std::acos(ratio);
  double u = c13r * std::cos(theta / 3.0);
  double w = std::sqrt(c1) * std::sin(theta / 3.0);
  double xi0 = (w == 0.0) ? 1.0 : std::sin(w) / w;
  double u2 = u * u, w2 = w * w, cw = std::cos(w);
  cd emiu(std::cos(u), -std::sin(u));
  cd e2iu(std::cos(2.0 * u), std::sin(2.0 * u));
  cd h0 = e2iu * (u2 - w2) + emiu * cd(8.0 * u2 * cw, 2.0 * u * (3.0 * u2 + w2) * xi0);
  cd h1 = e2iu * (2.0 * u) - emiu * cd(2.0 * u * cw, -(3.0 * u2 - w2) * xi0);
  cd h2 = e2iu - emiu * cd(cw, 3.0 * u * xi0);
  double fden = 1.0 / (9.0 * u2 - w2);
  cd f0 = h0 * fden, f1 = h1 * fden, f2 = h2 * fden;
  // exp(iQ) = f0 + f1 Q + f2 Q^2, Q = -i iQ, Q^2 = -iQ2
  M3 r;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j)
      r.a[i][j] = (i == j ? f0 : cd(0)) + f1 * cd(0, -1) * iQ.a[i][j] - f2 * iQ2.a[i][j];
  return r;
}

// brute force exp by 32 Horner steps of the Taylor series
// (aprojection_Stout_SU3-tmpl.h:980-1018, exp_iQ_bf), without reunitarisation
M3 exp_iQ_taylor(const M3& iQ) {
  const int Nprec = 32;
  M3 u1 = unit3();
  for (int k = 0; k < Nprec; ++k) {
    double exf = 1.0 / (Nprec - k);
    M3 u2 = mul(iQ, u1);
    for (int i = 0; i < 3; ++i) for (int j = 0; j < 3; ++j) u1.a[i][j] = u2.a[i][j] * exf + (i == j ? 1.0 : 0.0);
  }
  return u1;
}

struct GaugeGeom {
  int N[4];
  long V;
  long idx(const int x[4]) const { return x[0] + N[0] * (x[1] + (long)N[1] * (x[2] + (long)N[2] * x[3])); }
  void coord(long s, int x[4]) const {
    x[0] = s % N[0]; s /= N[0]; x[1] = s % N[1]; s /= N[1]; x[2] = s % N[2]; x[3] = s / N[2];
  }
  long shift(long s, int mu, int dir) const {
    int x[4]; coord(s, x);
    x[mu] = (x[mu] + dir + N[mu]) % N[mu];
    return idx(x);
  }
};

// ---- DWF fermion force (aforce_F_Domainwall_5din_eo-tmpl.h:299-451 and the
// lexical aforce_F_Domainwall_5din-tmpl.h:338-933), restated on the two parity
// fields instead of a lexical 5D field.

// y = -1/2 [b_s w_s + c_s (P- w_{s+1} + P+ w_{s-1})], the wall terms times -mq
// (aforce_F_Domainwall_5din-tmpl.h:355-365: b1 = -b/2, Fup = Fdn = -c/4 with 1 -/+ g5)
void m5_force_weight(Dwf<double>& d, double* y, const double* w) {
  const int Ls = d.Ls;
  const long Vh = d.g.Vh;
#pragma omp parallel for schedule(static)
  for (long s2 = 0; s2 < Vh; ++s2) {
    for (int s = 0; s < Ls; ++s) {
      const int su = (s + 1) % Ls, sd = (s - 1 + Ls) % Ls;
      const double fup = (s == Ls - 1) ? -d.mq : 1.0;
      const double fdn = (s == 0) ? -d.mq : 1.0;
      const double* ws = w + 24 * (s2 + Vh * s);
      const double* wu = w + 24 * (s2 + Vh * su);
      const double* wd = w + 24 * (s2 + Vh * sd);
      double* ys = y + 24 * (s2 + Vh * s);
      for (int k = 0; k < 12; ++k) {
        // P- v = (v - g5 v)/2, P+ v = (v + g5 v)/2; g5 exchanges k and k + 12
        const double pm_up = 0.5 * (wu[k] - wu[k + 12]);     // upper half of P- w_{s+1}
        const double pp_dn = 0.5 * (wd[k] + wd[k + 12]);     // upper half of P+ w_{s-1}
        const double mv_hi = fup * pm_up + fdn * pp_dn;
        const double mv_lo = -fup * pm_up + fdn * pp_dn;      // lower halves: P- -> -, P+ -> +
        ys[k] = -0.5 * (d.b[s] * ws[k] + d.c[s] * mv_hi);
        ys[k + 12] = -0.5 * (d.b[s] * ws[k + 12] + d.c[s] * mv_lo);
      }
    }
  }
}

// v <- (1 + sgn gamma_mu) v for one 4-spinor (Dirac basis, gammaMatrixSet_Dirac.cpp:37-65)
inline void one_pm_gamma(double* out, const double* v, int mu, int sgn) {
  cd x[4][3], g[4][3];
  for (int dd = 0; dd < 4; ++dd)
    for (int c = 0; c < 3; ++c) x[dd][c] = cd(v[2 * c + 6 * dd], v[2 * c + 6 * dd + 1]);
  const cd I(0, 1);
  for (int c = 0; c < 3; ++c) {
    if (mu == 0) { g[0][c] = -I * x[3][c]; g[1][c] = -I * x[2][c]; g[2][c] = I * x[1][c]; g[3][c] = I * x[0][c]; }
    else if (mu == 1) { g[0][c] = -x[3][c]; g[1][c] = x[2][c]; g[2][c] = x[1][c]; g[3][c] = -x[0][c]; }
    else if (mu == 2) { g[0][c] = -I * x[2][c]; g[1][c] = I * x[3][c]; g[2][c] = I * x[0][c]; g[3][c] = -I * x[1][c]; }
    else { g[0][c] = x[0][c]; g[1][c] = x[1][c]; g[2][c] = -x[2][c]; g[3][c] = -x[3][c]; }
  }
  for (int dd = 0; dd < 4; ++dd)
    for (int c = 0; c < 3; ++c) {
      const cd r = x[dd][c] + (double)sgn * g[dd][c];
      out[2 * c + 6 * dd] = r.real();
      out[2 * c + 6 * dd + 1] = r.imag();
    }
}

// F_mu(x) += sum_{s,d} W(x)_{a,d,s} conj(Z(x)_{b,d,s}),  W(x) = (1 + sgn g_mu) V(x+mu)
// times the boundary factor when x+mu wraps (tensorProd_5din, afield_Spinor-inc.h:322-370,
// after the backward shift with boundary conditions).  x runs over parity p.
void tensor_fwd(Dwf<double>& d, double* F, const double* Z, const double* Vn, int p, int sgn) {
  const int Ls = d.Ls;
  const long Vh = d.g.Vh, V = d.g.V;
#pragma omp parallel for schedule(static)
  for (long s2 = 0; s2 < Vh; ++s2) {
    for (int mu = 0; mu < 4; ++mu) {
      const int nb = d.nbr[p][mu][0][s2];
      const double bcv = d.bcf[p][mu][0][s2];
      const long lx = d.lnk[p][mu][0][s2];
      cd acc[3][3] = {};
      for (int s = 0; s < Ls; ++s) {
        double w[24];
        one_pm_gamma(w, Vn + 24 * (nb + Vh * s), mu, sgn);
        const double* z = Z + 24 * (s2 + Vh * s);
        for (int a = 0; a < 3; ++a)
          for (int b = 0; b < 3; ++b)
            for (int dd = 0; dd < 4; ++dd)
              acc[a][b] += cd(w[2 * a + 6 * dd], w[2 * a + 6 * dd + 1]) *
                           std::conj(cd(z[2 * b + 6 * dd], z[2 * b + 6 * dd + 1]));
      }
      double* f = F + 18 * (lx + V * mu);
      for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b) {
          f[2 * (3 * a + b)] += bcv * acc[a][b].real();
          f[2 * (3 * a + b) + 1] += bcv * acc[a][b].imag();
        }
    }
  }
}

// force_udiv1_D / force_udiv1_Ddag of the even-odd operator; output -F (the
// eo wrapper's sign), F lexical Field_G (18, V, 4).
void force_udiv1_eo(Dwf<double>& d, double* F, const double* zeta, const double* eta, bool dag) {
  const long n = n5(d), ng = 18L * d.g.V * 4;
  std::vector<double> a_e(n), a_o(n), b_e(n), b_o(n), t(n), y_e(n), y_o(n);
  // "plain" side: (x_e, Doo^-1 Doe x_e); "dagger" side: (Dee^-+ x, Doo^-+ Deo^+ Dee^-+ x)
  const double* plain = dag ? zeta : eta;
  const double* dside = dag ? eta : zeta;
  std::copy(plain, plain + n, a_e.begin());
  mult_mode(d, "Doe", t.data(), plain);
  mult_mode(d, "Doo_inv", a_o.data(), t.data());
  mult_dag_mode(d, "Dee_inv", b_e.data(), dside);
  mult_dag_mode(d, "Deo", t.data(), b_e.data());
  mult_dag_mode(d, "Doo_inv", b_o.data(), t.data());
  std::fill(F, F + ng, 0.0);
  if (!dag) {
    // W = (1 - g_mu) y(x + mu) with y the M5-weighted eta, Z = dagger-side zeta
    m5_force_weight(d, y_e.data(), a_e.data());
    m5_force_weight(d, y_o.data(), a_o.data());
    tensor_fwd(d, F, b_e.data(), y_o.data(), 0, -1);
    tensor_fwd(d, F, b_o.data(), y_e.data(), 1, -1);
  } else {
    // W = (1 + g_mu) eta'(x + mu), Z = the M5-weighted zeta
    m5_force_weight(d, y_e.data(), a_e.data());
    m5_force_weight(d, y_o.data(), a_o.data());
    tensor_fwd(d, F, y_e.data(), b_o.data(), 0, +1);
    tensor_fwd(d, F, y_o.data(), b_e.data(), 1, +1);
  }
  for (long i = 0; i < ng; ++i) F[i] = -F[i];
}

// mult_generator (lib_alt_QXS/Force/Fermion/aforce_F.cpp:97-110): F <- -2 at(U F)
void mult_generator(Dwf<double>& d, double* F) {
  const long V = d.g.V;
#pragma omp parallel for schedule(static)
  for (long site = 0; site < V; ++site)
    for (int mu = 0; mu < 4; ++mu) {
      double* f = F + 18 * (site + V * mu);
      M3 UF = mul(load3(d.U.data() + 18 * (site + V * mu)), load3(f));
      M3 A = at3(UF);
      for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) A.a[i][j] *= -2.0;
      store3(f, A);
    }
}

}  // namespace

// ===========================================================================
// C API
// ===========================================================================
struct orc_dwf_s {
  int prec;
  Dwf<double> d64;
  Dwf<float> d32;
};

template <typename R>
static void init_dwf(Dwf<R>& d, const int dims[4], int Ls, double mq, double M0,
                     const double* b, const double* c, double alpha, const int bc[4]) {
  for (int mu = 0; mu < 4; ++mu) d.g.N[mu] = dims[mu];
  d.g.Nx2 = dims[0] / 2;
  d.g.V = dims[0] * dims[1] * dims[2] * dims[3];
  d.g.Vh = d.g.V / 2;
  d.Ls = Ls;
  d.mq = (R)mq; d.M0 = (R)M0; d.alpha = (R)alpha;
  d.b.resize(Ls); d.c.resize(Ls);
  for (int s = 0; s < Ls; ++s) { d.b[s] = (R)b[s]; d.c[s] = (R)c[s]; }
  for (int mu = 0; mu < 4; ++mu) d.bc[mu] = bc[mu];
  d.kappa = 1.0 / (8.0 - 2.0 * M0);
  set_precond(d);
  build_tables(d);
  const long n4 = 24L * d.g.Vh;
  d.w4.assign(n4, 0); d.v4.assign(n4, 0); d.y4.assign(n4, 0); d.t4.assign(n4, 0);
  d.v1.assign(n4 * Ls, 0); d.v2.assign(n4 * Ls, 0); d.w1.assign(n4 * Ls, 0);
}

extern "C" {

orc_dwf orc_dwf_create(const int dims[4], int Ls, double mq, double M0,
                       const double* b, const double* c, double alpha,
                       const int bc[4], int prec) {
  if (dims[0] % 2 != 0 || Ls < 2) return nullptr;
  orc_dwf h = new orc_dwf_s;
  h->prec = prec;
  if (prec == 0) init_dwf(h->d64, dims, Ls, mq, M0, b, c, alpha, bc);
  else init_dwf(h->d32, dims, Ls, mq, M0, b, c, alpha, bc);
  return h;
}

void orc_dwf_destroy(orc_dwf h) { delete h; }

void orc_dwf_set_config(orc_dwf h, const double* U) {
  if (h->prec == 0) {
    Dwf<double>& d = h->d64;
    d.U.assign(U, U + 18L * d.g.V * 4);
    d.have_U = true;
  } else {
    Dwf<float>& d = h->d32;
    d.U.resize(18L * d.g.V * 4);
    for (long i = 0; i < 18L * d.g.V * 4; ++i) d.U[i] = (float)U[i];
    d.have_U = true;
  }
}

int orc_dwf_nvol2(orc_dwf h) { return h->prec == 0 ? h->d64.g.Vh : h->d32.g.Vh; }
int orc_dwf_ls(orc_dwf h) { return h->prec == 0 ? h->d64.Ls : h->d32.Ls; }

static int run_f32(orc_dwf h, const char* mode, double* v, const double* w, bool dagger) {
  Dwf<float>& d = h->d32;
  const long n = n5(d);
  std::vector<float> wf(n), vf(n);
  for (long i = 0; i < n; ++i) wf[i] = (float)w[i];
  int rc = dagger ? mult_dag_mode(d, mode, vf.data(), wf.data())
                  : mult_mode(d, mode, vf.data(), wf.data());
  for (long i = 0; i < n; ++i) v[i] = (double)vf[i];
  return rc;
}

int orc_mult(orc_dwf h, const char* mode, double* v, const double* w) {
  if (h->prec == 0) return mult_mode(h->d64, mode, v, w);
  return run_f32(h, mode, v, w, false);
}
int orc_mult_dag(orc_dwf h, const char* mode, double* v, const double* w) {
  if (h->prec == 0) return mult_dag_mode(h->d64, mode, v, w);
  return run_f32(h, mode, v, w, true);
}
void orc_mult_gm5(orc_dwf h, double* v, const double* w) { gm5_5d(h->d64, v, w); }
void orc_mult_gm5R(orc_dwf h, double* v, const double* w) { gm5R(h->d64, v, w); }
void orc_wilson_meo(orc_dwf h, double* v4, const double* w4, int ieo) { wilson_meo(h->d64, v4, w4, ieo); }

int orc_cg_solve(orc_dwf h, const char* mode, double* x, const double* b,
                 int niter, double stop_cond2, int init_mode, int* nconv, double* diff) {
  return cg_solve(h->d64, mode, x, b, niter, stop_cond2, init_mode, nconv, diff);
}

// ASolver_Richardson::solve (asolver_Richardson-tmpl.h:131-213) with
// APrecond_Mixedprec::mult (aprecond_Mixedprec-tmpl.h:55-108): the
// preconditioner rounds r to float, runs the float CG from a zero guess and
// widens the result back to double.
int orc_richardson_solve(orc_dwf h64, orc_dwf h32, const char* mode,
                         double* xq, const double* b, int niter_outer,
                         double stop_outer2, int niter_inner, double stop_inner2,
                         int* nconv_out, double* diff_out, int* inner_iters) {
  Dwf<double>& d = h64->d64;
  Dwf<float>& df = h32->d32;
  const long n = n5(d);
  std::vector<double> r(n), s(n);
  std::vector<float> rf(n), sf(n);
  double bnorm2 = norm2v(b, n);
  *inner_iters = 0;
  if (!(bnorm2 > 0.0)) {
    std::fill(xq, xq + n, 0.0);
    *nconv_out = 0; *diff_out = 0.0;
    return 0;
  }
  double bnorm2_inv = 1.0 / bnorm2;
  for (long i = 0; i < n; ++i) r[i] = -b[i];
  std::fill(xq, xq + n, 0.0);
  int nconv = -1;
  double r2 = bnorm2;
  for (int iter = 0; iter < niter_outer; ++iter) {
    if (r2 < bnorm2 * stop_outer2) { nconv = iter; break; }
    double normalization = std::sqrt(r2);
    scalv(r.data(), 1.0 / normalization, n);
    for (long i = 0; i < n; ++i) rf[i] = (float)r[i];
    int inc = 0; double idiff = 0;
    cg_solve(df, std::string(mode), sf.data(), rf.data(), niter_inner, stop_inner2, 2, &inc, &idiff);
    if (inc > 0) *inner_iters += inc; else *inner_iters += niter_inner;
    for (long i = 0; i < n; ++i) s[i] = (double)sf[i];
    axpyv(xq, -normalization, s.data(), n);
    mult_mode(d, mode, r.data(), xq);
    axpyv(r.data(), -1.0, b, n);
    r2 = norm2v(r.data(), n);
  }
  mult_mode(d, mode, r.data(), xq);
  axpyv(r.data(), -1.0, b, n);
  double diff2 = norm2v(r.data(), n);
  *nconv_out = nconv;
  *diff_out = std::sqrt(diff2 * bnorm2_inv);
  return nconv < 0 ? 1 : 0;
}

int orc_shiftcg_solve(orc_dwf h, const char* mode, int nshift, const double* sigma,
                      double** x, const double* b, int niter, double stop_cond2,
                      int* nconv, double* diff) {
  return shiftcg_solve(h->d64, mode, nshift, sigma, x, b, niter, stop_cond2, nconv, diff);
}

double orc_norm2(const double* x, long n) { return norm2v(x, n); }

// AForce_F_Domainwall_5din_eo::force_udiv1 in mode D, Ddag, H or Hdag
// (aforce_F_Domainwall_5din_eo-tmpl.h:318-451); H/Hdag apply gamma_5 R to
// zeta / eta first (aforce_F_Domainwall_5din-tmpl.h:913-933).
int orc_force_udiv1(orc_dwf h, const char* mode, double* F, const double* zeta,
                    const double* eta) {
  Dwf<double>& d = h->d64;
  std::string m(mode);
  if (m == "D") force_udiv1_eo(d, F, zeta, eta, false);
  else if (m == "Ddag") force_udiv1_eo(d, F, zeta, eta, true);
  else if (m == "H" || m == "Hdag") {
    std::vector<double> t(n5(d));
    if (m == "H") { gm5R(d, t.data(), zeta); force_udiv1_eo(d, F, t.data(), eta, false); }
    else { gm5R(d, t.data(), eta); force_udiv1_eo(d, F, zeta, t.data(), true); }
  } else return -1;
  return 0;
}
// force_udiv (mode D): z = D eta; udiv1_D(z, eta) + udiv1_Ddag(eta, z)
// (aforce_F_Domainwall_5din_eo-tmpl.h:299-315)
void orc_force_udiv(orc_dwf h, double* F, const double* eta) {
  Dwf<double>& d = h->d64;
  const long ng = 18L * d.g.V * 4;
  std::vector<double> z(n5(d)), F2(ng);
  mult_mode(d, "D", z.data(), eta);
  force_udiv1_eo(d, F, z.data(), eta, false);
  force_udiv1_eo(d, F2.data(), eta, z.data(), true);
  for (long i = 0; i < ng; ++i) F[i] += F2[i];
}
void orc_mult_generator(orc_dwf h, double* F) { mult_generator(h->d64, F); }
double orc_dot(const double* x, const double* y, long n) { return dotv(x, y, n); }

// One APE-staple + stout projection level (asmear_APE-tmpl.h:183-223):
//   C_mu = rho sum_{nu != mu} (upper + lower staples)
//   upper = U_nu(x) U_mu(x+nu) U_nu(x+mu)^dag
//   lower = U_nu(x-nu)^dag U_mu(x-nu) U_nu(x-nu+mu)   (astaple_lex-tmpl.h:221-256)
//   U'_mu(x) = exp(at(C_mu U_mu^dag)) U_mu(x)       (aprojection_Stout_SU3-tmpl.h:477-553)
void orc_stout_smear(const int dims[4], double rho, double* Uout, const double* Uin) {
  GaugeGeom g;
  for (int mu = 0; mu < 4; ++mu) g.N[mu] = dims[mu];
  g.V = (long)dims[0] * dims[1] * dims[2] * dims[3];
  const long V = g.V;
#pragma omp parallel for schedule(static)
  for (long site = 0; site < V; ++site) {
    for (int mu = 0; mu < 4; ++mu) {
      M3 C = zero3();
      for (int nu = 0; nu < 4; ++nu) {
        if (nu == mu) continue;
        long xpmu = g.shift(site, mu, +1);
        long xpnu = g.shift(site, nu, +1);
        long xmnu = g.shift(site, nu, -1);
        long xmnupmu = g.shift(xmnu, mu, +1);
        M3 up = mul(mul(load3(Uin + 18 * (site + V * nu)), load3(Uin + 18 * (xpnu + V * mu))),
                    dag(load3(Uin + 18 * (xpmu + V * nu))));
        M3 lo = mul(mul(dag(load3(Uin + 18 * (xmnu + V * nu))), load3(Uin + 18 * (xmnu + V * mu))),
                    load3(Uin + 18 * (xmnupmu + V * nu)));
        for (int i = 0; i < 3; ++i)
          for (int j = 0; j < 3; ++j) C.a[i][j] += rho * up.a[i][j] + rho * lo.a[i][j];
      }
      M3 U = load3(Uin + 18 * (site + V * mu));
      M3 iQ = at3(mul(C, dag(U)));
      store3(Uout + 18 * (site + V * mu), mul(exp_iQ(iQ), U));
    }
  }
}

// ---- stout force recursion --------------------------------------------------
// AForceSmear_APE::force_udiv for one level (lib_alt_QXS/Smear/aforceSmear_APE-tmpl.h:157-206):
// given Sigma' = dS/dU' of the smeared links U' = exp(iQ) U, returns Sigma = dS/dU.
// Per link, AProjection_Stout_SU3::force_recursive (aprojection_Stout_SU3-tmpl.h:691-895),
// the chain rule of Morningstar and Peardon, Phys. Rev. D 69 (2004) 054501, eqs. (54)-(75):
//   Gamma = tr(U S' B1) Q + tr(U S' B2) Q^2 + f1 U S' + f2 (Q U S' + U S' Q),
//   i Lambda = at(i Gamma),  iTheta = i Lambda U,  Xi = S' exp(iQ) + C^+ i Lambda;
// then the staple terms of force_each (aforceSmear_APE-tmpl.h:280-360).

namespace {

inline M3 add3(const M3& x, const M3& y) {
  M3 r;
  for (int i = 0; i < 3; ++i) for (int j = 0; j < 3; ++j) r.a[i][j] = x.a[i][j] + y.a[i][j];
  return r;
}
inline M3 scal3(cd s, const M3& x) {
  M3 r;
  for (int i = 0; i < 3; ++i) for (int j = 0; j < 3; ++j) r.a[i][j] = s * x.a[i][j];
  return r;
}
inline cd tr3(const M3& x) { return x.a[0][0] + x.a[1][1] + x.a[2][2]; }
inline M3 poly3(cd c0, cd c1, cd c2, const M3& Q, const M3& Q2) {
  M3 r = add3(scal3(c1, Q), scal3(c2, Q2));
  for (int i = 0; i < 3; ++i) r.a[i][i] += c0;
  return r;
}

// xi1(w) = (w cos w - sin w) / w^3, by its Taylor series for small w
double stout_xi1(double w) {
  if (w < 0.25) {
    const double w2 = w * w;
    return -1.0 / 3.0 + w2 * (1.0 / 30.0 + w2 * (-1.0 / 840.0 + w2 * (1.0 / 45360.0 - w2 / 3991680.0)));
  }
  return (w * std::cos(w) - std::sin(w)) / (w * w * w);
}

// one link: Xi and iTheta from C, U and Sigma'
void stout_link_force(const M3& C, const M3& U, const M3& Sp, M3& Xi, M3& iTheta) {
  M3 iQ = at3(mul(C, dag(U)));
  M3 iQ2 = mul(iQ, iQ), iQ3 = mul(iQ, iQ2);
  double n2 = 0;
  for (int i = 0; i < 3; ++i) for (int j = 0; j < 3; ++j) n2 += std::norm(iQ.a[i][j]);
  M3 eiQ = unit3(), iGam = zero3();
  if (n2 > 1.0e-10) {
    const M3 Q = scal3(cd(0, -1), iQ);
    const M3 Q2 = scal3(-1.0, iQ2);
    double c0 = -(iQ3.a[0][0].imag() + iQ3.a[1][1].imag() + iQ3.a[2][2].imag()) / 3.0;
    double c1 = -0.5 * (iQ2.a[0][0].real() + iQ2.a[1][1].real() + iQ2.a[2][2].real());
    double c13r = std::sqrt(c1 / 3.0);
    double ratio = c0 / (2.0 * c13r * c13r * c13r);
    if (ratio > 1.0) ratio = 1.0;
    if (ratio < -1.0) ratio = -1.0;
    double theta = std::acos(ratio);
    double u = c13r * std::cos(theta / 3.0), w = std::sqrt(c1) * std::sin(theta / 3.0);
    double u2 = u * u, w2 = w * w, cw = std::cos(w);
    double xi0 = (w == 0.0) ? 1.0 : std::sin(w) / w, xi1 = stout_xi1(w);
    cd emiu(std::cos(u), -std::sin(u)), e2iu(std::cos(2.0 * u), std::sin(2.0 * u));
    // f_j (eq. 29-30)
    cd h0 = e2iu * (u2 - w2) + emiu * cd(8.0 * u2 * cw, 2.0 * u * (3.0 * u2 + w2) * xi0);
    cd h1 = e2iu * (2.0 * u) - emiu * cd(2.0 * u * cw, -(3.0 * u2 - w2) * xi0);
    cd h2 = e2iu - emiu * cd(cw, 3.0 * u * xi0);
    double fd = 1.0 / (9.0 * u2 - w2);
    cd f0 = h0 * fd, f1 = h1 * fd, f2 = h2 * fd;
    // r_j^(1), r_j^(2) (eqs. 60-65)
    cd r01 = cd(2.0 * u, 2.0 * (u2 - w2)) * e2iu +
             emiu * cd(16.0 * u * cw + 2.0 * u * (3.0 * u2 + w2) * xi0,
                       -8.0 * u2 * cw + 2.0 * (9.0 * u2 + w2) * xi0);
    cd r11 = cd(2.0, 4.0 * u) * e2iu + emiu * cd(-2.0 * cw + (3.0 * u2 - w2) * xi0,
                                                2.0 * u * cw + 6.0 * u * xi0);
    cd r21 = cd(0.0, 2.0) * e2iu + emiu * cd(-3.0 * u * xi0, cw - 3.0 * xi0);
    cd r02 = cd(-2.0, 0.0) * e2iu + emiu * cd(-8.0 * u2 * xi0, 2.0 * u * (cw + xi0 + 3.0 * u2 * xi1));
    cd r12 = emiu * cd(2.0 * u * xi0, -cw - xi0 + 3.0 * u2 * xi1);
    cd r22 = emiu * cd(xi0, -3.0 * u * xi1);
    // b_ij (eqs. 57-58)
    double bd = 1.0 / (2.0 * (9.0 * u2 - w2) * (9.0 * u2 - w2));
    double g1 = 3.0 * u2 - w2, g2 = 30.0 * u2 + 2.0 * w2;
    cd b10 = (2.0 * u * r01 + g1 * r02 - g2 * f0) * bd;
    cd b11 = (2.0 * u * r11 + g1 * r12 - g2 * f1) * bd;
    cd b12 = (2.0 * u * r21 + g1 * r22 - g2 * f2) * bd;
    cd b20 = (r01 - 3.0 * u * r02 - 24.0 * u * f0) * bd;
    cd b21 = (r11 - 3.0 * u * r12 - 24.0 * u * f1) * bd;
    cd b22 = (r21 - 3.0 * u * r22 - 24.0 * u * f2) * bd;
    M3 B1 = poly3(b10, b11, b12, Q, Q2), B2 = poly3(b20, b21, b22, Q, Q2);
    M3 US = mul(U, Sp);
    cd t1 = tr3(mul(US, B1)), t2 = tr3(mul(US, B2));
    M3 Gam = add3(add3(scal3(t1, Q), scal3(t2, Q2)),
                  add3(scal3(f1, US), scal3(f2, add3(mul(Q, US), mul(US, Q)))));
    iGam = scal3(cd(0, 1), Gam);
    eiQ = poly3(f0, f1, f2, Q, Q2);
  }
  const M3 iLam = at3(iGam);
  iTheta = mul(iLam, U);
  Xi = add3(mul(Sp, eiQ), mul(dag(C), iLam));
}

}  // namespace

void orc_stout_force_udiv(const int dims[4], double rho, double* Sigma, const double* Sigmap,
                          const double* Uin) {
  GaugeGeom g;
  for (int mu = 0; mu < 4; ++mu) g.N[mu] = dims[mu];
  g.V = (long)dims[0] * dims[1] * dims[2] * dims[3];
  const long V = g.V;
  auto Ul = [&](long site, int mu) { return load3(Uin + 18 * (site + V * mu)); };
  std::vector<double> iTh(18L * V * 4);
  auto Th = [&](long site, int mu) { return load3(iTh.data() + 18 * (site + V * mu)); };
#pragma omp parallel for schedule(static)
  for (long site = 0; site < V; ++site) {
    for (int mu = 0; mu < 4; ++mu) {
      M3 C = zero3();
      for (int nu = 0; nu < 4; ++nu) {
        if (nu == mu) continue;
        long xpmu = g.shift(site, mu, +1), xpnu = g.shift(site, nu, +1);
        long xmnu = g.shift(site, nu, -1), xmnupmu = g.shift(xmnu, mu, +1);
        M3 up = mul(mul(Ul(site, nu), Ul(xpnu, mu)), dag(Ul(xpmu, nu)));
        M3 lo = mul(mul(dag(Ul(xmnu, nu)), Ul(xmnu, mu)), Ul(xmnupmu, nu));
        C = add3(C, scal3(rho, add3(up, lo)));
      }
      M3 Xi, iTheta;
      stout_link_force(C, Ul(site, mu), load3(Sigmap + 18 * (site + V * mu)), Xi, iTheta);
      store3(Sigma + 18 * (site + V * mu), Xi);
      store3(iTh.data() + 18 * (site + V * mu), iTheta);
    }
  }
  // staple terms: Sigma_mu(x) += rho sum_{nu != mu} (six products of U and iTheta)
#pragma omp parallel for schedule(static)
  for (long site = 0; site < V; ++site) {
    for (int mu = 0; mu < 4; ++mu) {
      M3 acc = zero3();
      for (int nu = 0; nu < 4; ++nu) {
        if (nu == mu) continue;
        long xpmu = g.shift(site, mu, +1), xpnu = g.shift(site, nu, +1);
        long xmnu = g.shift(site, nu, -1), xmnupmu = g.shift(xmnu, mu, +1);
        M3 t = zero3();
        // V_nu(x+mu) V_mu(x+nu)^+ iTheta_nu(x)^+
        t = add3(t, mul(mul(Ul(xpmu, nu), dag(Ul(xpnu, mu))), dag(Th(site, nu))));
        // V_nu(x-nu+mu)^+ iTheta_mu(x-nu)^+ V_nu(x-nu)
        t = add3(t, mul(mul(dag(Ul(xmnupmu, nu)), dag(Th(xmnu, mu))), Ul(xmnu, nu)));
        // V_nu(x-nu+mu)^+ V_mu(x-nu)^+ iTheta_nu(x-nu)
        t = add3(t, mul(mul(dag(Ul(xmnupmu, nu)), dag(Ul(xmnu, mu))), Th(xmnu, nu)));
        // iTheta_nu(x+mu) V_mu(x+nu)^+ V_nu(x)^+
        t = add3(t, mul(mul(Th(xpmu, nu), dag(Ul(xpnu, mu))), dag(Ul(site, nu))));
        // iTheta_nu(x-nu+mu)^+ V_mu(x-nu)^+ V_nu(x-nu)
        t = add3(t, mul(mul(dag(Th(xmnupmu, nu)), dag(Ul(xmnu, mu))), Ul(xmnu, nu)));
        // V_nu(x+mu) iTheta_mu(x+nu)^+ V_nu(x)^+
        t = add3(t, mul(mul(Ul(xpmu, nu), dag(Th(xpnu, mu))), dag(Ul(site, nu))));
        acc = add3(acc, scal3(rho, t));
      }
      double* out = Sigma + 18 * (site + V * mu);
      store3(out, add3(load3(out), acc));
    }
  }
}

// mult_generator with a given gauge field: F <- -2 at(U F)
void orc_mult_generator_gauge(const int dims[4], double* F, const double* U) {
  const long V = (long)dims[0] * dims[1] * dims[2] * dims[3];
#pragma omp parallel for schedule(static)
  for (long site = 0; site < V; ++site)
    for (int mu = 0; mu < 4; ++mu) {
      double* f = F + 18 * (site + V * mu);
      M3 A = at3(mul(load3(U + 18 * (site + V * mu)), load3(f)));
      store3(f, scal3(-2.0, A));
    }
}

double orc_plaquette(const int dims[4], const double* U) {
  GaugeGeom g;
  for (int mu = 0; mu < 4; ++mu) g.N[mu] = dims[mu];
  g.V = (long)dims[0] * dims[1] * dims[2] * dims[3];
  const long V = g.V;
  double sum = 0;
#pragma omp parallel for reduction(+ : sum) schedule(static)
  for (long site = 0; site < V; ++site) {
    for (int mu = 0; mu < 4; ++mu)
      for (int nu = mu + 1; nu < 4; ++nu) {
        long xpmu = g.shift(site, mu, +1), xpnu = g.shift(site, nu, +1);
        M3 P = mul(mul(load3(U + 18 * (site + V * mu)), load3(U + 18 * (xpmu + V * nu))),
                   mul(dag(load3(U + 18 * (xpnu + V * mu))), dag(load3(U + 18 * (site + V * nu)))));
        sum += (P.a[0][0] + P.a[1][1] + P.a[2][2]).real();
      }
  }
  return sum / (3.0 * 6.0 * (double)V);
}

void orc_exp_iQ(double* e, const double* iQ) { store3(e, exp_iQ(load3(iQ))); }
void orc_exp_iQ_taylor(double* e, const double* iQ) { store3(e, exp_iQ_taylor(load3(iQ))); }

int orc_num_threads(void) { return omp_get_max_threads(); }

}  // extern "C"
