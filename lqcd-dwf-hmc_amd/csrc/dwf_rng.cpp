// dwf_rng.cpp -- the random numbers of the HMC (host side, as in the reference).
//
// M-series generator: the lagged-Fibonacci XOR generator of J. Makino,
// "Lagged-Fibonacci random number generators on parallel computers",
// Parallel Computing 20 (1994) 1357-1367, with lags (521, 32): the word
// sequence w_n = w_{n-521} ^ w_{n-489} of 31-bit words, started from a
// 69069-LCG bit sequence and advanced ("delayed") by n steps through the
// characteristic polynomial x^521 + x^32 + 1 of the recurrence
// (RandomNumbers_Mseries::initset / delay3 / get,
// src/lib/Tools/randomNumbers_Mseries.cpp:29-145, randomNumbers_Mseries.h).
// Gaussian pairs by Box-Muller with variance 1/2 (RandomNumbers::gauss,
// src/lib/Tools/randomNumbers.cpp:64-76).
// Parallel streams: one generator per (z, t) slice, seeded (seed + 1) Lz Lt + z
// + Lz t and advanced 1000 draws, filling the lexical sites of its slice in x,
// y order, the extra index outermost (the streams run on host threads)
// (RandomNumbers_parallel::init_2d /
// set_lex_global_2d, src/lib/Tools/randomNumbers_parallel-tmpl.h:146-230,
// randomNumbers_parallel_Mseries.cpp:29-38).  The numbers must be
// bit-identical to the reference's for the HMC energies to agree, so the
// arithmetic follows those steps exactly.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <string>
#include <thread>
#include <vector>

#include "dwf_hip.h"

namespace {

class Mseries {
 public:
  static constexpr int Np = 521, Nq = 32;
  explicit Mseries(int ndelay) { initset(ndelay); }
  double get() {
    w_[jr_] = w_[jr_] ^ w_[kr_];
    const double rw = w_[jr_] * kFnorm;
    if (++jr_ >= Np) jr_ -= Np;
    if (++kr_ >= Np) kr_ -= Np;
    return rw;
  }
  // two Gaussian numbers of variance 1/2
  void gauss(double& r1, double& r2) {
    const double u1 = get();
    const double u2 = get();
    const double slg = std::sqrt(-std::log(1.0 - u1) * 2.0) * kSq2r;
    const double ang = kPi2 * u2;
    r1 = slg * std::cos(ang);
    r2 = slg * std::sin(ang);
  }

 private:
  static constexpr double kFnorm = 4.656612870908988e-10;
  static const double kSq2r, kPi2;
  int w_[Np];
  int jr_ = 0, kr_ = 0;

  void initset(int ndelay) {
    unsigned bits[Np];
    unsigned ix = 1;
    for (int i = 0; i < Np; ++i) {
      ix *= 69069u;
      bits[i] = ix >> 31;
    }
    int j0 = 0, k0 = Np - Nq;
    for (int j = 0; j < Np; ++j) {
      unsigned word = 0;
      for (int i = 0; i < 32; ++i) {
        word = word * 2 + bits[j0];
        bits[j0] ^= bits[k0];
        if (++j0 == Np) j0 = 0;
        if (++k0 == Np) k0 = 0;
      }
      w_[j] = (int)(word >> 1);
    }
    jr_ = j0;
    kr_ = k0;
    delay(ndelay);
  }

  // advance the word sequence by n * 2^40 steps: x^(n 2^40) mod
  // (x^Np + x^Nq + 1) by binary powering, then w'_j = sum_i c_i w_{j+i} over GF(2)
  void delay(int n) {
    constexpr int kShift = 40;   // the 2^40 spacing of the streams
    std::vector<unsigned> ext(2 * Np - 1);
    for (int i = 0; i < Np; ++i) ext[i] = (unsigned)w_[i];
    for (int i = Np; i < 2 * Np - 1; ++i) ext[i] = ext[i - Np] ^ ext[i - Nq];
    // binary digits of n 2^40 above the part of n that is < Np, least
    // significant first: kShift zeros, then the digits of n
    std::vector<int> dig(kShift, 0);
    int m = n;
    while (m > Np - 1) {
      dig.push_back(m % 2);
      m /= 2;
    }
    std::vector<unsigned> c(2 * Np, 0);
    c[m] = 1;
    for (int d = (int)dig.size() - 1; d >= 0; --d) {
      const int b = dig[d];
      for (int i = Np - 1; i >= 0; --i) {   // square, times x^b
        c[2 * i + b] = c[i];
        c[2 * i + 1 - b] = 0;
      }
      for (int i = 2 * Np - 1; i >= Np; --i) {   // reduce
        c[i - Np] ^= c[i];
        c[i - Nq] ^= c[i];
      }
    }
    for (int j = 0; j < Np; ++j) {
      unsigned word = 0;
      for (int i = 0; i < Np; ++i) word ^= c[i] * ext[j + i];
      w_[j] = (int)word;
    }
  }
};
const double Mseries::kSq2r = 1.0 / std::sqrt(2.0);
const double Mseries::kPi2 = 2.0 * (4.0 * std::atan(1.0));

// f(i) for i in [0, n) on up to 16 host threads (the (z, t) streams are
// independent, so each thread runs whole streams in their own order)
template <typename F> void parallel_for(int n, F f) {
  const int nt = std::max(1, std::min({n, 16, (int)std::thread::hardware_concurrency()}));
  if (nt == 1) {
    for (int i = 0; i < n; ++i) f(i);
    return;
  }
  std::vector<std::thread> th;
  for (int t = 0; t < nt; ++t)
    th.emplace_back([&, t]() {
      for (int i = t; i < n; i += nt) f(i);
    });
  for (auto& x : th) x.join();
}

}  // namespace

struct dwf_rng_s {
  int N[4] = {0, 0, 0, 0};
  std::unique_ptr<Mseries> master;
  std::vector<std::unique_ptr<Mseries>> slice;   // (z, t) streams, z fastest

  void init(unsigned long seed) {
    const int Nz = N[2], Nt = N[3];
    master.reset(new Mseries((int)seed));
    slice.clear();
    slice.resize((size_t)Nz * Nt);
    parallel_for(Nz * Nt, [&](int zt) {
      const unsigned long s = (seed + 1) * (unsigned long)Nz * Nt + (unsigned long)zt;
      auto& g = slice[zt];
      g.reset(new Mseries((int)s));
      for (int i = 0; i < 1000; ++i) g->get();
    });
  }
};

extern "C" {

int dwf_rng_create(const int lattice_size[4], unsigned long seed, dwf_rng* out) {
  if (!lattice_size || !out) return DWF_ERR_INVALID_ARG;
  auto* r = new dwf_rng_s;
  for (int mu = 0; mu < 4; ++mu) {
    if (lattice_size[mu] < 1) {
      delete r;
      return DWF_ERR_INVALID_ARG;
    }
    r->N[mu] = lattice_size[mu];
  }
  r->init(seed);
  *out = r;
  return DWF_OK;
}

int dwf_rng_reset(dwf_rng r, unsigned long seed) {
  if (!r) return DWF_ERR_INVALID_ARG;
  r->init(seed);
  return DWF_OK;
}

int dwf_rng_destroy(dwf_rng r) {
  delete r;
  return DWF_OK;
}

int dwf_rng_get(dwf_rng r, double* value) {
  if (!r || !value) return DWF_ERR_INVALID_ARG;
  *value = r->master->get();
  return DWF_OK;
}

int dwf_rng_gauss_lex_global(dwf_rng r, int nin, int nex, double* out) {
  if (!r || !out || nin < 2 || (nin & 1) || nex < 1) return DWF_ERR_INVALID_ARG;
  const int Nx = r->N[0], Ny = r->N[1], Nz = r->N[2], Nt = r->N[3];
  const size_t V = (size_t)Nx * Ny * Nz * Nt;
  // each (z, t) stream draws, in its own order, extra index j outermost
  parallel_for(Nz * Nt, [&](int zt) {
    Mseries& g = *r->slice[zt];
    for (int j = 0; j < nex; ++j) {
      size_t site = (size_t)Nx * Ny * zt;
      for (int yx = 0; yx < Nx * Ny; ++yx, ++site) {
        double* o = out + (size_t)nin * (site + V * j);
        for (int i = 0; i < nin; i += 2) g.gauss(o[i], o[i + 1]);
      }
    }
  });
  return DWF_OK;
}

// The numbers of gauss_lex_global, stored only for the sites of one parity in
// the even-odd order site2 = x/2 + Nx/2 (y + Ny (z + Nz t)) (gauss_lex_global
// followed by Index_eo::convertField(., ., parity), src/lib/Field/index_eo.cpp:
// 125-156): out[j][site2][i]; every site still draws its numbers.
int dwf_rng_gauss_eo_global(dwf_rng r, int nin, int nex, int parity, double* out) {
  if (!r || !out || nin < 2 || (nin & 1) || nex < 1 || (parity != 0 && parity != 1) ||
      (r->N[0] & 1))
    return DWF_ERR_INVALID_ARG;
  const int Nx = r->N[0], Ny = r->N[1], Nz = r->N[2], Nt = r->N[3];
  const size_t Vh = (size_t)Nx * Ny * Nz * Nt / 2;
  parallel_for(Nz * Nt, [&](int zt) {
    Mseries& g = *r->slice[zt];
    const int z = zt % Nz, t = zt / Nz;
    double skip[2];
    for (int j = 0; j < nex; ++j)
      for (int y = 0; y < Ny; ++y)
        for (int x = 0; x < Nx; ++x) {
          if (((x + y + z + t) & 1) == parity) {
            const size_t site2 = x / 2 + (size_t)(Nx / 2) * (y + (size_t)Ny * zt);
            double* o = out + (size_t)nin * (site2 + Vh * j);
            for (int i = 0; i < nin; i += 2) g.gauss(o[i], o[i + 1]);
          } else {
            for (int i = 0; i < nin; i += 2) g.gauss(skip[0], skip[1]);
          }
        }
  });
  return DWF_OK;
}

}  // extern "C"
