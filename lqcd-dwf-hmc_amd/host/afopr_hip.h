// afopr_hip.h -- C++ host mirror of the reference's operator/solver interface
// over the C ABI of libdwfhip (include/dwf_hip.h).
//
// The classes keep the names, arguments and call pattern of the reference's
//   AFopr<AFIELD>        src/lib/Fopr/afopr.h:47-277
//   ASolver<AFIELD>      src/lib_alt/Solver/asolver.h:22-79 (ASolver_CG, ASolver_Richardson)
//   AShiftsolver_CG      src/lib/Solver/ashiftsolver.h:26-54
//   Director_alt_Smear   src/lib_alt/Smear/director_alt_Smear.h
// so that code written against them (benchmark_hmc_dwf/hmc_dwf_eo.cpp and the
// actions/forces it builds) reads the same.  Errors throw dwfhip::Error carrying
// the library status and message; the reference aborts instead.
//
// Header only; link with -ldwfhip.
#pragma once

#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "dwf_hip.h"

namespace dwfhip {

class Error : public std::runtime_error {
 public:
  Error(const std::string& where, int code)
      : std::runtime_error(where + " failed (status " + std::to_string(code) + "): " +
                           dwf_last_error()),
        code_(code) {}
  int code() const { return code_; }

 private:
  int code_;
};

inline void check(int rc, const char* where) {
  if (rc != DWF_OK) throw Error(where, rc);
}

// Device-resident 5D even-odd field of one operator (the AFIELD of the
// reference: owns its buffer, RAII).
class AField {
 public:
  AField() = default;
  explicit AField(dwf_fopr op) : op_(op) { check(dwf_field_alloc(op, &p_), "dwf_field_alloc"); }
  ~AField() { reset(); }
  AField(const AField&) = delete;
  AField& operator=(const AField&) = delete;
  AField(AField&& o) noexcept : op_(o.op_), p_(o.p_) { o.p_ = nullptr; }
  AField& operator=(AField&& o) noexcept {
    if (this != &o) {
      reset();
      op_ = o.op_;
      p_ = o.p_;
      o.p_ = nullptr;
    }
    return *this;
  }
  void reset() {
    if (p_) dwf_field_free(p_);
    p_ = nullptr;
  }
  void* ptr() { return p_; }
  const void* ptr() const { return p_; }
  dwf_fopr fopr() const { return op_; }

  // BLAS-1 of AField (reductions in double)
  // AField::set(a): every element equals a (afield.h:158)
  void set(double a) { check(dwf_field_set(op_, p_, a), "dwf_field_set"); }
  double norm2() const {
    double r;
    check(dwf_field_norm2(op_, p_, &r), "dwf_field_norm2");
    return r;
  }

 private:
  dwf_fopr op_ = nullptr;
  void* p_ = nullptr;
};

inline void copy(AField& y, const AField& x) {
  check(dwf_field_copy(y.fopr(), y.ptr(), x.ptr()), "dwf_field_copy");
}
inline void axpy(AField& y, double a, const AField& x) {
  check(dwf_field_axpy(y.fopr(), y.ptr(), a, x.ptr()), "dwf_field_axpy");
}
inline void aypx(double a, AField& y, const AField& x) {
  check(dwf_field_aypx(y.fopr(), a, y.ptr(), x.ptr()), "dwf_field_aypx");
}
inline void scal(AField& x, double a) {
  check(dwf_field_scal(x.fopr(), x.ptr(), a), "dwf_field_scal");
}
inline double dot(const AField& x, const AField& y) {
  double r;
  check(dwf_field_dot(x.fopr(), x.ptr(), y.ptr(), &r), "dwf_field_dot");
  return r;
}

// Parameters with the YAML key names of the reference (HMC_DWF_Nf2p1_eo.yaml).
struct Params {
  std::vector<int> lattice_size{8, 8, 8, 8};
  int extent_of_5th_dimension = 8;
  double quark_mass = 0.02;
  double domain_wall_height = 1.6;
  double coefficient_b = 1.0;
  double coefficient_c = 0.0;
  std::vector<int> boundary_condition{1, 1, 1, -1};
  int precision = DWF_F64;
};

// AFopr<AFIELD> for fermion_type "Domainwall_5din_eo".
class AFopr_Domainwall_5din_eo {
 public:
  explicit AFopr_Domainwall_5din_eo(const Params& p) {
    dwf_fopr_params q{};
    for (int mu = 0; mu < 4; ++mu) {
      q.lattice_size[mu] = p.lattice_size.at(mu);
      q.boundary_condition[mu] = p.boundary_condition.at(mu);
    }
    q.extent_of_5th_dimension = p.extent_of_5th_dimension;
    q.quark_mass = p.quark_mass;
    q.domain_wall_height = p.domain_wall_height;
    q.b_uniform = p.coefficient_b;
    q.c_uniform = p.coefficient_c;
    q.parameter_alpha = 1.0;
    q.precision = p.precision;
    check(dwf_fopr_create(&q, &h_), "dwf_fopr_create");
  }
  ~AFopr_Domainwall_5din_eo() { dwf_fopr_destroy(h_); }
  AFopr_Domainwall_5din_eo(const AFopr_Domainwall_5din_eo&) = delete;
  AFopr_Domainwall_5din_eo& operator=(const AFopr_Domainwall_5din_eo&) = delete;

  dwf_fopr handle() const { return h_; }
  AField new_field() const { return AField(h_); }

  // U: corelib Field_G (18, V, 4) in host memory
  void set_config(const double* U) { check(dwf_fopr_set_config(h_, U), "dwf_fopr_set_config"); }
  void set_mode(const std::string& mode) {
    check(dwf_fopr_set_mode(h_, mode.c_str()), "dwf_fopr_set_mode");
  }
  std::string get_mode() const { return dwf_fopr_get_mode(h_); }
  void mult(AField& v, const AField& w) { check(dwf_fopr_mult(h_, v.ptr(), w.ptr()), "mult"); }
  void mult_dag(AField& v, const AField& w) {
    check(dwf_fopr_mult_dag(h_, v.ptr(), w.ptr()), "mult_dag");
  }
  void mult(AField& v, const AField& w, const std::string& mode) {
    check(dwf_fopr_mult_mode(h_, v.ptr(), w.ptr(), mode.c_str()), "mult(mode)");
  }
  void mult_dag(AField& v, const AField& w, const std::string& mode) {
    check(dwf_fopr_mult_dag_mode(h_, v.ptr(), w.ptr(), mode.c_str()), "mult_dag(mode)");
  }
  void mult_gm5(AField& v, const AField& w) {
    check(dwf_fopr_mult_gm5(h_, v.ptr(), w.ptr()), "mult_gm5");
  }
  int field_nin() const { return dwf_fopr_field_nin(h_); }
  int field_nvol() const { return dwf_fopr_field_nvol(h_); }
  int field_nex() const { return dwf_fopr_field_nex(h_); }
  double flop_count(const std::string& mode) const {
    return dwf_fopr_flop_count(h_, mode.c_str());
  }
  bool needs_convert() const { return true; }
  // corelib Field(24, V/2, Ls) host array <-> device field
  void convert(AField& v, const double* w) {
    check(dwf_fopr_convert(h_, v.ptr(), w), "convert");
  }
  void reverse(double* v, const AField& w) {
    check(dwf_fopr_reverse(h_, v, w.ptr()), "reverse");
  }

 private:
  dwf_fopr h_ = nullptr;
};

enum class InitialGuess { RHS = DWF_GUESS_RHS, GIVEN = DWF_GUESS_GIVEN, ZERO = DWF_GUESS_ZERO };

// ASolver<AFIELD>::solve(x, b, nconv, diff): nconv = -1 when not converged
// (the reference logs the failure and continues; so does this class).
class ASolver {
 public:
  virtual ~ASolver() {
    if (h_) dwf_solver_destroy(h_);
  }
  void solve(AField& x, const AField& b, int& nconv, double& diff) {
    const int rc = dwf_solver_solve(h_, x.ptr(), b.ptr(), &nconv, &diff);
    if (rc != DWF_OK && rc != DWF_ERR_NOT_CONVERGED) throw Error("solve", rc);
  }

 protected:
  dwf_solver h_ = nullptr;
};

class ASolver_CG : public ASolver {
 public:
  ASolver_CG(AFopr_Domainwall_5din_eo& op, int maximum_number_of_iteration,
             double convergence_criterion_squared, InitialGuess g = InitialGuess::RHS) {
    check(dwf_solver_create_cg(op.handle(), maximum_number_of_iteration,
                               convergence_criterion_squared, static_cast<int>(g), &h_),
          "dwf_solver_create_cg");
  }
};

// ASolver_Richardson with APrecond_Mixedprec (inner ASolver_CG from a zero guess),
// as Fprop_alt_Standard_eo_Richardson builds it.
class ASolver_Richardson : public ASolver {
 public:
  ASolver_Richardson(AFopr_Domainwall_5din_eo& op_d, AFopr_Domainwall_5din_eo& op_f,
                     int maximum_number_of_iteration, double convergence_criterion_squared,
                     int inner_iteration = 100000, double inner_criterion_squared = 1.0e-8) {
    check(dwf_solver_create_richardson(op_d.handle(), op_f.handle(),
                                       maximum_number_of_iteration,
                                       convergence_criterion_squared, inner_iteration,
                                       inner_criterion_squared, &h_),
          "dwf_solver_create_richardson");
  }
};

class AShiftsolver_CG {
 public:
  AShiftsolver_CG(AFopr_Domainwall_5din_eo& op, int maximum_number_of_iteration,
                  double convergence_criterion_squared)
      : op_(op), niter_(maximum_number_of_iteration), stop_(convergence_criterion_squared) {}
  void solve(std::vector<AField>& x, const std::vector<double>& sigma, const AField& b,
             int& nconv, double& diff) {
    std::vector<void*> px;
    for (auto& f : x) px.push_back(f.ptr());
    check(dwf_shiftsolver_cg_solve(op_.handle(), niter_, stop_, (int)sigma.size(),
                                   sigma.data(), px.data(), b.ptr(), &nconv, &diff),
          "dwf_shiftsolver_cg_solve");
  }

 private:
  AFopr_Domainwall_5din_eo& op_;
  int niter_;
  double stop_;
};

// Device force field in the Field_G (18, V, 4) lexical layout (FP64).
class AForceField {
 public:
  explicit AForceField(const std::vector<int>& lattice_size) {
    bytes_ = sizeof(double) * 18 * 4;
    for (int v : lattice_size) bytes_ *= (size_t)v;
    check(dwf_device_alloc(bytes_, &p_), "dwf_device_alloc");
  }
  ~AForceField() { dwf_device_free(p_); }
  AForceField(const AForceField&) = delete;
  AForceField& operator=(const AForceField&) = delete;
  double* ptr() { return static_cast<double*>(p_); }
  size_t bytes() const { return bytes_; }
  void to_host(double* F) const { check(dwf_memcpy_to_host(F, p_, bytes_), "dwf_memcpy_to_host"); }

 private:
  void* p_ = nullptr;
  size_t bytes_ = 0;
};

// AForce_F<AFIELD> of the even-odd DWF operator
// (src/lib/Force/Fermion/aforce_F.h:55-110,
//  src/lib_alt_QXS/Force/Fermion/aforce_F_Domainwall_5din_eo-tmpl.h:299-451).
class AForce_F_Domainwall_5din_eo {
 public:
  explicit AForce_F_Domainwall_5din_eo(AFopr_Domainwall_5din_eo& op) {
    check(dwf_force_create(op.handle(), &h_), "dwf_force_create");
  }
  ~AForce_F_Domainwall_5din_eo() { dwf_force_destroy(h_); }
  AForce_F_Domainwall_5din_eo(const AForce_F_Domainwall_5din_eo&) = delete;
  AForce_F_Domainwall_5din_eo& operator=(const AForce_F_Domainwall_5din_eo&) = delete;
  void set_mode(const std::string& mode) {
    check(dwf_force_set_mode(h_, mode.c_str()), "dwf_force_set_mode");
  }
  void force_udiv(AForceField& F, const AField& eta) {
    check(dwf_force_udiv(h_, F.ptr(), eta.ptr()), "dwf_force_udiv");
  }
  void force_udiv1(AForceField& F, const AField& zeta, const AField& eta) {
    check(dwf_force_udiv1(h_, F.ptr(), zeta.ptr(), eta.ptr()), "dwf_force_udiv1");
  }
  void force_core(AForceField& F, const AField& eta) {
    check(dwf_force_core(h_, F.ptr(), eta.ptr()), "dwf_force_core");
  }
  void force_core1(AForceField& F, const AField& zeta, const AField& eta) {
    check(dwf_force_core1(h_, F.ptr(), zeta.ptr(), eta.ptr()), "dwf_force_core1");
  }

 private:
  dwf_force h_ = nullptr;
};

class Director_alt_Smear {
 public:
  Director_alt_Smear(const std::vector<int>& lattice_size, double rho_uniform,
                     int number_of_smearing) {
    int d[4] = {lattice_size.at(0), lattice_size.at(1), lattice_size.at(2), lattice_size.at(3)};
    check(dwf_smear_create(d, rho_uniform, number_of_smearing, &h_), "dwf_smear_create");
  }
  ~Director_alt_Smear() { dwf_smear_destroy(h_); }
  Director_alt_Smear(const Director_alt_Smear&) = delete;
  Director_alt_Smear& operator=(const Director_alt_Smear&) = delete;
  dwf_smear handle() const { return h_; }
  void set_config(const double* U) { check(dwf_smear_set_config(h_, U), "dwf_smear_set_config"); }
  void get_config(int level, double* U) {
    check(dwf_smear_get_config(h_, level, U), "dwf_smear_get_config");
  }

 private:
  dwf_smear h_ = nullptr;
};

// AFopr_Smeared_alt::set_config: smear, then give the top level to the kernel.
inline void set_config_smeared(AFopr_Domainwall_5din_eo& op, Director_alt_Smear& sm,
                               const double* U) {
  check(dwf_fopr_set_config_smeared(op.handle(), sm.handle(), U),
        "dwf_fopr_set_config_smeared");
}

}  // namespace dwfhip
