/*
 * dwf_hip.h -- C ABI of libdwfhip, the MI355X (gfx950) implementation of the
 * Bridge++ domain-wall HMC hot path.
 *
 * Every entry point returns an int status (DWF_OK = 0) instead of exiting; the
 * message of the last failure on the calling thread is dwf_last_error().
 * Plain pointers and sizes only.  "host" pointers are ordinary CPU memory;
 * "device" pointers are HIP device memory on the current device, e.g. from
 * dwf_field_alloc().  All device work is ordered on one library stream
 * (dwf_stream()); calls that return host values synchronise it.
 *
 * Interchange formats at the boundary are the reference's corelib layouts:
 *   gauge field  Field_G(18, V, 4): double U[mu][site][ab], ab = ri+2*(b+3*a),
 *                lexical site x + Nx*(y + Ny*(z + Nz*t))   (src/lib/Field/field.h:64-66)
 *   5D eo spinor Field(24, V/2, Ls): double w[s][site2][in], in = ri+2*(c+3*d),
 *                site2 = x/2 + Nx/2*(y + Ny*(z + Nz*t))    (src/lib/Field/index_eo.h:66)
 * The device-resident field of an operator ("AFIELD") has its own layout
 * (site-major, spin-colour, Ls innermost) and precision; only convert/reverse
 * move data between the two.
 */
#ifndef DWF_HIP_H
#define DWF_HIP_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

enum dwf_status {
  DWF_OK = 0,
  DWF_ERR_INVALID_ARG = 1,
  DWF_ERR_UNKNOWN_MODE = 2,   /* reference: vout.crucial + abort() */
  DWF_ERR_HIP = 3,
  DWF_ERR_NOT_CONVERGED = 4,  /* reference: logged, nconv = -1 */
  DWF_ERR_NO_CONFIG = 5,
  DWF_ERR_UNSUPPORTED = 6,
  DWF_ERR_NO_DEVICE = 7
};

enum dwf_precision {
  DWF_F64 = 0,   /* double storage, double arithmetic */
  DWF_F32 = 1,   /* float storage, float arithmetic */
  DWF_F16 = 2,   /* half storage (spinors and links), float arithmetic */
  DWF_BF16 = 3   /* bfloat16 storage, float arithmetic */
};

enum dwf_initial_guess {   /* src/lib_alt/Solver/asolver.h InitialGuess */
  DWF_GUESS_RHS = 0,
  DWF_GUESS_GIVEN = 1,
  DWF_GUESS_ZERO = 2
};

const char* dwf_last_error(void);
const char* dwf_version(void);
/* Library stream (hipStream_t as void*), for callers that time or order work. */
void* dwf_stream(void);
/* Block until all library work has finished. */
int dwf_synchronize(void);

/* ------------------------------------------------------------------------
 * AFopr_Domainwall_5din_eo  (src/lib/Fopr/afopr.h:47-277 interface;
 * operator of src/lib_alt_QXS/Fopr/afopr_Domainwall_5din_eo-tmpl.h and
 * src/lib/Fopr/afopr_Domainwall_eo-tmpl.h)
 * ---------------------------------------------------------------------- */
typedef struct dwf_fopr_s* dwf_fopr;

typedef struct {
  int lattice_size[4];        /* Lx Ly Lz Lt (one GPU holds the lattice)        */
  int extent_of_5th_dimension;/* Ls                                             */
  double quark_mass;          /* mq                                             */
  double domain_wall_height;  /* M0                                             */
  const double* coefficient_b;/* Ls values, or NULL to use b_uniform            */
  const double* coefficient_c;/* Ls values, or NULL to use c_uniform            */
  double b_uniform, c_uniform;/* coefficient_b / coefficient_c of the YAML      */
  double parameter_alpha;     /* 1.0 unless the Accel alpha form is wanted      */
  int boundary_condition[4];  /* +1 periodic, -1 antiperiodic, 0 Dirichlet      */
  int precision;              /* enum dwf_precision                             */
} dwf_fopr_params;

/* AFopr::New("Domainwall_5din_eo", params) (afopr_Domainwall_5din_eo.h:214-221) */
int dwf_fopr_create(const dwf_fopr_params* params, dwf_fopr* out);
int dwf_fopr_destroy(dwf_fopr op);
/* set_config(Field*) (afopr_Domainwall_5din_eo-tmpl.h:367-424): U host Field_G. */
int dwf_fopr_set_config(dwf_fopr op, const double* U_host);
/* Same, U already in device memory in the Field_G layout. */
int dwf_fopr_set_config_device(dwf_fopr op, const double* U_device);
/* set_mode/get_mode: D, Ddag, DdagD, DDdag, H, Hdag, Deo, Doe, Dee, Doo,
 * Dee_inv, Doo_inv (afopr_Domainwall_eo-tmpl.h:384-431). */
int dwf_fopr_set_mode(dwf_fopr op, const char* mode);
const char* dwf_fopr_get_mode(dwf_fopr op);
/* mult(v, w) / mult_dag(v, w) in the current mode; v, w device fields. */
int dwf_fopr_mult(dwf_fopr op, void* v, const void* w);
int dwf_fopr_mult_dag(dwf_fopr op, void* v, const void* w);
/* mult(v, w) in the current mode without any allocation.  No counterpart in
 * the reference: for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops).  The intermediate fields live in a
 * static per-GPU scratch array of 256 MiB (5 fields of the operator, each
 * rounded up to 256 bytes, must fit: e.g. FP16 storage at 16^3 x 32, Ls = 16);
 * larger operators are rejected with DWF_ERR_INVALID_ARG.  Every call on one
 * GPU shares that array: calls that overlap on one GPU race.  The library must
 * be initialised (an operator created) on the current device. */
int dwf_fopr_mult_noalloc(dwf_fopr op, void* v, const void* w);
/* mult(v, w, mode) / mult_dag(v, w, mode) (afopr_Domainwall_eo-tmpl.h:475-525) */
int dwf_fopr_mult_mode(dwf_fopr op, void* v, const void* w, const char* mode);
int dwf_fopr_mult_dag_mode(dwf_fopr op, void* v, const void* w, const char* mode);
/* mult_gm5 (gamma_5 on every s) and mult_gm5R (with s -> Ls-1-s reflection) */
int dwf_fopr_mult_gm5(dwf_fopr op, void* v, const void* w);
int dwf_fopr_mult_gm5R(dwf_fopr op, void* v, const void* w);
/* field_nin() = 24*Ls, field_nvol() = V/2, field_nex() = 1 */
int dwf_fopr_field_nin(dwf_fopr op);
int dwf_fopr_field_nvol(dwf_fopr op);
int dwf_fopr_field_nex(dwf_fopr op);
/* bytes of one device field of this operator */
size_t dwf_fopr_field_bytes(dwf_fopr op);
int dwf_fopr_precision(dwf_fopr op);
/* flop_count(mode): the reference's flop model
 * (afopr_Domainwall_5din_eo-tmpl.h:1196-1250). */
double dwf_fopr_flop_count(dwf_fopr op, const char* mode);
/* convert(AFIELD& v, const Field& w) / reverse(Field& v, const AFIELD& w)
 * (afopr_Domainwall_5din_eo-tmpl.h:478-567): corelib Field(24, V/2, Ls) double
 * host array <-> device field of this operator. */
int dwf_fopr_convert(dwf_fopr op, void* v_device, const double* w_host);
int dwf_fopr_reverse(dwf_fopr op, double* v_host, const void* w_device);
/* Same with the corelib array in device memory. */
int dwf_fopr_convert_device(dwf_fopr op, void* v_device, const double* w_corelib_device);
int dwf_fopr_reverse_device(dwf_fopr op, double* v_corelib_device, const void* w_device);

/* ------------------------------------------------------------------------
 * AField BLAS-1 on device fields of an operator (src/lib_alt_Accel/Field/
 * afield-tmpl.h:637-742; reductions in double).
 * ---------------------------------------------------------------------- */
int dwf_field_alloc(dwf_fopr op, void** field);
int dwf_field_free(void* field);
int dwf_field_set_zero(dwf_fopr op, void* x);
/* AField::set(a): every stored real of x equals a
 * (src/lib_alt_Accel/Field/afield.h:158, afield-tmpl.h:182-193). */
int dwf_field_set(dwf_fopr op, void* x, double a);
int dwf_field_copy(dwf_fopr op, void* y, const void* x);
int dwf_field_axpy(dwf_fopr op, void* y, double a, const void* x);   /* y += a x   */
int dwf_field_aypx(dwf_fopr op, double a, void* y, const void* x);   /* y = a y + x */
int dwf_field_scal(dwf_fopr op, void* x, double a);
int dwf_field_norm2(dwf_fopr op, const void* x, double* result);
int dwf_field_dot(dwf_fopr op, const void* x, const void* y, double* result); /* Re<x,y> */
int dwf_field_dotc(dwf_fopr op, const void* x, const void* y, double* re, double* im);
/* convert_h: field of op_from -> field of op_to (precision cast;
 * aprecond_Mixedprec-tmpl.h:55-108).  Both operators on the same lattice. */
int dwf_field_convert_precision(dwf_fopr op_to, void* v, dwf_fopr op_from, const void* w);

/* ------------------------------------------------------------------------
 * Solvers.  ASolver<AFIELD> (src/lib_alt/Solver/asolver.h:22-79).
 * ---------------------------------------------------------------------- */
typedef struct dwf_solver_s* dwf_solver;

/* ASolver_CG (asolver_CG-tmpl.h:118-234) on op->mult in op's current mode.
 * Stops on |r|^2/|b|^2 < stop_cond (convergence_criterion_squared). */
int dwf_solver_create_cg(dwf_fopr op, int max_iter, double stop_cond,
                         int initial_guess, dwf_solver* out);
/* ASolver_Richardson (asolver_Richardson-tmpl.h:131-213) on op_d with
 * APrecond_Mixedprec (aprecond_Mixedprec-tmpl.h:55-108) running ASolver_CG on
 * op_f from a zero guess to stop_cond_inner, as Fprop_alt_Standard_eo_Richardson
 * sets it up (fprop_alt_Standard_eo_Richardson-tmpl.h:27-102, inner 1e-8). */
int dwf_solver_create_richardson(dwf_fopr op_d, dwf_fopr op_f, int max_iter,
                                 double stop_cond, int max_iter_inner,
                                 double stop_cond_inner, dwf_solver* out);
/* Schedule autotuner (no counterpart in the reference; for callers that choose
 * the mixed-precision scheme per lattice): for each inner operator inner[i]
 * (same lattice and parameters as op_d, any storage type) and inner criterion
 * tols[j], one untimed and one timed solve of op_d's system with right-hand
 * side b (x: a field of op_d, overwritten); times[i * n_tol + j] = seconds of
 * the timed solve, -1 if it did not reach stop_cond.  *out = the Richardson
 * solver of the fastest candidate, *best = its index. */
int dwf_solver_tune_richardson(dwf_fopr op_d, int n_inner, const dwf_fopr* inner, int n_tol,
                               const double* tols, int max_iter, double stop_cond,
                               int max_iter_inner, const void* b, void* x, double* times,
                               int* best, dwf_solver* out);
/* solve(x, b, nconv, diff): diff = |A x - b| / |b| recomputed in the solver's
 * outer precision.  Returns DWF_ERR_NOT_CONVERGED (nconv = -1) when the
 * iteration limit is hit; x then holds the last iterate. */
int dwf_solver_solve(dwf_solver s, void* x, const void* b, int* nconv, double* diff);
/* Iterations of the inner solver summed over the last solve (0 for CG). */
int dwf_solver_inner_iterations(dwf_solver s);
int dwf_solver_destroy(dwf_solver s);

/* AShiftsolver_CG::solve (src/lib/Solver/ashiftsolver_CG-tmpl.h:89-329):
 * multishift CG on (A + sigma_i), A = op->mult in op's current mode, with the
 * reference's shift truncation.  x: nshift device fields.  nconv follows the
 * reference's count; diff = max_i |(A + sigma_i) x_i - b| / |b|. */
int dwf_shiftsolver_cg_solve(dwf_fopr op, int max_iter, double stop_cond,
                             int nshift, const double* sigma, void** x,
                             const void* b, int* nconv, double* diff);

/* Batched multishift CG: nsys right-hand sides ("systems") solved together
 * for the same shifts, each with the recurrences of AShiftsolver_CG::solve
 * (ashiftsolver_CG-tmpl.h:89-286) -- e.g. the three shift solves of one
 * AForce_F_Rational_Ratio::force_udiv (aforce_F_Rational_Ratio-tmpl.h:371,396,
 * 420) or several pseudofermions, and independent lattices.  ops[k] is the
 * operator of system k (mode DdagD; FP64 or FP32; all with the same precision,
 * Ls and b, c; at most 4 different operators, which then need Ls = 12 or 16).
 * Systems on one operator share every hop launch (links read once for all of
 * them); different operators share grouped launches.  Scalars stay on the
 * device; DWF_MSBATCH_GRAPH replays the iteration from a captured HIP graph.
 * x: nsys * nshift fields of the system operators, x[k * nshift + j] solving
 * (A_k + sigma_j) x = b[k]; nconv[k] as the reference reports it (-1 if not
 * converged), diff[k] = max_j |(A_k + sigma_j) x_kj - b_k| / |b_k|. */
typedef struct dwf_msbatch_s* dwf_msbatch;
enum { DWF_MSBATCH_GRAPH = 1 };
int dwf_msbatch_create(int nsys, const dwf_fopr* ops, int nshift, int flags, dwf_msbatch* out);
int dwf_msbatch_solve(dwf_msbatch m, const double* sigma, void** x, const void* const* b,
                      int max_iter, double stop_cond, int* nconv, double* diff);
/* iterations of the last solve (the largest over the systems) */
int dwf_msbatch_iterations(dwf_msbatch m);
int dwf_msbatch_destroy(dwf_msbatch m);
/* Mixed precision (m created on FP32 operators, ops64[k] the FP64 operator of
 * system k on the same lattice, mode DdagD): the batched multishift in FP32 to
 * stop_cond_low, then every (system, shift) refined in FP64 with the
 * Richardson / Mixedprec scheme (asolver_Richardson-tmpl.h:131-213,
 * aprecond_Mixedprec-tmpl.h:55-108) on the shifted operator: the residual
 * b - (A + sigma_j) x in FP64, the correction by CG on the FP32 operator
 * (A + sigma_j) to stop_cond_inner, until |r|^2/|b|^2 < stop_cond.  x and b
 * are fields of ops64; nconv[k] = iterations of the FP32 multishift;
 * diff[k] = max_j |(A + sigma_j) x_kj - b_k| / |b_k| in FP64. */
int dwf_msbatch_solve_mixed(dwf_msbatch m, const dwf_fopr* ops64, const double* sigma, void** x,
                            const void* const* b, int max_iter, double stop_cond,
                            double stop_cond_low, double stop_cond_inner, int* nconv,
                            double* diff);
/* CG iterations of the last mixed solve's refinement, summed */
int dwf_msbatch_refine_iterations(dwf_msbatch m);

/* ------------------------------------------------------------------------
 * Learned polynomial preconditioner (no counterpart in the reference; it is a
 * new solver component for the same operator).
 *   K = sum_{k=0}^{L} W_k H^k,  H = 1 - D,  D the even-odd Schur operator of op,
 * W_k per-chirality Ls x Ls real matrices (weights), 1 <= L <= 4, Ls <= 16.
 * K approximates D^-1 and M = K K^+ approximates (D^+ D)^-1 (symmetric by
 * construction).  dwf_lpc_train fits the weights on the current gauge field
 * by least squares, min sum_j |K D v_j - v_j|^2 over nvec random fields v_j
 * (uniform in [-1, 1), counter hash of seed), with ridge * mean(diag) added to
 * the normal matrix; loss = the relative residual of the fit.  Training runs
 * with opt (any precision, same lattice and parameters); the preconditioner
 * is applied in op's precision (FP16/BF16 storage is meant for it).
 * Weights layout: [k][chi][s_out][s_in], chi = 0 for P+, 1 for P-; a new
 * preconditioner starts from W_k = 1 (the truncated Neumann series of D^-1).
 * ---------------------------------------------------------------------- */
typedef struct dwf_lpc_s* dwf_lpc;

int dwf_lpc_create(dwf_fopr op, int degree, dwf_lpc* out);
int dwf_lpc_destroy(dwf_lpc p);
int dwf_lpc_degree(dwf_lpc p);
int dwf_lpc_set_weights(dwf_lpc p, const double* W);
int dwf_lpc_get_weights(dwf_lpc p, double* W);
int dwf_lpc_train(dwf_lpc p, dwf_fopr opt, int nvec, unsigned long long seed, double ridge,
                  double* loss);
/* which: "K", "Kdag" or "M" (= K K^+); v != w, fields of p's operator */
int dwf_lpc_apply(dwf_lpc p, const char* which, void* v, const void* w);
/* chiral Gram matrices of nfields <= 6 fields of op (all pairs i <= j, row-major
 * over i): out[p][chi][s][t] = sum_{site, k<12} p_chi(f_i)(s) p_chi(f_j)(t),
 * p_+ = f_k + f_{k+12}, p_- = f_k - f_{k+12} (host array of nfields(nfields+1)/2 * 2 Ls^2) */
int dwf_lpc_chiral_gram(dwf_fopr op, int nfields, const void* const* fields, double* out_host);
/* x_i = uniform [-1, 1) from the counter hash of (seed, i), i the device index */
int dwf_field_fill_uniform(dwf_fopr op, void* x, unsigned long long seed);

/* Flexible CG on op (mode DdagD) preconditioned by M = K K^+, zero initial
 * guess; stop when |r|^2 / |b|^2 < stop_cond.  The preconditioner may use a
 * lower precision than op. */
int dwf_solver_create_fcg(dwf_fopr op, dwf_lpc prec, int max_iter, double stop_cond,
                          dwf_solver* out);
/* ASolver_Richardson (FP64 op_d) whose inner solver is the flexible CG on op_f
 * with the preconditioner: the mixed-precision solve of Fprop_alt_Standard_eo_Richardson
 * with the inner CG replaced. */
int dwf_solver_create_richardson_lpc(dwf_fopr op_d, dwf_fopr op_f, dwf_lpc prec, int max_iter,
                                     double stop_cond, int max_iter_inner,
                                     double stop_cond_inner, dwf_solver* out);

/* ------------------------------------------------------------------------
 * Learned preconditioner: a gauge-covariant parallel-transport convolution
 * network (no counterpart in the reference; a new solver component for the
 * same operator).  Layer l maps channels[l] to channels[l + 1] 5D spinor
 * fields on the full lattice:
 *   out[a](x, s) = sum_{q < 11} sum_b W_l[q][a][b] T_q(in[b])(x, s),
 * T_0 = 1, T_{1+dir} = the 8 gauge hops of the operator (U(x, dir) in(x + dir),
 * boundary phases included), T_9 / T_10 = the shifts s -> s + 1 / s - 1 (zero
 * beyond the walls); W_l[q][a][b] a complex 4 x 4 spin matrix, stored as
 * bfloat16 (round to nearest even), applied in FP32.  The input is an even
 * field (odd parity zero), the output the even part of the last layer.
 * channels[0] = channels[nlayers] = 1, others <= 16.  op: an FP32 operator
 * whose links (after set_config) the network transports with.
 * Weights (host float / double): per layer [q][c_out][c_in][16 x (re, im)],
 * row-major spin entries, layers one after another (dwf_ptc_num_weights).
 * ---------------------------------------------------------------------- */
typedef struct dwf_ptc_s* dwf_ptc;
int dwf_ptc_create(dwf_fopr op, int nlayers, const int* channels, dwf_ptc* out);
int dwf_ptc_destroy(dwf_ptc n);
size_t dwf_ptc_num_weights(dwf_ptc n);
/* W rounded to bfloat16; get returns the rounded values */
int dwf_ptc_set_weights(dwf_ptc n, const float* W);
int dwf_ptc_get_weights(dwf_ptc n, float* W);
/* out = N(in), FP32 fields of op */
int dwf_ptc_apply(dwf_ptc n, void* out, const void* in);
/* loss = sum_k |N(b[k]) - x[k]|^2 over nbatch pairs of FP32 fields of op and
 * grad = d loss / d W (host array of dwf_ptc_num_weights doubles, d/d re and
 * d/d im of every entry), by the network's forward and adjoint kernels. */
int dwf_ptc_loss_grad(dwf_ptc n, int nbatch, const void* const* b, const void* const* x,
                      double* loss, double* grad);
/* Flexible CG on op (mode DdagD) preconditioned by the network, and the FP64
 * Richardson around it (the inner flexible CG on op_f), as for the polynomial. */
int dwf_solver_create_fcg_ptc(dwf_fopr op, dwf_ptc prec, int max_iter, double stop_cond,
                              dwf_solver* out);
int dwf_solver_create_richardson_ptc(dwf_fopr op_d, dwf_fopr op_f, dwf_ptc prec, int max_iter,
                                     double stop_cond, int max_iter_inner,
                                     double stop_cond_inner, dwf_solver* out);

/* ------------------------------------------------------------------------
 * DWF fermion force, FP64: AForce_F_Domainwall_5din_eo
 * (src/lib_alt_QXS/Force/Fermion/aforce_F_Domainwall_5din_eo-tmpl.h:299-451,
 *  lexical part aforce_F_Domainwall_5din-tmpl.h:338-933, tensorProd_5din
 *  afield_Spinor-inc.h:322-370, mult_generator aforce_F.cpp:97-139).
 * Force fields are device arrays in the Field_G (18, V, 4) lexical layout.
 * ---------------------------------------------------------------------- */
typedef struct dwf_force_s* dwf_force;

int dwf_force_create(dwf_fopr op, dwf_force* out);
int dwf_force_destroy(dwf_force f);
/* set_mode: D, Ddag, H or Hdag (the mode of force_udiv1) */
int dwf_force_set_mode(dwf_force f, const char* mode);
/* force_udiv1(F, zeta, eta) in the current mode; F overwritten */
int dwf_force_udiv1(dwf_force f, double* F, const void* zeta, const void* eta);
/* force_udiv(F, eta) = udiv1_D(D eta, eta) + udiv1_Ddag(eta, D eta) */
int dwf_force_udiv(dwf_force f, double* F, const void* eta);
/* mult_generator: F <- -2 at(U F) */
int dwf_force_mult_generator(dwf_force f, double* F);
/* force_core = mult_generator(force_udiv), force_core1 = mult_generator(force_udiv1) */
int dwf_force_core(dwf_force f, double* F, const void* eta);
int dwf_force_core1(dwf_force f, double* F, const void* zeta, const void* eta);

/* ------------------------------------------------------------------------
 * Rational functions of A^dag A, FP64.  One rational function is
 *   r(x) = a0 + sum_i bl[i] / (x + cl[i]),  i < np,
 * with the coefficients of a Math_Rational file (src/lib/Tools/math_Rational.cpp:
 * a0 = norm, bl = res, cl = pole), evaluated by multishift CG on mode DdagD.
 * ---------------------------------------------------------------------- */
typedef struct {
  int np;               /* number_of_poles */
  double a0;
  const double* bl;     /* np residues */
  const double* cl;     /* np poles (shifts; cl[0] is the base system) */
  int max_iter;         /* maximum_number_of_iteration of the shift solver */
  double stop_cond;     /* convergence_criterion_squared of the shift solver */
} dwf_rational_params;

/* AFopr_Rational::mult (src/lib/Fopr/afopr_Rational-tmpl.h:181-219):
 * v = r(A^dag A) b with A = op; sets op's mode to DdagD as the reference does.
 * nconv, diff: those of the shift solver.  DWF_ERR_NOT_CONVERGED leaves v
 * computed from the unconverged solutions. */
int dwf_rational_mult(dwf_fopr op, const dwf_rational_params* r, void* v, const void* b,
                      int* nconv, double* diff);

/* AForce_F_Rational_Ratio (src/lib_alt/Force/Fermion/aforce_F_Rational_Ratio-tmpl.h:
 * 21-459): the force of S = <phi| r1(A^dag A) r2(B^dag B) r1(A^dag A) |phi>,
 * A = fopr1 (numerator, r1 ~ (A^dag A)^(1/4)), B = fopr2 (denominator,
 * r2 ~ (B^dag B)^(-1/2)).  Both operators FP64 on the same lattice. */
typedef struct dwf_ratio_force_s* dwf_ratio_force;

int dwf_force_rational_ratio_create(dwf_fopr fopr1, dwf_fopr fopr2,
                                    const dwf_rational_params* numerator,
                                    const dwf_rational_params* denominator,
                                    dwf_ratio_force* out);
int dwf_force_rational_ratio_destroy(dwf_ratio_force rf);
/* force_udiv (-tmpl.h:345-447): F = sum_i bl2[i] force_udiv^B(eta2_i)
 *   + sum_i bl1[i] [udiv1_H(H eta1_i, zeta1_i) + udiv1_H(H zeta1_i, eta1_i)
 *                 + udiv1_Hdag(eta1_i, H zeta1_i) + udiv1_Hdag(zeta1_i, H eta1_i)];
 * leaves fopr1 in mode H and fopr2 in mode DdagD, as the reference does.
 * DWF_ERR_NOT_CONVERGED if a shift solve did not converge (F still computed). */
int dwf_force_rational_ratio_udiv(dwf_ratio_force rf, double* F, const void* eta);
/* force_core = mult_generator(force_udiv)   (lib_alt_QXS/Force/Fermion/aforce_F.cpp:124) */
int dwf_force_rational_ratio_core(dwf_ratio_force rf, double* F, const void* eta);

/* Plain device memory helpers (synchronous copies on the library stream). */
int dwf_device_alloc(size_t bytes, void** ptr);
int dwf_device_free(void* ptr);
int dwf_memcpy_to_device(void* dst_device, const void* src_host, size_t bytes);
int dwf_memcpy_to_host(void* dst_host, const void* src_device, size_t bytes);

/* ------------------------------------------------------------------------
 * Stout smearing: Director_alt_Smear + ASmear_APE + AProjection_Stout_SU3
 * (src/lib_alt/Smear/director_alt_Smear-tmpl.h:19-88,273-286,
 *  src/lib_alt_QXS/Smear/asmear_APE-tmpl.h:183-223,
 *  src/lib_alt_QXS/Smear/aprojection_Stout_SU3-tmpl.h:477-553)
 * ---------------------------------------------------------------------- */
typedef struct dwf_smear_s* dwf_smear;

int dwf_smear_create(const int lattice_size[4], double rho_uniform,
                     int number_of_smearing, dwf_smear* out);
int dwf_smear_destroy(dwf_smear sm);
/* Director::set_config: keeps U (host Field_G) and computes all levels. */
int dwf_smear_set_config(dwf_smear sm, const double* U_host);
int dwf_smear_set_config_device(dwf_smear sm, const double* U_device);
/* Copy smearing level i (0 = thin links) to a host Field_G. */
int dwf_smear_get_config(dwf_smear sm, int level, double* U_host);
/* Device pointer of level i (valid until the next set_config / destroy). */
const double* dwf_smear_config_device(dwf_smear sm, int level);
/* AFopr_Smeared_alt::set_config (src/lib/Fopr/afopr_Smeared_alt-tmpl.h:91-98):
 * smear U, then give the top level to the kernel operator. */
int dwf_fopr_set_config_smeared(dwf_fopr op, dwf_smear sm, const double* U_host);

/* Stout force recursion (FP64, device Field_G (18, V, 4) arrays).
 * dwf_smear_force_udiv: Director_alt_Smear::force_udiv of one level
 *   (src/lib_alt/Smear/director_alt_Smear-tmpl.h:253-270 -> AForceSmear_APE::force_udiv,
 *   src/lib_alt_QXS/Smear/aforceSmear_APE-tmpl.h:157-206, AProjection_Stout_SU3::
 *   force_recursive, aprojection_Stout_SU3-tmpl.h:691-895): Sigma = dS/dU_level from
 *   Sigmap = dS/dU_(level+1); Sigma may alias Sigmap.
 * dwf_smear_mult_jacobian: AForce_F_Smeared_alt::mult_jacobian
 *   (src/lib_alt/Force/Fermion/aforce_F_Smeared_alt-tmpl.h:170-191), in place, all levels.
 * dwf_smear_mult_generator: F <- -2 at(U_0 F) with the thin links U_0.
 * The force of an operator on smeared links (AForce_F_Smeared_alt::force_core):
 *   dwf_force_udiv(f, F, eta); dwf_smear_mult_jacobian(sm, F); dwf_smear_mult_generator(sm, F). */
int dwf_smear_force_udiv(dwf_smear sm, int level, double* Sigma, const double* Sigmap);
int dwf_smear_mult_jacobian(dwf_smear sm, double* F);
int dwf_smear_mult_generator(dwf_smear sm, double* F);

/* ------------------------------------------------------------------------
 * Random numbers of the HMC (host; no device needed):
 * RandomNumbers_parallel<RandomNumbers_Mseries>
 * (src/lib/Tools/randomNumbers_parallel-tmpl.h:146-230,
 *  randomNumbers_parallel_Mseries.cpp:29-38, randomNumbers_Mseries.cpp:29-145):
 * an M-series master generator and one per (z, t) slice of the lattice.
 * ---------------------------------------------------------------------- */
typedef struct dwf_rng_s* dwf_rng;
/* RandomNumbers_parallel(seed) on a lattice of these dimensions (one process) */
int dwf_rng_create(const int lattice_size[4], unsigned long seed, dwf_rng* out);
/* reset(seed): re-seed every generator (init_2d) */
int dwf_rng_reset(dwf_rng r, unsigned long seed);
int dwf_rng_destroy(dwf_rng r);
/* get(): one uniform number in [0, 1) from the master generator */
int dwf_rng_get(dwf_rng r, double* value);
/* gauss_lex_global(Field(nin, V, nex)): Gaussian numbers of variance 1/2 into
 * the host array out[j][site][i] (lexical sites; nin even). */
int dwf_rng_gauss_lex_global(dwf_rng r, int nin, int nex, double* out);
/* The same numbers kept for the sites of one parity only, in the even-odd order
 * site2 = x/2 + Nx/2*(y + Ny*(z + Nz*t)): out[j][site2][i] (gauss_lex_global then
 * Index_eo::convertField(., ., parity), src/lib/Field/index_eo.cpp:125-156, as the
 * pseudofermion noise of action_F_alt_Ratio_eo-tmpl.h langevin; Nx even). */
int dwf_rng_gauss_eo_global(dwf_rng r, int nin, int nex, int parity, double* out);

/* ------------------------------------------------------------------------
 * HMC gauge part, FP64, device arrays in the Field_G layout (4, V, 18).
 * ---------------------------------------------------------------------- */
/* Langevin_Momentum::set_iP (src/lib/HMC/langevin_Momentum.cpp:146-200): iP from
 * 8 Gaussian numbers per link, host array gauss[mu][site][8] (as
 * dwf_rng_gauss_lex_global(r, 8, 4, .) gives them); H_kin = |iP|^2 / 2. */
int dwf_hmc_set_iP(const int lattice_size[4], double* iP, const double* gauss_host,
                   double* H_kin);
/* Integrator_UpdateU::evolve: U <- exp(alpha iP) U, nested Taylor form of order
 * nprec and reunitarisation (mult_exp_Field_G, field_G_imp.cpp:496-520). */
int dwf_hmc_exp_update(const int lattice_size[4], double* U, double alpha, const double* iP,
                       int nprec);
/* Action_G_Rectangle::calcH (src/lib/Action/Gauge/action_G_Rectangle.cpp:113-193):
 * H = beta [c_plaq (6V - sum Re tr P / 3) + c_rect (12V - sum Re tr R / 3)];
 * plaq = the average plaquette (may be NULL). */
int dwf_hmc_rect_action(const int lattice_size[4], const double* U, double beta, double c_plaq,
                        double c_rect, double* H, double* plaq);
/* Force_G_Rectangle::force_core (src/lib/Force/Gauge/force_G_Rectangle.cpp:84-216):
 * F_mu(x) = -(beta / 3) at(U_mu(x) S_mu(x)^+), S the staple sum (F overwritten). */
int dwf_hmc_rect_force(const int lattice_size[4], double* F, const double* U, double beta,
                       double c_plaq, double c_rect);
/* y += a x, y = x, |x|^2 for device double arrays of n elements */
int dwf_device_axpy(double* y, double a, const double* x, size_t n);
int dwf_device_copy(double* y, const double* x, size_t n);
int dwf_device_norm2(const double* x, size_t n, double* result);
/* Average plaquette of a device Field_G. */
int dwf_plaquette_device(const int lattice_size[4], const double* U, double* result);

/* Gauge configuration files, "Binary" (ILDG order, big-endian; GaugeConfig
 * ("Binary"), src/lib/IO/gaugeConfig.cpp:61-107, fieldIO_Binary.cpp:29-176,
 * io_format_gauge.h:43-70): per lexical site the 4 x 18 doubles of its links.
 * U_host is a Field_G (4, V, 18).  DWF_ERR_INVALID_ARG if the file cannot be
 * opened, is short, or cannot be written.  No device needed. */
int dwf_gauge_read_binary(const char* path, const int lattice_size[4], double* U_host);
int dwf_gauge_write_binary(const char* path, const int lattice_size[4], const double* U_host);

/* Average plaquette of a host Field_G (1 for the unit field). */
int dwf_plaquette(const int lattice_size[4], const double* U_host, double* result);

#ifdef __cplusplus
}
#endif
#endif
