"""dwf_amd -- Python mirror of the reference's AFopr / ASolver / AShiftsolver /
Director_alt_Smear interfaces, over the C ABI of libdwfhip (HIP, gfx950).

Class and method names follow the reference (src/lib/Fopr/afopr.h:47-277,
src/lib_alt/Solver/asolver.h:22-79, src/lib/Solver/ashiftsolver.h:26-54,
src/lib_alt/Smear/director_alt_Smear.h); fields are device resident (AFIELD),
host data crosses the boundary only through convert/reverse in the corelib
layouts described in dwf_amd.lattice.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from ._lib import DwfError, F64, F32, F16, BF16, GUESS_RHS, GUESS_ZERO, GUESS_GIVEN

_PREC = {"double": F64, "float": F32, "half": F16, "bfloat16": BF16,
         F64: F64, F32: F32, F16: F16, BF16: BF16}
_dp = C.POINTER(C.c_double)


def _ptr(a):
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(_dp)


def stream():
    """The library's HIP stream (hipStream_t as an int)."""
    return _lib.load().dwf_stream()


def synchronize():
    _lib.check("dwf_synchronize", _lib.load().dwf_synchronize())


class Field:
    """A device-resident 5D even-odd field of one operator (AFIELD)."""

    def __init__(self, fopr):
        L = _lib.load()
        self.fopr = fopr
        p = C.c_void_p()
        _lib.check("dwf_field_alloc", L.dwf_field_alloc(fopr._h, C.byref(p)))
        self._p = p.value

    @property
    def ptr(self):
        return self._p

    def free(self):
        if getattr(self, "_p", None):
            _lib.load().dwf_field_free(self._p)
            self._p = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass

    # BLAS-1 (AField methods)
    def set_zero(self):
        _lib.check("dwf_field_set_zero", _lib.load().dwf_field_set_zero(self.fopr._h, self._p))

    def set(self, a):
        """AField::set(a): every stored element equals a."""
        _lib.check("dwf_field_set", _lib.load().dwf_field_set(self.fopr._h, self._p, float(a)))

    def copy_from(self, other):
        _lib.check("dwf_field_copy", _lib.load().dwf_field_copy(self.fopr._h, self._p, other._p))

    def axpy(self, a, x):
        _lib.check("dwf_field_axpy",
                   _lib.load().dwf_field_axpy(self.fopr._h, self._p, float(a), x._p))

    def aypx(self, a, x):
        _lib.check("dwf_field_aypx",
                   _lib.load().dwf_field_aypx(self.fopr._h, float(a), self._p, x._p))

    def scal(self, a):
        _lib.check("dwf_field_scal", _lib.load().dwf_field_scal(self.fopr._h, self._p, float(a)))

    def norm2(self):
        r = np.zeros(1)
        _lib.check("dwf_field_norm2", _lib.load().dwf_field_norm2(self.fopr._h, self._p, _ptr(r)))
        return float(r[0])

    def dot(self, other):
        r = np.zeros(1)
        _lib.check("dwf_field_dot",
                   _lib.load().dwf_field_dot(self.fopr._h, self._p, other._p, _ptr(r)))
        return float(r[0])

    def dotc(self, other):
        re, im = np.zeros(1), np.zeros(1)
        _lib.check("dwf_field_dotc", _lib.load().dwf_field_dotc(
            self.fopr._h, self._p, other._p, _ptr(re), _ptr(im)))
        return complex(re[0], im[0])

    def convert_precision_from(self, other):
        _lib.check("dwf_field_convert_precision", _lib.load().dwf_field_convert_precision(
            self.fopr._h, self._p, other.fopr._h, other._p))


class AFopr_Domainwall_5din_eo:
    """Even-odd Moebius domain-wall operator on the GPU.

    Mirrors AFopr<AFIELD> for fermion_type "Domainwall_5din_eo"; parameters
    carry the YAML names of the reference (quark_mass, domain_wall_height,
    extent_of_5th_dimension, coefficient_b, coefficient_c, boundary_condition).
    """

    def __init__(self, lattice_size, extent_of_5th_dimension, quark_mass,
                 domain_wall_height, coefficient_b=1.0, coefficient_c=0.0,
                 boundary_condition=(1, 1, 1, -1), precision="double", parameter_alpha=1.0):
        L = _lib.load()
        Ls = int(extent_of_5th_dimension)
        prm = _lib.FoprParams()
        for i in range(4):
            prm.lattice_size[i] = int(lattice_size[i])
            prm.boundary_condition[i] = int(boundary_condition[i])
        prm.extent_of_5th_dimension = Ls
        prm.quark_mass = float(quark_mass)
        prm.domain_wall_height = float(domain_wall_height)
        self._b = None
        self._c = None
        if np.isscalar(coefficient_b):
            prm.b_uniform = float(coefficient_b)
        else:
            self._b = np.ascontiguousarray(coefficient_b, np.float64)
            prm.coefficient_b = _ptr(self._b)
        if np.isscalar(coefficient_c):
            prm.c_uniform = float(coefficient_c)
        else:
            self._c = np.ascontiguousarray(coefficient_c, np.float64)
            prm.coefficient_c = _ptr(self._c)
        prm.parameter_alpha = float(parameter_alpha)
        prm.precision = _PREC[precision]
        h = C.c_void_p()
        _lib.check("dwf_fopr_create", L.dwf_fopr_create(C.byref(prm), C.byref(h)))
        self._h = h.value
        self.lattice_size = tuple(int(v) for v in lattice_size)
        self.Ls = Ls
        self.Vh = L.dwf_fopr_field_nvol(self._h)
        self.precision = prm.precision

    def __del__(self):
        if getattr(self, "_h", None):
            try:
                _lib.load().dwf_fopr_destroy(self._h)
            except Exception:
                pass
            self._h = None

    @property
    def handle(self):
        return self._h

    def new_field(self):
        return Field(self)

    def set_config(self, U):
        U = np.ascontiguousarray(U, np.float64)
        _lib.check("dwf_fopr_set_config", _lib.load().dwf_fopr_set_config(self._h, _ptr(U)))

    def set_config_smeared(self, smear, U=None):
        """AFopr_Smeared_alt::set_config: smear U, give the top level to the kernel."""
        Up = None if U is None else _ptr(np.ascontiguousarray(U, np.float64))
        _lib.check("dwf_fopr_set_config_smeared",
                   _lib.load().dwf_fopr_set_config_smeared(self._h, smear._h, Up))

    def set_mode(self, mode):
        _lib.check("dwf_fopr_set_mode", _lib.load().dwf_fopr_set_mode(self._h, mode.encode()))

    def get_mode(self):
        return _lib.load().dwf_fopr_get_mode(self._h).decode()

    def mult(self, v, w, mode=None):
        L = _lib.load()
        if mode is None:
            _lib.check("dwf_fopr_mult", L.dwf_fopr_mult(self._h, v.ptr, w.ptr))
        else:
            _lib.check("dwf_fopr_mult_mode",
                       L.dwf_fopr_mult_mode(self._h, v.ptr, w.ptr, mode.encode()))

    def mult_dag(self, v, w, mode=None):
        L = _lib.load()
        if mode is None:
            _lib.check("dwf_fopr_mult_dag", L.dwf_fopr_mult_dag(self._h, v.ptr, w.ptr))
        else:
            _lib.check("dwf_fopr_mult_dag_mode",
                       L.dwf_fopr_mult_dag_mode(self._h, v.ptr, w.ptr, mode.encode()))

    def mult_gm5(self, v, w):
        _lib.check("dwf_fopr_mult_gm5", _lib.load().dwf_fopr_mult_gm5(self._h, v.ptr, w.ptr))

    def mult_gm5R(self, v, w):
        _lib.check("dwf_fopr_mult_gm5R", _lib.load().dwf_fopr_mult_gm5R(self._h, v.ptr, w.ptr))

    def field_nin(self):
        return _lib.load().dwf_fopr_field_nin(self._h)

    def field_nvol(self):
        return _lib.load().dwf_fopr_field_nvol(self._h)

    def field_nex(self):
        return _lib.load().dwf_fopr_field_nex(self._h)

    def field_bytes(self):
        return _lib.load().dwf_fopr_field_bytes(self._h)

    def flop_count(self, mode):
        return _lib.load().dwf_fopr_flop_count(self._h, mode.encode())

    def convert(self, v, w):
        """corelib Field(24, V/2, Ls) host array w -> device field v."""
        w = np.ascontiguousarray(w, np.float64)
        assert w.size == 24 * self.Vh * self.Ls
        _lib.check("dwf_fopr_convert", _lib.load().dwf_fopr_convert(self._h, v.ptr, _ptr(w)))

    def reverse(self, w):
        """device field w -> corelib Field(24, V/2, Ls) host array (Ls, V/2, 24)."""
        out = np.empty((self.Ls, self.Vh, 24))
        _lib.check("dwf_fopr_reverse", _lib.load().dwf_fopr_reverse(self._h, _ptr(out), w.ptr))
        return out

    # convenience: host in, host out
    def apply(self, mode, w_host, dagger=False):
        w = self.new_field()
        v = self.new_field()
        self.convert(w, w_host)
        saved = self.get_mode()
        if mode in ("Deo", "Doe", "Dee", "Doo", "Dee_inv", "Doo_inv"):
            (self.mult_dag if dagger else self.mult)(v, w, mode)
        else:
            self.set_mode(mode)
            (self.mult_dag if dagger else self.mult)(v, w)
            self.set_mode(saved)
        return self.reverse(v)


class ASolver_CG:
    """ASolver_CG on fopr.mult in fopr's current mode (asolver_CG-tmpl.h)."""

    def __init__(self, fopr, maximum_number_of_iteration, convergence_criterion_squared,
                 initial_guess_mode="RHS"):
        g = {"RHS": GUESS_RHS, "ZERO": GUESS_ZERO, "GIVEN": GUESS_GIVEN}[initial_guess_mode]
        h = C.c_void_p()
        _lib.check("dwf_solver_create_cg", _lib.load().dwf_solver_create_cg(
            fopr._h, int(maximum_number_of_iteration), float(convergence_criterion_squared), g,
            C.byref(h)))
        self._h = h.value
        self.fopr = fopr

    def solve(self, x, b, raise_on_nonconvergence=False):
        """Returns (nconv, diff); nconv = -1 when not converged."""
        n = C.c_int(0)
        d = np.zeros(1)
        rc = _lib.load().dwf_solver_solve(self._h, x.ptr, b.ptr, C.byref(n), _ptr(d))
        _lib.check("dwf_solver_solve", rc,
                   allow=() if raise_on_nonconvergence else (_lib.ERR_NOT_CONVERGED,))
        return n.value, float(d[0])

    def inner_iterations(self):
        return _lib.load().dwf_solver_inner_iterations(self._h)

    def __del__(self):
        if getattr(self, "_h", None):
            try:
                _lib.load().dwf_solver_destroy(self._h)
            except Exception:
                pass
            self._h = None


class ASolver_Richardson(ASolver_CG):
    """Richardson on fopr_d, preconditioned by APrecond_Mixedprec with an inner
    ASolver_CG on fopr_f from a zero guess (fprop_alt_Standard_eo_Richardson)."""

    def __init__(self, fopr_d, fopr_f, maximum_number_of_iteration,
                 convergence_criterion_squared, inner_iteration=100000,
                 inner_convergence_criterion_squared=1.0e-8):
        h = C.c_void_p()
        _lib.check("dwf_solver_create_richardson", _lib.load().dwf_solver_create_richardson(
            fopr_d._h, fopr_f._h, int(maximum_number_of_iteration),
            float(convergence_criterion_squared), int(inner_iteration),
            float(inner_convergence_criterion_squared), C.byref(h)))
        self._h = h.value
        self.fopr = fopr_d
        self.fopr_f = fopr_f


class _SolverHandle(ASolver_CG):
    """A solver created by the library (dwf_solver_tune_richardson)."""

    def __init__(self, handle, fopr, fopr_f=None):
        self._h = handle
        self.fopr = fopr
        self.fopr_f = fopr_f


def tune_richardson(fopr_d, inner_foprs, inner_criteria, b, x, maximum_number_of_iteration=100,
                    convergence_criterion_squared=1.0e-20, inner_iteration=100000):
    """Schedule autotuner of the mixed-precision solve (no counterpart in the
    reference): times one solve of fopr_d's system (right-hand side b, x
    overwritten) per inner operator (storage type) and inner criterion, after an
    untimed one, and returns (solver of the fastest candidate that converged,
    times[i][j] in seconds or -1, (i, j) of the choice)."""
    ops = list(inner_foprs)
    tols = np.ascontiguousarray(inner_criteria, np.float64)
    arr = (C.c_void_p * len(ops))(*[o._h for o in ops])
    times = np.zeros(len(ops) * len(tols))
    best = C.c_int(-1)
    h = C.c_void_p()
    _lib.check("dwf_solver_tune_richardson", _lib.load().dwf_solver_tune_richardson(
        fopr_d._h, len(ops), arr, len(tols), _ptr(tols), int(maximum_number_of_iteration),
        float(convergence_criterion_squared), int(inner_iteration), b.ptr, x.ptr, _ptr(times),
        C.byref(best), C.byref(h)))
    i, j = divmod(best.value, len(tols))
    return _SolverHandle(h.value, fopr_d, ops[i]), times.reshape(len(ops), len(tols)), (i, j)


class APrecond_LearnedPolynomial:
    """Learned polynomial preconditioner K = sum_k W_k H^k, H = 1 - D (no
    counterpart in the reference).  Applied in fopr's precision; trained by
    least squares with any operator of the same lattice and parameters."""

    def __init__(self, fopr, degree):
        h = C.c_void_p()
        _lib.check("dwf_lpc_create", _lib.load().dwf_lpc_create(fopr._h, int(degree), C.byref(h)))
        self._h = h.value
        self.fopr = fopr
        self.degree = int(degree)
        self.loss = None

    @property
    def weights_shape(self):
        return (self.degree + 1, 2, self.fopr.Ls, self.fopr.Ls)

    def get_weights(self):
        W = np.zeros(self.weights_shape)
        _lib.check("dwf_lpc_get_weights", _lib.load().dwf_lpc_get_weights(self._h, _ptr(W)))
        return W

    def set_weights(self, W):
        W = np.ascontiguousarray(W, np.float64)
        assert W.shape == self.weights_shape
        _lib.check("dwf_lpc_set_weights", _lib.load().dwf_lpc_set_weights(self._h, _ptr(W)))

    def train(self, fopr_train, nvec=2, seed=12345, ridge=1e-10):
        loss = np.zeros(1)
        _lib.check("dwf_lpc_train", _lib.load().dwf_lpc_train(
            self._h, fopr_train._h, int(nvec), int(seed), float(ridge), _ptr(loss)))
        self.loss = float(loss[0])
        return self.loss

    def apply(self, which, v, w):
        _lib.check("dwf_lpc_apply", _lib.load().dwf_lpc_apply(self._h, which.encode(), v.ptr, w.ptr))

    def __del__(self):
        if getattr(self, "_h", None):
            try:
                _lib.load().dwf_lpc_destroy(self._h)
            except Exception:
                pass
            self._h = None


def chiral_gram(fopr, fields):
    """Chiral Gram matrices of up to 6 fields (all pairs i <= j)."""
    F = len(fields)
    NP = F * (F + 1) // 2
    out = np.zeros((NP, 2, fopr.Ls, fopr.Ls))
    arr = (C.c_void_p * F)(*[f.ptr for f in fields])
    _lib.check("dwf_lpc_chiral_gram", _lib.load().dwf_lpc_chiral_gram(fopr._h, F, arr, _ptr(out)))
    return out


def fill_uniform(field, seed):
    _lib.check("dwf_field_fill_uniform",
               _lib.load().dwf_field_fill_uniform(field.fopr._h, field.ptr, int(seed)))


class ASolver_FCG(ASolver_CG):
    """Flexible CG on fopr (mode DdagD) preconditioned by M = K K^+."""

    def __init__(self, fopr, prec, maximum_number_of_iteration, convergence_criterion_squared):
        h = C.c_void_p()
        _lib.check("dwf_solver_create_fcg", _lib.load().dwf_solver_create_fcg(
            fopr._h, prec._h, int(maximum_number_of_iteration),
            float(convergence_criterion_squared), C.byref(h)))
        self._h = h.value
        self.fopr = fopr
        self.prec = prec


class ASolver_Richardson_LPC(ASolver_CG):
    """Richardson on fopr_d with an inner flexible CG on fopr_f preconditioned
    by the learned preconditioner prec."""

    def __init__(self, fopr_d, fopr_f, prec, maximum_number_of_iteration,
                 convergence_criterion_squared, inner_iteration=100000,
                 inner_convergence_criterion_squared=1.0e-8):
        h = C.c_void_p()
        _lib.check("dwf_solver_create_richardson_lpc",
                   _lib.load().dwf_solver_create_richardson_lpc(
                       fopr_d._h, fopr_f._h, prec._h, int(maximum_number_of_iteration),
                       float(convergence_criterion_squared), int(inner_iteration),
                       float(inner_convergence_criterion_squared), C.byref(h)))
        self._h = h.value
        self.fopr = fopr_d
        self.fopr_f = fopr_f
        self.prec = prec


class AShiftsolver_CG:
    """Multishift CG (ashiftsolver_CG-tmpl.h) on fopr.mult + sigma_i."""

    def __init__(self, fopr, maximum_number_of_iteration, convergence_criterion_squared):
        self.fopr = fopr
        self.niter = int(maximum_number_of_iteration)
        self.stop = float(convergence_criterion_squared)

    def solve(self, xs, sigma, b, raise_on_nonconvergence=True):
        sig = np.ascontiguousarray(sigma, np.float64)
        arr = (C.c_void_p * len(xs))(*[x.ptr for x in xs])
        n = C.c_int(0)
        d = np.zeros(1)
        rc = _lib.load().dwf_shiftsolver_cg_solve(self.fopr._h, self.niter, self.stop, len(xs),
                                                  _ptr(sig), arr, b.ptr, C.byref(n), _ptr(d))
        _lib.check("dwf_shiftsolver_cg_solve", rc,
                   allow=() if raise_on_nonconvergence else (_lib.ERR_NOT_CONVERGED,))
        return n.value, float(d[0])


class AShiftsolver_CG_Batch:
    """Batched multishift CG: several right-hand sides (on one operator or on
    up to four lattices) with the same shifts advance together; every hop of
    D^+D is one launch for all of them and the scalar recurrences of
    AShiftsolver_CG stay on the device.  graph=True replays the iteration from
    a captured HIP graph."""

    def __init__(self, fopr_list, nshift, graph=False):
        self.fopr_list = list(fopr_list)
        self.nshift = int(nshift)
        ops = (C.c_void_p * len(self.fopr_list))(*[f._h for f in self.fopr_list])
        h = C.c_void_p()
        _lib.check("dwf_msbatch_create", _lib.load().dwf_msbatch_create(
            len(self.fopr_list), ops, self.nshift, 1 if graph else 0, C.byref(h)))
        self._h = h.value

    def __del__(self):
        try:
            if getattr(self, "_h", None):
                _lib.load().dwf_msbatch_destroy(self._h)
                self._h = None
        except Exception:
            pass

    def solve(self, xs, sigma, bs, maximum_number_of_iteration, convergence_criterion_squared,
              raise_on_nonconvergence=True):
        """xs[k][j] solves (A_k + sigma_j) x = bs[k]; returns (nconv list, diff list)."""
        nsys = len(self.fopr_list)
        assert len(bs) == nsys and len(xs) == nsys and len(sigma) == self.nshift
        sig = np.ascontiguousarray(sigma, np.float64)
        xa = (C.c_void_p * (nsys * self.nshift))(*[x.ptr for row in xs for x in row])
        ba = (C.c_void_p * nsys)(*[b.ptr for b in bs])
        n = (C.c_int * nsys)()
        d = np.zeros(nsys)
        rc = _lib.load().dwf_msbatch_solve(self._h, _ptr(sig), xa, ba,
                                           int(maximum_number_of_iteration),
                                           float(convergence_criterion_squared), n, _ptr(d))
        _lib.check("dwf_msbatch_solve", rc,
                   allow=() if raise_on_nonconvergence else (_lib.ERR_NOT_CONVERGED,))
        return [int(v) for v in n], [float(v) for v in d]

    def iterations(self):
        return int(_lib.load().dwf_msbatch_iterations(self._h))

    def solve_mixed(self, fopr64_list, xs, sigma, bs, maximum_number_of_iteration,
                    convergence_criterion_squared, low=1e-10, inner=1e-10,
                    raise_on_nonconvergence=True):
        """FP32 batched multishift (this object, on FP32 operators) to |r|^2/|b|^2 < low,
        then every shift refined in FP64 (xs, bs: fields of the FP64 operators)."""
        nsys = len(self.fopr_list)
        assert len(fopr64_list) == nsys and len(bs) == nsys and len(xs) == nsys
        sig = np.ascontiguousarray(sigma, np.float64)
        ops = (C.c_void_p * nsys)(*[f._h for f in fopr64_list])
        xa = (C.c_void_p * (nsys * self.nshift))(*[x.ptr for row in xs for x in row])
        ba = (C.c_void_p * nsys)(*[b.ptr for b in bs])
        n = (C.c_int * nsys)()
        d = np.zeros(nsys)
        rc = _lib.load().dwf_msbatch_solve_mixed(self._h, ops, _ptr(sig), xa, ba,
                                                 int(maximum_number_of_iteration),
                                                 float(convergence_criterion_squared),
                                                 float(low), float(inner), n, _ptr(d))
        _lib.check("dwf_msbatch_solve_mixed", rc,
                   allow=() if raise_on_nonconvergence else (_lib.ERR_NOT_CONVERGED,))
        return [int(v) for v in n], [float(v) for v in d]

    def refine_iterations(self):
        return int(_lib.load().dwf_msbatch_refine_iterations(self._h))


class DeviceArray:
    """A plain device buffer of float64 (e.g. a Field_G force field)."""

    def __init__(self, shape):
        self.shape = tuple(shape)
        self.nbytes = int(np.prod(self.shape)) * 8
        p = C.c_void_p()
        _lib.check("dwf_device_alloc", _lib.load().dwf_device_alloc(self.nbytes, C.byref(p)))
        self._p = p.value

    @property
    def ptr(self):
        return self._p

    def to_host(self):
        out = np.empty(self.shape)
        _lib.check("dwf_memcpy_to_host", _lib.load().dwf_memcpy_to_host(
            out.ctypes.data_as(C.c_void_p), self._p, self.nbytes))
        return out

    def from_host(self, a):
        a = np.ascontiguousarray(a, np.float64)
        assert a.size * 8 == self.nbytes
        _lib.check("dwf_memcpy_to_device", _lib.load().dwf_memcpy_to_device(
            self._p, a.ctypes.data_as(C.c_void_p), self.nbytes))

    def __del__(self):
        if getattr(self, "_p", None):
            try:
                _lib.load().dwf_device_free(self._p)
            except Exception:
                pass
            self._p = None


class AForce_F_Domainwall_5din_eo:
    """DWF fermion force on the GPU (FP64): force_udiv1 / force_udiv /
    force_core / force_core1 into a device Field_G (4, V, 18)."""

    def __init__(self, fopr):
        h = C.c_void_p()
        _lib.check("dwf_force_create", _lib.load().dwf_force_create(fopr._h, C.byref(h)))
        self._h = h.value
        self.fopr = fopr
        V = int(np.prod(fopr.lattice_size))
        self.gauge_shape = (4, V, 18)

    def new_force(self):
        return DeviceArray(self.gauge_shape)

    def set_mode(self, mode):
        _lib.check("dwf_force_set_mode", _lib.load().dwf_force_set_mode(self._h, mode.encode()))

    def force_udiv1(self, F, zeta, eta):
        _lib.check("dwf_force_udiv1", _lib.load().dwf_force_udiv1(self._h, F.ptr, zeta.ptr, eta.ptr))

    def force_udiv(self, F, eta):
        _lib.check("dwf_force_udiv", _lib.load().dwf_force_udiv(self._h, F.ptr, eta.ptr))

    def mult_generator(self, F):
        _lib.check("dwf_force_mult_generator", _lib.load().dwf_force_mult_generator(self._h, F.ptr))

    def force_core(self, F, eta):
        _lib.check("dwf_force_core", _lib.load().dwf_force_core(self._h, F.ptr, eta.ptr))

    def force_core1(self, F, zeta, eta):
        _lib.check("dwf_force_core1", _lib.load().dwf_force_core1(self._h, F.ptr, zeta.ptr, eta.ptr))

    def __del__(self):
        if getattr(self, "_h", None):
            try:
                _lib.load().dwf_force_destroy(self._h)
            except Exception:
                pass
            self._h = None


class Rational:
    """Coefficients of r(x) = a0 + sum_i bl[i] / (x + cl[i]) with the shift
    solver's parameters (Math_Rational + the solver block of the YAML)."""

    def __init__(self, a0, bl, cl, maximum_number_of_iteration=10000,
                 convergence_criterion_squared=1e-24):
        self.a0 = float(a0)
        self.bl = np.ascontiguousarray(bl, np.float64)
        self.cl = np.ascontiguousarray(cl, np.float64)
        assert self.bl.shape == self.cl.shape and self.bl.size >= 1
        self.max_iter = int(maximum_number_of_iteration)
        self.stop_cond = float(convergence_criterion_squared)

    @classmethod
    def from_file(cls, path, **solver):
        """A Math_Rational coefficient file (see dwf_amd.lattice.read_rational)."""
        from .lattice import read_rational
        d = read_rational(path)
        return cls(d["a0"], d["res"], d["pole"], **solver)

    def params(self):
        p = _lib.RationalParams()
        p.np = self.bl.size
        p.a0 = self.a0
        p.bl = _ptr(self.bl)
        p.cl = _ptr(self.cl)
        p.max_iter = self.max_iter
        p.stop_cond = self.stop_cond
        return p


class AFopr_Rational:
    """AFopr_Rational (src/lib/Fopr/afopr_Rational-tmpl.h:181-219): v = r(A^dag A) b."""

    def __init__(self, fopr, rational):
        self.fopr = fopr
        self.rational = rational
        self.nconv = None
        self.diff = None

    def mult(self, v, b):
        nconv, diff = C.c_int(0), np.zeros(1)
        p = self.rational.params()
        _lib.check("dwf_rational_mult", _lib.load().dwf_rational_mult(
            self.fopr._h, C.byref(p), v.ptr, b.ptr, C.byref(nconv), _ptr(diff)))
        self.nconv, self.diff = nconv.value, float(diff[0])


class AForce_F_Rational_Ratio:
    """AForce_F_Rational_Ratio (src/lib_alt/Force/Fermion/aforce_F_Rational_Ratio-tmpl.h):
    force of <phi| r1(A^dag A) r2(B^dag B) r1(A^dag A) |phi>, A = fopr1, B = fopr2."""

    def __init__(self, fopr1, fopr2, numerator, denominator):
        h = C.c_void_p()
        p1, p2 = numerator.params(), denominator.params()
        _lib.check("dwf_force_rational_ratio_create", _lib.load().dwf_force_rational_ratio_create(
            fopr1._h, fopr2._h, C.byref(p1), C.byref(p2), C.byref(h)))
        self._h = h.value
        self.fopr1, self.fopr2 = fopr1, fopr2
        V = int(np.prod(fopr1.lattice_size))
        self.gauge_shape = (4, V, 18)

    def new_force(self):
        return DeviceArray(self.gauge_shape)

    def force_udiv(self, F, eta):
        _lib.check("dwf_force_rational_ratio_udiv",
                   _lib.load().dwf_force_rational_ratio_udiv(self._h, F.ptr, eta.ptr))

    def force_core(self, F, eta):
        _lib.check("dwf_force_rational_ratio_core",
                   _lib.load().dwf_force_rational_ratio_core(self._h, F.ptr, eta.ptr))

    def __del__(self):
        if getattr(self, "_h", None):
            try:
                _lib.load().dwf_force_rational_ratio_destroy(self._h)
            except Exception:
                pass
            self._h = None


class Director_alt_Smear:
    """Stout smearing levels (APE staples + stout projection) on the GPU."""

    def __init__(self, lattice_size, rho_uniform, number_of_smearing):
        h = C.c_void_p()
        dims = (C.c_int * 4)(*[int(v) for v in lattice_size])
        _lib.check("dwf_smear_create", _lib.load().dwf_smear_create(
            dims, float(rho_uniform), int(number_of_smearing), C.byref(h)))
        self._h = h.value
        self.lattice_size = tuple(int(v) for v in lattice_size)
        self.Nsmear = int(number_of_smearing)

    def set_config(self, U):
        U = np.ascontiguousarray(U, np.float64)
        _lib.check("dwf_smear_set_config", _lib.load().dwf_smear_set_config(self._h, _ptr(U)))

    def get_config(self, level):
        V = int(np.prod(self.lattice_size))
        out = np.empty((4, V, 18))
        _lib.check("dwf_smear_get_config",
                   _lib.load().dwf_smear_get_config(self._h, int(level), _ptr(out)))
        return out

    def force_udiv(self, Sigma, Sigmap, level):
        """Director_alt_Smear::force_udiv of one level (device arrays; may alias)."""
        _lib.check("dwf_smear_force_udiv", _lib.load().dwf_smear_force_udiv(
            self._h, int(level), Sigma.ptr, Sigmap.ptr))

    def mult_jacobian(self, F):
        _lib.check("dwf_smear_mult_jacobian", _lib.load().dwf_smear_mult_jacobian(self._h, F.ptr))

    def mult_generator(self, F):
        _lib.check("dwf_smear_mult_generator",
                   _lib.load().dwf_smear_mult_generator(self._h, F.ptr))

    def __del__(self):
        if getattr(self, "_h", None):
            try:
                _lib.load().dwf_smear_destroy(self._h)
            except Exception:
                pass
            self._h = None


class AForce_F_Smeared_alt:
    """AForce_F_Smeared_alt (src/lib_alt/Force/Fermion/aforce_F_Smeared_alt-tmpl.h):
    the force of an operator whose links are the top level of a smearing
    director: the inner force on the smeared links, mult_jacobian down to the
    thin links, then mult_generator with the thin links."""

    def __init__(self, force, smear):
        self.force = force
        self.smear = smear

    def new_force(self):
        return self.force.new_force()

    def set_mode(self, mode):
        self.force.set_mode(mode)

    def force_udiv(self, F, eta):
        self.force.force_udiv(F, eta)
        self.smear.mult_jacobian(F)

    def force_udiv1(self, F, zeta, eta):
        self.force.force_udiv1(F, zeta, eta)
        self.smear.mult_jacobian(F)

    def force_core(self, F, eta):
        self.force_udiv(F, eta)
        self.smear.mult_generator(F)

    def force_core1(self, F, zeta, eta):
        self.force_udiv1(F, zeta, eta)
        self.smear.mult_generator(F)


def plaquette(lattice_size, U):
    U = np.ascontiguousarray(U, np.float64)
    r = np.zeros(1)
    dims = (C.c_int * 4)(*[int(v) for v in lattice_size])
    _lib.check("dwf_plaquette", _lib.load().dwf_plaquette(dims, _ptr(U), _ptr(r)))
    return float(r[0])
