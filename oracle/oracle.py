"""ctypes binding of the CPU oracle (liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, always as the checker or the timed CPU baseline,
never as the product path.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

_dp = C.POINTER(C.c_double)


def _ptr(a):
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(_dp)


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError(f"oracle library not built: {path} (run make -C oracle)")
        L = C.CDLL(path)
        L.orc_dwf_create.restype = C.c_void_p
        L.orc_dwf_create.argtypes = [C.POINTER(C.c_int), C.c_int, C.c_double, C.c_double,
                                     _dp, _dp, C.c_double, C.POINTER(C.c_int), C.c_int]
        L.orc_dwf_destroy.argtypes = [C.c_void_p]
        L.orc_dwf_set_config.argtypes = [C.c_void_p, _dp]
        for f in ("orc_mult", "orc_mult_dag"):
            getattr(L, f).argtypes = [C.c_void_p, C.c_char_p, _dp, _dp]
            getattr(L, f).restype = C.c_int
        for f in ("orc_mult_gm5", "orc_mult_gm5R"):
            getattr(L, f).argtypes = [C.c_void_p, _dp, _dp]
        L.orc_wilson_meo.argtypes = [C.c_void_p, _dp, _dp, C.c_int]
        L.orc_cg_solve.argtypes = [C.c_void_p, C.c_char_p, _dp, _dp, C.c_int, C.c_double,
                                   C.c_int, C.POINTER(C.c_int), _dp]
        L.orc_cg_solve.restype = C.c_int
        L.orc_richardson_solve.argtypes = [C.c_void_p, C.c_void_p, C.c_char_p, _dp, _dp,
                                           C.c_int, C.c_double, C.c_int, C.c_double,
                                           C.POINTER(C.c_int), _dp, C.POINTER(C.c_int)]
        L.orc_richardson_solve.restype = C.c_int
        L.orc_shiftcg_solve.argtypes = [C.c_void_p, C.c_char_p, C.c_int, _dp, C.POINTER(_dp),
                                        _dp, C.c_int, C.c_double, C.POINTER(C.c_int), _dp]
        L.orc_shiftcg_solve.restype = C.c_int
        L.orc_force_udiv1.argtypes = [C.c_void_p, C.c_char_p, _dp, _dp, _dp]
        L.orc_force_udiv1.restype = C.c_int
        L.orc_force_udiv.argtypes = [C.c_void_p, _dp, _dp]
        L.orc_mult_generator.argtypes = [C.c_void_p, _dp]
        L.orc_stout_smear.argtypes = [C.POINTER(C.c_int), C.c_double, _dp, _dp]
        L.orc_stout_force_udiv.argtypes = [C.POINTER(C.c_int), C.c_double, _dp, _dp, _dp]
        L.orc_mult_generator_gauge.argtypes = [C.POINTER(C.c_int), _dp, _dp]
        L.orc_plaquette.argtypes = [C.POINTER(C.c_int), _dp]
        L.orc_plaquette.restype = C.c_double
        L.orc_exp_iQ.argtypes = [_dp, _dp]
        L.orc_exp_iQ_taylor.argtypes = [_dp, _dp]
        L.orc_num_threads.restype = C.c_int
        _LIB = L
    return _LIB


class OracleDWF:
    """The oracle's even-odd Moebius domain-wall operator (corelib semantics)."""

    def __init__(self, dims, Ls, mq, M0, b=1.0, c=0.0, alpha=1.0, bc=(1, 1, 1, -1),
                 precision="double"):
        L = lib()
        self.dims = tuple(int(v) for v in dims)
        self.Ls = int(Ls)
        bv = np.full(Ls, float(b)) if np.isscalar(b) else np.asarray(b, np.float64)
        cv = np.full(Ls, float(c)) if np.isscalar(c) else np.asarray(c, np.float64)
        self.b, self.c = np.ascontiguousarray(bv), np.ascontiguousarray(cv)
        d = (C.c_int * 4)(*self.dims)
        bcv = (C.c_int * 4)(*[int(v) for v in bc])
        prec = 0 if precision == "double" else 1
        self._h = L.orc_dwf_create(d, self.Ls, float(mq), float(M0), _ptr(self.b),
                                   _ptr(self.c), float(alpha), bcv, prec)
        if not self._h:
            raise ValueError("invalid oracle operator parameters")
        self.Vh = (self.dims[0] * self.dims[1] * self.dims[2] * self.dims[3]) // 2
        self.shape = (self.Ls, self.Vh, 24)

    def __del__(self):
        if getattr(self, "_h", None):
            lib().orc_dwf_destroy(self._h)
            self._h = None

    def set_config(self, U):
        U = np.ascontiguousarray(U, np.float64)
        lib().orc_dwf_set_config(self._h, _ptr(U))

    def mult(self, mode, w):
        w = np.ascontiguousarray(w, np.float64)
        v = np.empty_like(w)
        if lib().orc_mult(self._h, mode.encode(), _ptr(v), _ptr(w)) != 0:
            raise ValueError(f"unknown mode {mode}")
        return v

    def mult_dag(self, mode, w):
        w = np.ascontiguousarray(w, np.float64)
        v = np.empty_like(w)
        if lib().orc_mult_dag(self._h, mode.encode(), _ptr(v), _ptr(w)) != 0:
            raise ValueError(f"unknown mode {mode}")
        return v

    def mult_gm5(self, w):
        w = np.ascontiguousarray(w, np.float64)
        v = np.empty_like(w)
        lib().orc_mult_gm5(self._h, _ptr(v), _ptr(w))
        return v

    def mult_gm5R(self, w):
        w = np.ascontiguousarray(w, np.float64)
        v = np.empty_like(w)
        lib().orc_mult_gm5R(self._h, _ptr(v), _ptr(w))
        return v

    def wilson_meo(self, w4, ieo):
        w4 = np.ascontiguousarray(w4, np.float64)
        v4 = np.empty_like(w4)
        lib().orc_wilson_meo(self._h, _ptr(v4), _ptr(w4), int(ieo))
        return v4

    def _gauge_shape(self):
        V = self.dims[0] * self.dims[1] * self.dims[2] * self.dims[3]
        return (4, V, 18)

    def force_udiv1(self, mode, zeta, eta):
        zeta = np.ascontiguousarray(zeta, np.float64)
        eta = np.ascontiguousarray(eta, np.float64)
        F = np.empty(self._gauge_shape())
        if lib().orc_force_udiv1(self._h, mode.encode(), _ptr(F), _ptr(zeta), _ptr(eta)) != 0:
            raise ValueError(f"unknown force mode {mode}")
        return F

    def force_udiv(self, eta):
        eta = np.ascontiguousarray(eta, np.float64)
        F = np.empty(self._gauge_shape())
        lib().orc_force_udiv(self._h, _ptr(F), _ptr(eta))
        return F

    def mult_generator(self, F):
        F = np.ascontiguousarray(F, np.float64).copy()
        lib().orc_mult_generator(self._h, _ptr(F))
        return F

    def force_core(self, eta):
        return self.mult_generator(self.force_udiv(eta))

    def force_core1(self, mode, zeta, eta):
        return self.mult_generator(self.force_udiv1(mode, zeta, eta))

    def cg(self, b, mode="DdagD", niter=10000, stop_cond2=1e-24, init="RHS"):
        b = np.ascontiguousarray(b, np.float64)
        x = np.empty_like(b)
        nconv = C.c_int(0)
        diff = np.zeros(1)
        lib().orc_cg_solve(self._h, mode.encode(), _ptr(x), _ptr(b), int(niter),
                           float(stop_cond2), 0 if init == "RHS" else 2, C.byref(nconv),
                           _ptr(diff))
        return x, nconv.value, float(diff[0])

    def shift_cg(self, b, sigma, mode="DdagD", niter=10000, stop_cond2=1e-24):
        b = np.ascontiguousarray(b, np.float64)
        sig = np.ascontiguousarray(sigma, np.float64)
        xs = [np.empty_like(b) for _ in sig]
        arr = (_dp * len(xs))(*[_ptr(x) for x in xs])
        nconv = C.c_int(0)
        diff = np.zeros(1)
        rc = lib().orc_shiftcg_solve(self._h, mode.encode(), len(sig), _ptr(sig), arr,
                                     _ptr(b), int(niter), float(stop_cond2), C.byref(nconv),
                                     _ptr(diff))
        return xs, nconv.value, float(diff[0]), rc


def rational_mult(orc, a0, bl, cl, b, niter=10000, stop_cond2=1e-24):
    """AFopr_Rational::mult (src/lib/Fopr/afopr_Rational-tmpl.h:181-219):
    returns (v, x) with x_i = (A^dag A + cl_i)^-1 b and v = a0 b + sum_i bl_i x_i."""
    xs, _, _, _ = orc.shift_cg(b, cl, "DdagD", niter, stop_cond2)
    v = a0 * b
    for bi, xi in zip(bl, xs):
        v = v + bi * xi
    return v, xs


def rational_ratio_force_udiv(orc1, orc2, rat1, rat2, eta, niter=10000, stop_cond2=1e-24):
    """AForce_F_Rational_Ratio::force_udiv_impl
    (src/lib_alt/Force/Fermion/aforce_F_Rational_Ratio-tmpl.h:345-447), composed
    from the oracle's multishift CG, force_udiv and force_udiv1.
    rat = (a0, bl, cl) of the numerator (orc1 = A) and denominator (orc2 = B)."""
    a01, bl1, cl1 = rat1
    a02, bl2, cl2 = rat2
    v1, eta1 = rational_mult(orc1, a01, bl1, cl1, eta, niter, stop_cond2)
    v2, eta2 = rational_mult(orc2, a02, bl2, cl2, v1, niter, stop_cond2)
    zeta1, _, _, _ = orc1.shift_cg(v2, cl1, "DdagD", niter, stop_cond2)
    F = np.zeros(orc1._gauge_shape())
    for b, e2 in zip(bl2, eta2):
        F += b * orc2.force_udiv(e2)
    for b, e1, z1 in zip(bl1, eta1, zeta1):
        h1 = orc1.mult("H", e1)
        h2 = orc1.mult("H", z1)
        F += b * orc1.force_udiv1("H", h1, z1)
        F += b * orc1.force_udiv1("H", h2, e1)
        F += b * orc1.force_udiv1("Hdag", e1, h2)
        F += b * orc1.force_udiv1("Hdag", z1, h1)
    return F


def richardson(op64, op32, b, mode="DdagD", niter_outer=100, stop_outer2=1e-28,
               niter_inner=10000, stop_inner2=1e-8):
    b = np.ascontiguousarray(b, np.float64)
    x = np.empty_like(b)
    nconv = C.c_int(0)
    inner = C.c_int(0)
    diff = np.zeros(1)
    lib().orc_richardson_solve(op64._h, op32._h, mode.encode(), _ptr(x), _ptr(b),
                               int(niter_outer), float(stop_outer2), int(niter_inner),
                               float(stop_inner2), C.byref(nconv), _ptr(diff), C.byref(inner))
    return x, nconv.value, float(diff[0]), inner.value


def stout_smear(dims, rho, U):
    U = np.ascontiguousarray(U, np.float64)
    out = np.empty_like(U)
    lib().orc_stout_smear((C.c_int * 4)(*dims), float(rho), _ptr(out), _ptr(U))
    return out


def stout_force_udiv(dims, rho, Sigmap, U):
    """AForceSmear_APE::force_udiv of one stout level
    (lib_alt_QXS/Smear/aforceSmear_APE-tmpl.h:157-206): dS/dU from dS/dU'."""
    Sigmap = np.ascontiguousarray(Sigmap, np.float64)
    U = np.ascontiguousarray(U, np.float64)
    out = np.empty_like(U)
    lib().orc_stout_force_udiv((C.c_int * 4)(*dims), float(rho), _ptr(out), _ptr(Sigmap), _ptr(U))
    return out


def stout_mult_jacobian(dims, rho, levels, F):
    """AForce_F_Smeared_alt::mult_jacobian (aforce_F_Smeared_alt-tmpl.h:170-191):
    levels[i] is the input of smearing level i (levels[0] = the thin links)."""
    for ismear in range(len(levels) - 1, -1, -1):
        F = stout_force_udiv(dims, rho, F, levels[ismear])
    return F


def mult_generator_gauge(dims, F, U):
    F = np.ascontiguousarray(F, np.float64).copy()
    U = np.ascontiguousarray(U, np.float64)
    lib().orc_mult_generator_gauge((C.c_int * 4)(*dims), _ptr(F), _ptr(U))
    return F


def plaquette(dims, U):
    U = np.ascontiguousarray(U, np.float64)
    return lib().orc_plaquette((C.c_int * 4)(*dims), _ptr(U))


def exp_iQ(iQ, taylor=False):
    iQ = np.ascontiguousarray(iQ, np.float64).reshape(18)
    e = np.empty(18)
    (lib().orc_exp_iQ_taylor if taylor else lib().orc_exp_iQ)(_ptr(e), _ptr(iQ))
    return e


def num_threads():
    return lib().orc_num_threads()
