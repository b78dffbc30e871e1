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


# ---- learned polynomial preconditioner (no reference counterpart) ----------
# K = sum_k W_k H^k, H = 1 - D; W_k = (W+, W-) per-chirality Ls x Ls matrices
# acting on P+ x and P- x (Dirac basis: P+- x = (x_k +- x_{k+12}) / 2 on both halves).

def chiral_apply(W, x):
    """(W x) = W+ P+ x + W- P- x for a corelib field x (Ls, Vh, 24); W (2, Ls, Ls)."""
    u = x[..., :12] + x[..., 12:]
    l = x[..., :12] - x[..., 12:]
    ou = np.einsum("st,tvk->svk", W[0], u)
    ol = np.einsum("st,tvk->svk", W[1], l)
    out = np.empty_like(x)
    out[..., :12] = 0.5 * (ou + ol)
    out[..., 12:] = 0.5 * (ou - ol)
    return out


def lpc_K(orc, W, x):
    acc = chiral_apply(W[0], x)
    z = x
    for k in range(1, len(W)):
        z = z - orc.mult("D", z)
        acc = acc + chiral_apply(W[k], z)
    return acc


def lpc_Kdag(orc, W, x):
    L = len(W) - 1
    Wt = [np.transpose(w, (0, 2, 1)) for w in W]
    t = chiral_apply(Wt[L], x)
    for k in range(L - 1, -1, -1):
        t = chiral_apply(Wt[k], x) + (t - orc.mult_dag("D", t))
    return t


def chiral_gram(fields):
    """All pairs i <= j: G[p][chi][s][t] = sum p_chi(f_i)(s) p_chi(f_j)(t)."""
    P = [(f[..., :12] + f[..., 12:], f[..., :12] - f[..., 12:]) for f in fields]
    out = []
    for i in range(len(fields)):
        for j in range(i, len(fields)):
            out.append([np.einsum("svk,tvk->st", P[i][c], P[j][c]) for c in (0, 1)])
    return np.array(out)


_M64 = (1 << 64) - 1


def hash_uniform(seed, idx):
    """x = uniform [-1, 1) of the counter hash (splitmix64) of (seed, idx)."""
    idx = np.asarray(idx, dtype=np.uint64)
    with np.errstate(over="ignore"):
        z = np.uint64(seed & _M64) + (idx + np.uint64(1)) * np.uint64(0x9E3779B97F4A7C15)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    return (z >> np.uint64(11)).astype(np.float64) * 2.0 ** -52 - 1.0


def uniform_field(seed, Vh, Ls):
    """The device fill of (seed) in the corelib layout (Ls, Vh, 24): device
    index ((site * 24) + k) * Ls + s."""
    s = np.arange(Ls, dtype=np.uint64)[:, None, None]
    site = np.arange(Vh, dtype=np.uint64)[None, :, None]
    k = np.arange(24, dtype=np.uint64)[None, None, :]
    return hash_uniform(seed, (site * np.uint64(24) + k) * np.uint64(Ls) + s)


def lpc_train(orc, L, vs, ridge=0.0):
    """Least-squares weights of K (degree L) with K D v ~ v over the fields vs;
    returns (W, relative loss)."""
    Ls = orc.Ls
    d = (L + 1) * Ls
    A = np.zeros((2, d, d))
    B = np.zeros((2, d, Ls))
    vv = np.zeros(2)
    for v in vs:
        z = [orc.mult("D", v)]
        for _ in range(L):
            z.append(z[-1] - orc.mult("D", z[-1]))
        G = chiral_gram([v] + z)
        F = L + 2
        pid = lambda i, j: i * F - i * (i - 1) // 2 + (j - i)
        for c in (0, 1):
            for a in range(L + 1):
                for b in range(L + 1):
                    i, j = 1 + a, 1 + b
                    blk = G[pid(i, j)][c] if i <= j else G[pid(j, i)][c].T
                    A[c, a * Ls:(a + 1) * Ls, b * Ls:(b + 1) * Ls] += blk
                B[c, a * Ls:(a + 1) * Ls, :] += G[pid(0, 1 + a)][c].T
            vv[c] += np.trace(G[pid(0, 0)][c])
    W = np.zeros((L + 1, 2, Ls, Ls))
    lsum = 0.0
    for c in (0, 1):
        Ar = A[c] + ridge * np.trace(A[c]) / d * np.eye(d)
        X = np.linalg.solve(Ar, B[c])            # X[(k, t)][s] = W_k[s][t]
        lsum += 0.5 * (vv[c] - 2 * np.sum(X * B[c]) + np.sum(X * (A[c] @ X)))
        for k in range(L + 1):
            W[k, c] = X[k * Ls:(k + 1) * Ls, :].T
    return W, lsum / (0.5 * vv.sum())


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
