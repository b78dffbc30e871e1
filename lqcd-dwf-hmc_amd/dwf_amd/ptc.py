"""The learned preconditioner: a gauge-covariant parallel-transport
convolution network (libdwfhip's dwf_ptc_*), its training and its use in
flexible CG.

The network N maps an even 5D field r to N(r) ~ (D^+D)^-1 r; it is linear in r
(no activation), so it is a learned linear preconditioner that commutes with
gauge transformations.  Training minimises sum_k |N(b_k) - x_k|^2 over
(source, solution) pairs of the operator, b_k = D^+D x_k, with Adam on FP32
master weights; the device holds bfloat16 copies, and the loss and gradient
come from the network's own forward and adjoint kernels on the GPU.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from . import ASolver_CG, Field

NPATH = 11


def _fp(a):
    return a.ctypes.data_as(C.POINTER(C.c_float))


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


class ParallelTransportNet:
    """N = L_{n-1} ... L_0 with channels (1, c_1, ..., 1) on the FP32 operator fopr32."""

    def __init__(self, fopr32, channels=(1, 8, 8, 8, 1)):
        self.fopr = fopr32
        self.channels = tuple(int(c) for c in channels)
        ch = (C.c_int * len(self.channels))(*self.channels)
        h = C.c_void_p()
        _lib.check("dwf_ptc_create", _lib.load().dwf_ptc_create(
            fopr32._h, len(self.channels) - 1, ch, C.byref(h)))
        self._h = h.value
        self.nw = int(_lib.load().dwf_ptc_num_weights(self._h))

    def __del__(self):
        if getattr(self, "_h", None):
            try:
                _lib.load().dwf_ptc_destroy(self._h)
            except Exception:
                pass
            self._h = None

    def layout(self):
        """[(offset, c_in, c_out)] per layer in the flat weight vector."""
        out, off = [], 0
        for l in range(len(self.channels) - 1):
            cin, cout = self.channels[l], self.channels[l + 1]
            out.append((off, cin, cout))
            off += NPATH * cin * cout * 32
        return out

    def identity_weights(self, scale=1.0, noise=0.0, seed=0):
        """Channel 0 carried through every layer by the identity path, times
        scale in the last layer (N = scale * 1); other entries ~ noise."""
        rng = np.random.default_rng(seed)
        w = (noise * rng.standard_normal(self.nw)).astype(np.float32)
        lay = self.layout()
        for l, (off, cin, cout) in enumerate(lay):
            s = scale if l == len(lay) - 1 else 1.0
            base = off   # q = 0, a = 0, b = 0
            for i in range(4):
                w[base + 2 * (i * 4 + i)] = s
        return w

    def set_weights(self, w):
        w = np.ascontiguousarray(w, np.float32)
        assert w.size == self.nw
        _lib.check("dwf_ptc_set_weights", _lib.load().dwf_ptc_set_weights(self._h, _fp(w)))

    def get_weights(self):
        w = np.empty(self.nw, np.float32)
        _lib.check("dwf_ptc_get_weights", _lib.load().dwf_ptc_get_weights(self._h, _fp(w)))
        return w

    def apply(self, out, inp):
        _lib.check("dwf_ptc_apply", _lib.load().dwf_ptc_apply(self._h, out.ptr, inp.ptr))

    def loss_grad(self, bs, xs):
        n = len(bs)
        ba = (C.c_void_p * n)(*[b.ptr for b in bs])
        xa = (C.c_void_p * n)(*[x.ptr for x in xs])
        loss = np.zeros(1)
        grad = np.zeros(self.nw)
        _lib.check("dwf_ptc_loss_grad", _lib.load().dwf_ptc_loss_grad(
            self._h, n, ba, xa, _dp(loss), _dp(grad)))
        return float(loss[0]), grad


def operator_pairs(fopr32, n, seed):
    """n (source, solution) pairs (b = D^+D x / |D^+D x|, x / |D^+D x|), x with
    uniform random entries, on the GPU (fopr32 in mode DdagD)."""
    from . import fill_uniform
    out = []
    for k in range(n):
        x, b = fopr32.new_field(), fopr32.new_field()
        fill_uniform(x, seed * 1000003 + k)
        fopr32.mult(b, x)
        s = 1.0 / np.sqrt(b.norm2())
        b.scal(s)
        x.scal(s)
        out.append((b, x))
    return out


def solver_pairs(fopr32, n, seed, tol=1e-12, maxit=20000):
    """n (source, solution) pairs from CG solves of D^+D x = b for random b."""
    from . import fill_uniform
    out = []
    cg = ASolver_CG(fopr32, maxit, tol)
    for k in range(n):
        x, b = fopr32.new_field(), fopr32.new_field()
        fill_uniform(b, seed * 1000003 + k)
        b.scal(1.0 / np.sqrt(b.norm2()))
        cg.solve(x, b)
        out.append((b, x))
    return out


class Adam:
    """Adam on FP32 master weights (Kingma and Ba, arXiv:1412.6980)."""

    def __init__(self, w, lr=1e-3, beta1=0.9, beta2=0.999, eps=1e-8):
        self.w = np.asarray(w, np.float64).copy()
        self.m = np.zeros_like(self.w)
        self.v = np.zeros_like(self.w)
        self.lr, self.b1, self.b2, self.eps, self.t = lr, beta1, beta2, eps, 0

    def step(self, g):
        self.t += 1
        self.m = self.b1 * self.m + (1 - self.b1) * g
        self.v = self.b2 * self.v + (1 - self.b2) * g * g
        mh = self.m / (1 - self.b1 ** self.t)
        vh = self.v / (1 - self.b2 ** self.t)
        self.w -= self.lr * mh / (np.sqrt(vh) + self.eps)
        return self.w.astype(np.float32)


def train(net, pairs, steps, batch=4, lr=2e-3, seed=0, log=None):
    """Adam over random batches of the pairs; returns the loss per step."""
    rng = np.random.default_rng(seed)
    opt = Adam(net.get_weights(), lr=lr)
    hist = []
    for it in range(steps):
        idx = rng.choice(len(pairs), size=min(batch, len(pairs)), replace=False)
        loss, g = net.loss_grad([pairs[i][0] for i in idx], [pairs[i][1] for i in idx])
        net.set_weights(opt.step(g / len(idx)))
        hist.append(loss / len(idx))
        if log and (it % 50 == 0 or it == steps - 1):
            log(f"  step {it}: loss {loss / len(idx):.5f}")
    return hist


class ASolver_FCG_PTC(ASolver_CG):
    """Flexible CG on fopr (mode DdagD) preconditioned by the network."""

    def __init__(self, fopr, net, maximum_number_of_iteration, convergence_criterion_squared):
        self.fopr = fopr
        self.net = net
        h = C.c_void_p()
        _lib.check("dwf_solver_create_fcg_ptc", _lib.load().dwf_solver_create_fcg_ptc(
            fopr._h, net._h, int(maximum_number_of_iteration),
            float(convergence_criterion_squared), C.byref(h)))
        self._h = h.value


class ASolver_Richardson_PTC(ASolver_CG):
    """FP64 Richardson around the network-preconditioned flexible CG on fopr_f."""

    def __init__(self, fopr_d, fopr_f, net, maximum_number_of_iteration,
                 convergence_criterion_squared, inner_iteration, inner_convergence_criterion_squared):
        self.fopr = fopr_d
        self.fopr_f = fopr_f
        self.net = net
        h = C.c_void_p()
        _lib.check("dwf_solver_create_richardson_ptc", _lib.load().dwf_solver_create_richardson_ptc(
            fopr_d._h, fopr_f._h, net._h, int(maximum_number_of_iteration),
            float(convergence_criterion_squared), int(inner_iteration),
            float(inner_convergence_criterion_squared), C.byref(h)))
        self._h = h.value
