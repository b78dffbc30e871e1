"""CPU restatement (numpy, FP64) of the parallel-transport convolution network
used as the learned preconditioner.

TEST INFRASTRUCTURE ONLY: imported by tests/ as the checker of the GPU
network kernels, never by the product path.

The network (no counterpart in the reference) acts on full-lattice 5D spinor
fields psi[x][s][spin][colour] (lexical sites) with channels:
    out[a](x, s) = sum_q sum_b W[q][a][b] T_q(in[b])(x, s),
    T_0 = 1,
    T_{1 + 2 mu}(psi)(x) = U_mu(x) psi(x + mu)         (times bc_mu on x_mu = N_mu - 1)
    T_{2 + 2 mu}(psi)(x) = U_mu(x - mu)^+ psi(x - mu)  (times bc_mu on x_mu = 0)
    T_9(psi)(x, s) = psi(x, s + 1),  T_10(psi)(x, s) = psi(x, s - 1)  (zero beyond 0, Ls - 1),
W[q][a][b] a complex 4x4 matrix on spin.  The boundary factors are those of the
even-odd operator's links (afopr_Domainwall_5din_eo-tmpl.h:367-424 set_config:
the phase on the last slice of the forward link).  Input: an even 5D field
(corelib Field(24, V/2, Ls), odd parity zero); output: the even part of the last
layer's single channel.  The gradient is the adjoint (reversed transports, W^+)
chain, the weight gradient sum g conj(T in) over the lattice.
"""
from __future__ import annotations

import numpy as np

NPATH = 11


def weight_layout(channels):
    """[(offset, c_in, c_out)] of each layer in the flat weight vector."""
    out, off = [], 0
    for l in range(len(channels) - 1):
        out.append((off, channels[l], channels[l + 1]))
        off += NPATH * channels[l] * channels[l + 1] * 32
    return out, off


def unpack(wflat, channels):
    """flat float vector -> list of complex arrays W[l] of shape (11, c_out, c_in, 4, 4)."""
    lay, _ = weight_layout(channels)
    Ws = []
    for off, cin, cout in lay:
        n = NPATH * cout * cin * 32
        w = np.asarray(wflat[off:off + n], np.float64).reshape(NPATH, cout, cin, 16, 2)
        Ws.append((w[..., 0] + 1j * w[..., 1]).reshape(NPATH, cout, cin, 4, 4))
    return Ws


def pack(Ws, channels):
    lay, total = weight_layout(channels)
    flat = np.zeros(total)
    for (off, cin, cout), G in zip(lay, Ws):
        n = NPATH * cout * cin * 32
        g = G.reshape(NPATH, cout, cin, 16)
        flat[off:off + n] = np.stack([g.real, g.imag], axis=-1).reshape(-1)
    return flat


class Lattice:
    """Geometry, links and the even-odd <-> lexical maps of one lattice."""

    def __init__(self, dims, U, Ls, bc=(1, 1, 1, -1)):
        self.dims = tuple(int(v) for v in dims)
        Nx, Ny, Nz, Nt = self.dims
        self.Ls = int(Ls)
        self.shape = (Nt, Nz, Ny, Nx)
        M = np.asarray(U, np.float64).reshape(4, -1, 3, 3, 2)
        self.U = (M[..., 0] + 1j * M[..., 1]).reshape((4,) + self.shape + (3, 3))
        self.bc = bc
        t, z, y, x = np.meshgrid(np.arange(Nt), np.arange(Nz), np.arange(Ny), np.arange(Nx),
                                 indexing="ij")
        self.coord = (x, y, z, t)
        # even half-site index -> lexical coordinates (index_eo.cpp:125-156)
        x2, y2, z2, t2 = np.meshgrid(np.arange(Nx // 2), np.arange(Ny), np.arange(Nz),
                                     np.arange(Nt), indexing="ij")
        self.even_idx = []
        for par in (0, 1):
            xs = 2 * x2 + (y2 + z2 + t2 + par) % 2
            # half-site order x2 fastest
            order = np.argsort((x2 + (Nx // 2) * (y2 + Ny * (z2 + Nz * t2))).reshape(-1))
            self.even_idx.append(tuple(a.reshape(-1)[order] for a in (t2, z2, y2, xs)))

    def axis(self, mu):
        return {0: 3, 1: 2, 2: 1, 3: 0}[mu]

    def transport(self, psi, q):
        """T_q on psi (Nt, Nz, Ny, Nx, Ls, 4, 3)."""
        if q == 0:
            return psi.copy()
        if q >= 9:
            out = np.zeros_like(psi)
            if q == 9:
                out[..., :-1, :, :] = psi[..., 1:, :, :]
            else:
                out[..., 1:, :, :] = psi[..., :-1, :, :]
            return out
        d = q - 1
        mu, fb = d // 2, d % 2
        ax = self.axis(mu)
        N = self.dims[mu]
        c = self.coord[mu]
        if fb == 0:
            src = np.roll(psi, -1, axis=ax)
            link = self.U[mu]
            ph = np.where(c == N - 1, float(self.bc[mu]), 1.0)
        else:
            src = np.roll(psi, 1, axis=ax)
            link = np.conj(np.swapaxes(np.roll(self.U[mu], 1, axis=ax), -1, -2))
            ph = np.where(c == 0, float(self.bc[mu]), 1.0)
        out = np.einsum("...ab,...sdb->...sda", link, src)
        return out * ph[..., None, None, None]

    def embed_even(self, even):
        """corelib even field (Ls, V/2, 24) -> full lexical complex field, odd zero."""
        e = np.asarray(even, np.float64).reshape(self.Ls, -1, 4, 3, 2)
        ec = e[..., 0] + 1j * e[..., 1]                      # (Ls, Vh, 4, 3)
        psi = np.zeros(self.shape + (self.Ls, 4, 3), np.complex128)
        psi[self.even_idx[0]] = np.transpose(ec, (1, 0, 2, 3))
        return psi

    def even_part(self, psi):
        ec = np.transpose(psi[self.even_idx[0]], (1, 0, 2, 3))   # (Ls, Vh, 4, 3)
        return np.stack([ec.real, ec.imag], axis=-1).reshape(self.Ls, -1, 24)


def layer(lat, W, act):
    """act: (c_in, ...) complex -> (c_out, ...)"""
    cout = W.shape[1]
    out = np.zeros((cout,) + act.shape[1:], np.complex128)
    for q in range(NPATH):
        for b in range(act.shape[0]):
            t = lat.transport(act[b], q)
            for a in range(cout):
                out[a] += np.einsum("ij,...sjc->...sic", W[q, a, b], t)
    return out


def layer_adjoint(lat, W, g):
    """the adjoint of layer(): g (c_out, ...) -> (c_in, ...)"""
    rev = [0, 2, 1, 4, 3, 6, 5, 8, 7, 10, 9]
    cin = W.shape[2]
    out = np.zeros((cin,) + g.shape[1:], np.complex128)
    for q in range(NPATH):
        for a in range(g.shape[0]):
            t = lat.transport(g[a], rev[q])
            for b in range(cin):
                out[b] += np.einsum("ij,...sjc->...sic", np.conj(W[q, a, b]).T, t)
    return out


def forward(lat, Ws, even_in):
    acts = [lat.embed_even(even_in)[None]]
    for W in Ws:
        acts.append(layer(lat, W, acts[-1]))
    return acts


def apply(lat, wflat, channels, even_in):
    acts = forward(lat, unpack(wflat, channels), even_in)
    return lat.even_part(acts[-1][0])


def loss(lat, wflat, channels, pairs):
    tot = 0.0
    for b, x in pairs:
        y = apply(lat, wflat, channels, b)
        tot += float(np.sum((y - x) ** 2))
    return tot


def loss_grad(lat, wflat, channels, pairs):
    """(loss, d loss / d wflat) with the adjoint chain."""
    Ws = unpack(wflat, channels)
    grads = [np.zeros_like(W) for W in Ws]
    tot = 0.0
    for b, x in pairs:
        acts = forward(lat, Ws, b)
        y = lat.even_part(acts[-1][0])
        d = y - x
        tot += float(np.sum(d * d))
        g = np.zeros_like(acts[-1])
        g[0] = lat.embed_even(2.0 * d)
        for l in range(len(Ws) - 1, -1, -1):
            W = Ws[l]
            for q in range(NPATH):
                for bb in range(acts[l].shape[0]):
                    t = lat.transport(acts[l][bb], q)
                    for a in range(W.shape[1]):
                        # dL/dRe W + i dL/dIm W = sum g conj(t)^T  (g = dL/dRe out
                        # + i dL/dIm out: dL = Re(conj(g) dW t))
                        G = np.einsum("tzyxsic,tzyxsjc->ij", g[a], np.conj(t))
                        grads[l][q, a, bb] += G
            if l > 0:
                g = layer_adjoint(lat, W, g)
    return tot, pack(grads, channels)


def bf16_round(w):
    """round float64 values to bfloat16 (through float32, nearest even)"""
    f = np.asarray(w, np.float32).view(np.uint32).astype(np.uint64)
    r = ((f >> 16) & 1) + 0x7FFF
    return ((((f + r) >> 16) << 16).astype(np.uint32)).view(np.float32).astype(np.float64)
