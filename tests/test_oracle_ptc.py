"""Pins the oracle's parallel-transport network (CPU only).

* identity weights (channel 0 through the identity path) give N(b) = b;
* each layer's adjoint is the adjoint: <u, L v> = <L^+ u, v>;
* gauge covariance: with U_mu(x) -> g(x) U_mu(x) g(x + mu)^+ and b -> g b, the
  output is g N(b) (boundary phases included);
* the analytic gradient equals central finite differences of the loss, and a
  planted mistake (the weight gradient conjugated) fails it.
"""
import numpy as np
import pytest

import ptc_oracle as po
from dwf_amd import lattice

DIMS, LS = (2, 2, 2, 4), 4
CH = (1, 2, 2, 1)


def setup(seed=1):
    U = lattice.random_gauge(DIMS, 0.6, seed)
    lat = po.Lattice(DIMS, U, LS)
    rng = np.random.default_rng(seed)
    _, nw = po.weight_layout(CH)
    w = po.bf16_round(0.3 * rng.standard_normal(nw))
    return U, lat, w, rng


def test_identity_weights_pass_the_input():
    U, lat, w, rng = setup()
    w = np.zeros_like(w)
    for off, cin, cout in po.weight_layout(CH)[0]:
        for i in range(4):
            w[off + 2 * (i * 4 + i)] = 1.0
    b = lattice.random_spinor5(DIMS, LS, 3)
    assert np.abs(po.apply(lat, w, CH, b) - b).max() < 1e-14


def test_layer_adjoint():
    U, lat, w, rng = setup(2)
    W = po.unpack(w, CH)[1]
    shape = (2,) + lat.shape + (LS, 4, 3)
    v = rng.standard_normal(shape) + 1j * rng.standard_normal(shape)
    u = rng.standard_normal(shape) + 1j * rng.standard_normal(shape)
    lhs = np.vdot(u, po.layer(lat, W, v))
    rhs = np.vdot(po.layer_adjoint(lat, W, u), v)
    assert abs(lhs - rhs) < 1e-12 * abs(lhs)


def test_gauge_covariance():
    U, lat, w, rng = setup(3)
    V = lattice.volume(DIMS)
    g = lattice.as_su3(lattice.random_gauge(DIMS, 1.0, 9))[0]      # (V, 3, 3)
    Uc = lattice.as_su3(U.copy())
    Nx, Ny, Nz, Nt = DIMS
    idx = np.arange(V)
    x, y, z, t = idx % Nx, (idx // Nx) % Ny, (idx // (Nx * Ny)) % Nz, idx // (Nx * Ny * Nz)
    Ug = np.empty_like(Uc)
    for mu in range(4):
        c = [x, y, z, t]
        c[mu] = (c[mu] + 1) % DIMS[mu]
        nb = c[0] + Nx * (c[1] + Ny * (c[2] + Nz * c[3]))
        Ug[mu] = g @ Uc[mu] @ np.conj(np.swapaxes(g[nb], -1, -2))
    latg = po.Lattice(DIMS, lattice.from_su3(Ug), LS)
    b = lattice.random_spinor5(DIMS, LS, 4)
    # g acts on the colour of every (s, spin) of the even sites
    even = lat.even_idx[0]
    lex_even = even[3] + Nx * (even[2] + Ny * (even[1] + Nz * even[0]))
    bs = lattice.as_spinor(b.copy())                        # (Ls, Vh, 4, 3)
    gb = np.einsum("vab,svdb->svda", g[lex_even], bs)
    out = lattice.as_spinor(po.apply(lat, w, CH, b).copy())
    outg = lattice.as_spinor(po.apply(latg, w, CH, lattice.from_spinor(gb)).copy())
    want = np.einsum("vab,svdb->svda", g[lex_even], out)
    assert np.abs(outg - want).max() < 1e-12 * np.abs(want).max()


@pytest.mark.parametrize("plant", [False, True])
def test_gradient_matches_finite_differences(plant):
    U, lat, w, rng = setup(4)
    pairs = [(lattice.random_spinor5(DIMS, LS, 10 + k), lattice.random_spinor5(DIMS, LS, 20 + k))
             for k in range(2)]
    L0, G = po.loss_grad(lat, w, CH, pairs)
    if plant:   # a plausible mistake: the weight gradient conjugated
        Ws = po.unpack(G, CH)
        G = po.pack([np.conj(W) for W in Ws], CH)
    h = 1e-6
    ok = True
    for i in rng.choice(w.size, size=12, replace=False):
        wp, wm = w.copy(), w.copy()
        wp[i] += h
        wm[i] -= h
        fd = (po.loss(lat, wp, CH, pairs) - po.loss(lat, wm, CH, pairs)) / (2 * h)
        ok = ok and abs(fd - G[i]) <= 1e-6 * max(1.0, abs(fd))
    assert ok != plant
