"""The parallel-transport network on the GPU against the oracle, its training,
and what it gains inside flexible CG.

Bars: the FP32 network output within 2e-5 relative of the FP64 oracle with
the same bfloat16 weights; the loss and the gradient (forward + adjoint
kernels) within 1e-4 relative.  The trained network must lower the flexible
CG's iteration count on a gauge field it was not trained on, against the same
solver with the identity network (= CG).
"""
import numpy as np
import pytest

import ptc_oracle as po
import dwf_amd
from dwf_amd import lattice
from dwf_amd.ptc import (ASolver_FCG_PTC, ParallelTransportNet, operator_pairs, train)

pytestmark = pytest.mark.gpu


def rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


def make(dims, Ls, seed, mq=0.05, M0=1.6, b=1.0, c=0.0):
    U = lattice.random_gauge(dims, 0.4, seed)
    op = dwf_amd.AFopr_Domainwall_5din_eo(dims, Ls, mq, M0, b, c, (1, 1, 1, -1), "float")
    op.set_config(U)
    op.set_mode("DdagD")
    return U, op


# the matrix-core layer kernel with 2, 3 and 4 tiles (4, 6 and 8 channels; 3 and
# 1 channels pad the last tile; at Ls = 16 a 16-point tile is one site, at Ls = 12
# tiles straddle sites), and the VALU kernel that large lattices use
# (4-, 2- and 1-channel groups); the weight gradient's 2- and 1-channel groups
@pytest.mark.parametrize("dims,Ls,ch,kernel", [((4, 2, 2, 4), 4, (1, 4, 4, 1), "matrix"),
                                               ((4, 2, 2, 4), 4, (1, 6, 3, 1), "matrix"),
                                               ((4, 4, 2, 4), 8, (1, 8, 8, 8, 1), "matrix"),
                                               ((2, 2, 2, 6), 5, (1, 4, 2, 1), "matrix"),
                                               ((4, 2, 2, 4), 12, (1, 4, 4, 1), "matrix"),
                                               ((4, 2, 2, 4), 16, (1, 8, 8, 1), "matrix"),
                                               ((4, 2, 2, 4), 4, (1, 6, 3, 1), "valu"),
                                               ((4, 4, 2, 4), 8, (1, 8, 8, 8, 1), "valu")])
def test_forward_matches_oracle(dims, Ls, ch, kernel, monkeypatch):
    print("case ptc forward", dims, Ls, ch, kernel, flush=True)
    if kernel == "valu":
        monkeypatch.setenv("DWF_PTC_KERNEL", "valu")
    U, op = make(dims, Ls, 1)
    net = ParallelTransportNet(op, ch)
    rng = np.random.default_rng(2)
    net.set_weights((0.2 * rng.standard_normal(net.nw)).astype(np.float32))
    w = net.get_weights().astype(np.float64)       # the bfloat16-rounded weights
    assert np.array_equal(w, po.bf16_round(w))
    x = lattice.random_spinor5(dims, Ls, 3)
    fi, fo = op.new_field(), op.new_field()
    op.convert(fi, x)
    net.apply(fo, fi)
    lat = po.Lattice(dims, U, Ls)
    want = po.apply(lat, w, ch, op.reverse(fi))
    assert rel(op.reverse(fo), want) < 2e-5


# 2x2x2x6 at Ls = 5: 120 points per parity, so the last 16-point tile is partial
@pytest.mark.parametrize("dims,Ls,ch,kernel", [((4, 2, 2, 4), 4, (1, 4, 4, 1), "matrix"),
                                               ((4, 2, 2, 4), 4, (1, 6, 3, 1), "matrix"),
                                               ((4, 2, 2, 4), 4, (1, 8, 8, 1), "matrix"),
                                               ((2, 2, 2, 6), 5, (1, 4, 2, 1), "matrix"),
                                               ((4, 2, 2, 4), 4, (1, 6, 3, 1), "valu"),
                                               ((4, 2, 2, 4), 4, (1, 8, 8, 1), "valu")])
def test_loss_and_gradient_match_oracle(dims, Ls, ch, kernel, monkeypatch):
    print("case ptc gradient", dims, Ls, ch, kernel, flush=True)
    if kernel == "valu":
        monkeypatch.setenv("DWF_PTC_KERNEL", "valu")
    U, op = make(dims, Ls, 4)
    net = ParallelTransportNet(op, ch)
    rng = np.random.default_rng(5)
    net.set_weights((0.2 * rng.standard_normal(net.nw)).astype(np.float32))
    w = net.get_weights().astype(np.float64)
    fb, fx, hb, hx = [], [], [], []
    for k in range(2):
        b, x = op.new_field(), op.new_field()
        op.convert(b, lattice.random_spinor5(dims, Ls, 10 + k))
        op.convert(x, lattice.random_spinor5(dims, Ls, 20 + k))
        fb.append(b)
        fx.append(x)
        hb.append(op.reverse(b))
        hx.append(op.reverse(x))
    loss, grad = net.loss_grad(fb, fx)
    lat = po.Lattice(dims, U, Ls)
    lo, go = po.loss_grad(lat, w, ch, list(zip(hb, hx)))
    assert abs(loss - lo) / lo < 1e-5
    assert rel(grad, go) < 1e-4


def test_training_lowers_fcg_iterations_on_held_out_links():
    dims, Ls, ch = (4, 4, 4, 8), 8, (1, 8, 8, 8, 1)
    U, op = make(dims, Ls, 7)
    net = ParallelTransportNet(op, ch)
    w0 = net.identity_weights(scale=0.5, noise=1e-3, seed=1)
    net.set_weights(w0)
    pairs = operator_pairs(op, 64, seed=3)
    hist = train(net, pairs, steps=150, batch=4, lr=2e-3, seed=2,
                 log=lambda m: print(m, flush=True))
    assert np.mean(hist[-10:]) < 0.7 * np.mean(hist[:5])
    w = net.get_weights()
    # held-out links, same action parameters
    U2, op2 = make(dims, Ls, 99)
    net2 = ParallelTransportNet(op2, ch)
    b = op2.new_field()
    op2.convert(b, lattice.random_spinor5(dims, Ls, 11))
    its = {}
    for name, ww in (("identity", w0), ("trained", w)):
        net2.set_weights(ww)
        x = op2.new_field()
        n, d = ASolver_FCG_PTC(op2, net2, 5000, 1e-12).solve(x, b)
        assert n > 0
        its[name] = n
    n_cg, _ = dwf_amd.ASolver_CG(op2, 5000, 1e-12).solve(op2.new_field(), b)
    print("  FCG iterations:", its, "CG:", n_cg, flush=True)
    assert its["trained"] < 0.9 * its["identity"]
