"""GPU learned polynomial preconditioner against the oracle's restatement.

Bars: the hash fill bit-exact; Gram matrices 1e-12 (FP64, vector path) and
1e-5 relative (FP32 fields on the matrix cores, float accumulation); K, K^+,
M = K K^+ 1e-12 (FP64) / 1e-5 (FP32) / 5e-3 (FP16 storage); trained weights
1e-6 relative to the oracle's least squares on the same fields; the solvers by
their FP64 true residual.
"""
import numpy as np
import pytest

from oracle import (OracleDWF, lpc_K, lpc_Kdag, chiral_gram as orc_gram, uniform_field,
                    lpc_train)
import dwf_amd
from dwf_amd import lattice

pytestmark = pytest.mark.gpu

CASES = [
    ((4, 4, 2, 4), 6, 0.05, 1.6, 1.0, 0.0),     # Shamir: Mp spin-scalar
    ((4, 2, 4, 4), 8, 0.05, 1.6, 1.5, 0.5),     # Moebius: Mp not spin-scalar
    ((8, 8, 8, 8), 16, 0.02, 1.6, 1.5, 0.5),
]


def rel(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


def ops(case, precisions=("double",)):
    dims, Ls, mq, M0, b, c = case
    U = lattice.random_gauge(dims, 0.4, 91)
    orc = OracleDWF(dims, Ls, mq, M0, b, c)
    orc.set_config(U)
    g = {}
    for p in precisions:
        g[p] = dwf_amd.AFopr_Domainwall_5din_eo(dims, Ls, mq, M0, b, c, precision=p)
        g[p].set_config(U)
    return orc, g


@pytest.mark.parametrize("prec", ["double", "float", "half"])
def test_fill_uniform_bit_exact(prec):
    case = CASES[1]
    orc, g = ops(case, (prec,))
    op = g[prec]
    f = op.new_field()
    dwf_amd.fill_uniform(f, 0x1234567887654321)
    want = uniform_field(0x1234567887654321, op.Vh, op.Ls)
    # the kernel rounds double -> compute type (float for the 16-bit storage) -> storage
    if prec == "double":
        stored = want
    elif prec == "float":
        stored = want.astype(np.float32)
    else:
        stored = want.astype(np.float32).astype(np.float16)
    assert np.array_equal(op.reverse(f), stored.astype(np.float64))


@pytest.mark.parametrize("prec,tol", [("double", 1e-12), ("float", 1e-5)])
@pytest.mark.parametrize("nf", [1, 3, 6])
def test_chiral_gram(prec, tol, nf):
    case = CASES[2]
    print("case", case, prec, nf, flush=True)
    orc, g = ops(case, (prec,))
    op = g[prec]
    xs = [lattice.random_spinor5(case[0], case[1], 100 + i) for i in range(nf)]
    fs = []
    for x in xs:
        f = op.new_field()
        op.convert(f, x)
        fs.append(f)
    got = dwf_amd.chiral_gram(op, fs)
    xr = [op.reverse(f) for f in fs]    # the stored (rounded) values
    want = orc_gram(xr)
    assert rel(got, want) < tol


@pytest.mark.parametrize("case", CASES)
@pytest.mark.parametrize("prec,tol", [("double", 1e-12), ("float", 1e-5), ("half", 5e-3)])
def test_apply_matches_oracle(case, prec, tol):
    print("case", case, prec, flush=True)
    orc, g = ops(case, (prec,))
    op = g[prec]
    L = 3
    P = dwf_amd.APrecond_LearnedPolynomial(op, L)
    rng = np.random.default_rng(92)
    W = np.eye(case[1])[None, None] + 0.3 * rng.standard_normal(P.weights_shape)
    P.set_weights(W)
    assert np.array_equal(P.get_weights(), W)
    x = lattice.random_spinor5(case[0], case[1], 93)
    fx, fy = op.new_field(), op.new_field()
    op.convert(fx, x)
    xr = op.reverse(fx)
    P.apply("K", fy, fx)
    assert rel(op.reverse(fy), lpc_K(orc, W, xr)) < tol
    P.apply("Kdag", fy, fx)
    assert rel(op.reverse(fy), lpc_Kdag(orc, W, xr)) < tol
    P.apply("M", fy, fx)
    assert rel(op.reverse(fy), lpc_K(orc, W, lpc_Kdag(orc, W, xr))) < tol * 3


@pytest.mark.parametrize("case", CASES[:2])
def test_training_matches_oracle(case):
    orc, g = ops(case, ("double",))
    op = g["double"]
    L, nvec, seed = 2, 2, 777
    P = dwf_amd.APrecond_LearnedPolynomial(op, L)
    loss = P.train(op, nvec=nvec, seed=seed, ridge=0.0)
    M64 = (1 << 64) - 1
    vs = [uniform_field((seed + j * 0xD1B54A32D192ED03) & M64, op.Vh, op.Ls) for j in range(nvec)]
    W, want_loss = lpc_train(orc, L, vs, ridge=0.0)
    Wg = P.get_weights()
    assert np.abs(Wg - W).max() < 1e-6 * np.abs(W).max()
    assert abs(loss - want_loss) < 1e-6 * want_loss
    # the fit beats the truncated Neumann series of the same degree
    I = np.array([np.stack([np.eye(op.Ls)] * 2)] * (L + 1))
    v = vs[0]
    res_fit = np.linalg.norm(lpc_K(orc, W, orc.mult("D", v)) - v)
    res_neu = np.linalg.norm(lpc_K(orc, I, orc.mult("D", v)) - v)
    assert res_fit < res_neu


def test_fcg_and_mixed_solver_converge():
    case = CASES[2]
    dims, Ls = case[0], case[1]
    orc, g = ops(case, ("double", "float", "half"))
    b = lattice.random_spinor5(dims, Ls, 94)
    want, _, _ = orc.cg(b, "DdagD", 10000, 1e-24, "ZERO")
    for op in g.values():
        op.set_mode("DdagD")
    # FP64 flexible CG with an FP64 preconditioner
    P64 = dwf_amd.APrecond_LearnedPolynomial(g["double"], 2)
    P64.train(g["double"], nvec=2, seed=5)
    fb, fx = g["double"].new_field(), g["double"].new_field()
    g["double"].convert(fb, b)
    s = dwf_amd.ASolver_FCG(g["double"], P64, 5000, 1e-24)
    nconv, diff = s.solve(fx, fb, raise_on_nonconvergence=True)
    assert diff < 1e-11 and nconv > 0
    assert rel(g["double"].reverse(fx), want) < 1e-9
    # FP64 Richardson, FP32 flexible CG, FP16 preconditioner trained in FP32
    P16 = dwf_amd.APrecond_LearnedPolynomial(g["half"], 2)
    P16.train(g["float"], nvec=2, seed=6)
    s2 = dwf_amd.ASolver_Richardson_LPC(g["double"], g["float"], P16, 100, 1e-20, 10000, 1e-6)
    nconv2, diff2 = s2.solve(fx, fb, raise_on_nonconvergence=True)
    assert diff2 < 1e-10
    assert rel(g["double"].reverse(fx), want) < 1e-8
