"""Pins the oracle's solvers and stout smearing (CPU only).

Solvers are checked by the residual recomputed outside the solver (through the
operator alone), by agreement between solvers that must give the same answer
(CG from RHS/zero guesses, the sigma_0 multishift solution, Richardson with a
float inner CG), and by the reference's failure behaviour (nconv = -1).
Smearing is checked against the 32-term Taylor exponential the reference uses
as its own brute-force check (aprojection_Stout_SU3-tmpl.h:980-1018), against
scipy's expm, by unitarity, gauge covariance and the rho = 0 identity; the gauge
reader and plaquette against the plaquette the reference's sample runs print
for conf_04040408.txt.
"""
import json
import os

import numpy as np
import pytest
import scipy.linalg

from oracle import OracleDWF, richardson, stout_smear, plaquette, exp_iQ
from dwf_amd import lattice

GOLD = os.path.join(os.path.dirname(__file__), "golden")
DIMS, LS = (4, 4, 4, 4), 6


def rel(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


@pytest.fixture(scope="module")
def setup():
    U = lattice.random_gauge(DIMS, 0.5, 21)
    op = OracleDWF(DIMS, LS, 0.05, 1.6, 1.0, 0.0)
    op.set_config(U)
    b = lattice.random_spinor5(DIMS, LS, 22)
    return op, U, b


def test_cg_true_residual(setup):
    op, U, b = setup
    x, nconv, diff = op.cg(b, "DdagD", niter=2000, stop_cond2=1e-24)
    r = op.mult("DdagD", x) - b
    true = np.linalg.norm(r) / np.linalg.norm(b)
    assert nconv > 0
    assert true < 1e-11
    assert abs(true - diff) < 1e-13
    x0, n0, d0 = op.cg(b, "DdagD", niter=2000, stop_cond2=1e-24, init="ZERO")
    assert rel(x0, x) < 1e-9


def test_cg_not_converged_reports_minus_one(setup):
    op, U, b = setup
    x, nconv, diff = op.cg(b, "DdagD", niter=3, stop_cond2=1e-30)
    assert nconv == -1
    assert diff > 1e-3


def test_multishift_against_operator(setup):
    op, U, b = setup
    sigma = np.array([0.01, 0.05, 0.3, 2.0])
    xs, nconv, diff, rc = op.shift_cg(b, sigma, "DdagD", niter=2000, stop_cond2=1e-24)
    assert rc == 0
    worst = 0
    for s, x in zip(sigma, xs):
        r = op.mult("DdagD", x) + s * x - b
        worst = max(worst, np.linalg.norm(r) / np.linalg.norm(b))
    assert worst < 1e-10
    assert abs(worst - diff) < 1e-12
    # the base system sigma_0 alone equals a CG solve of (A + sigma_0); check
    # through the operator on a shift-0 problem as well
    xs0, _, _, _ = op.shift_cg(b, np.array([0.0]), "DdagD", niter=2000, stop_cond2=1e-24)
    xcg, _, _ = op.cg(b, "DdagD", niter=2000, stop_cond2=1e-24, init="ZERO")
    assert rel(xs0[0], xcg) < 1e-9


def test_richardson_mixed_precision(setup):
    op, U, b = setup
    opf = OracleDWF(DIMS, LS, 0.05, 1.6, 1.0, 0.0, precision="float")
    opf.set_config(U)
    x, nconv, diff, inner = richardson(op, opf, b, "DdagD", niter_outer=100,
                                       stop_outer2=1e-28, niter_inner=2000, stop_inner2=1e-8)
    true = np.linalg.norm(op.mult("DdagD", x) - b) / np.linalg.norm(b)
    assert nconv >= 2 and inner > 0
    assert true < 1e-13
    assert abs(true - diff) < 1e-14
    xcg, _, _ = op.cg(b, "DdagD", niter=2000, stop_cond2=1e-26)
    assert rel(x, xcg) < 1e-9


def random_su3_algebra(rng, scale):
    h = lattice.random_hermitian_traceless(rng, 1)[0] * scale
    return 1j * h   # anti-hermitian traceless


@pytest.mark.parametrize("scale", [1e-3, 0.1, 0.7, 2.0])
def test_exp_iQ_cayley_hamilton(scale):
    rng = np.random.default_rng(int(scale * 1000) + 3)
    for _ in range(5):
        iQ = random_su3_algebra(rng, scale)
        arr = np.ascontiguousarray(iQ).view(np.float64).reshape(18)
        e = exp_iQ(arr).view(np.complex128).reshape(3, 3)
        et = exp_iQ(arr, taylor=True).view(np.complex128).reshape(3, 3)
        ref = scipy.linalg.expm(iQ)
        assert np.abs(e - ref).max() < 1e-13
        assert np.abs(et - ref).max() < 1e-13
        assert np.abs(e @ np.conj(e.T) - np.eye(3)).max() < 1e-13
        assert abs(np.linalg.det(e) - 1) < 1e-13


def test_stout_smear_properties():
    dims = (4, 4, 2, 4)
    U = lattice.random_gauge(dims, 0.5, 31)
    # rho = 0 is the identity
    assert np.abs(stout_smear(dims, 0.0, U) - U).max() < 1e-15
    Us = stout_smear(dims, 0.1, U)
    M = lattice.as_su3(Us)
    assert np.abs(M @ np.conj(np.swapaxes(M, -1, -2)) - np.eye(3)).max() < 1e-13
    assert np.abs(np.linalg.det(M) - 1).max() < 1e-13
    assert plaquette(dims, Us) > plaquette(dims, U)
    # unit field is a fixed point
    I = lattice.unit_gauge(dims)
    assert np.abs(stout_smear(dims, 0.1, I) - I).max() < 1e-15


def test_stout_smear_gauge_covariance():
    dims = (4, 2, 4, 2)
    V = lattice.volume(dims)
    U = lattice.random_gauge(dims, 0.5, 41)
    g = lattice.as_su3(lattice.random_gauge(dims, 1.0, 42))[0]
    Nx, Ny, Nz, Nt = dims
    site = np.arange(V)
    coords = [site % Nx, (site // Nx) % Ny, (site // (Nx * Ny)) % Nz, site // (Nx * Ny * Nz)]

    def transform(Uf):
        Um = lattice.as_su3(Uf)
        out = np.empty_like(Um)
        for mu in range(4):
            c2 = [a.copy() for a in coords]
            c2[mu] = (c2[mu] + 1) % dims[mu]
            nb = c2[0] + Nx * (c2[1] + Ny * (c2[2] + Nz * c2[3]))
            out[mu] = g @ Um[mu] @ np.conj(np.swapaxes(g[nb], -1, -2))
        return lattice.from_su3(out)
    a = stout_smear(dims, 0.1, transform(U))
    b = transform(stout_smear(dims, 0.1, U))
    assert np.abs(a - b).max() < 1e-13


def test_known_answer_plaquette_conf_04040408():
    """Reader + plaquette against the value the reference's 8^4 runs print."""
    known = json.load(open(os.path.join(GOLD, "known_answers.json")))
    data = np.load(os.path.join(GOLD, "conf_04040408.npy"))
    U = lattice.tile_4x4x4x8(data, (8, 8, 8, 8))
    p = plaquette((8, 8, 8, 8), U)
    assert abs(p - known["plaq_init_8x8x8x8_from_conf_04040408"]["value"]) < 1e-12
    assert abs(p - known["plaq_init_8x8x8x8_sample_1gpu"]["value"]) < 1e-12
