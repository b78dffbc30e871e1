"""Pins the oracle's DWF fermion force by finite differences (CPU only).

For S(U) = |D[U] eta|^2 with D the even-odd preconditioned operator, the
reference's force_core (force_udiv in mode D followed by mult_generator,
-2 at(U F); aforce_F_Domainwall_5din_eo-tmpl.h:299-315,
lib_alt_QXS/Force/Fermion/aforce_F.cpp:97-139) satisfies, for the link
variation U_mu(x) -> exp(eps X) U_mu(x) with X anti-hermitian traceless,

    dS/deps = -Re tr(X F_mu(x)).

The H-mode pair udiv1_H(H eta, eta) + udiv1_Hdag(eta, H eta) gives the same
force (|H eta| = |D eta|), which pins the gamma_5 R handling.
"""
import numpy as np
import pytest
import scipy.linalg

from oracle import OracleDWF
from dwf_amd import lattice

PARAMS = [
    ((4, 2, 2, 4), 4, 0.05, 1.6, 1.0, 0.0, (1, 1, 1, -1)),
    ((2, 4, 2, 4), 5, 0.1, 1.0, 1.5, 0.5, (1, 1, 1, -1)),
    ((4, 2, 4, 2), 4, 0.3, 1.3, 1.2, 0.3, (-1, 1, -1, -1)),
]


def action(dims, Ls, mq, M0, b, c, bc, U, eta):
    op = OracleDWF(dims, Ls, mq, M0, b, c, bc=bc)
    op.set_config(U)
    v = op.mult("D", eta)
    return float(np.sum(v * v))


@pytest.mark.parametrize("prm", PARAMS)
def test_force_core_matches_finite_differences(prm):
    dims, Ls, mq, M0, b, c, bc = prm
    rng = np.random.default_rng(3)
    U = lattice.random_gauge(dims, 0.5, 4)
    eta = lattice.random_spinor5(dims, Ls, 5)
    op = OracleDWF(dims, Ls, mq, M0, b, c, bc=bc)
    op.set_config(U)
    F = op.force_core(eta)
    Fm = lattice.as_su3(F)
    V = lattice.volume(dims)
    eps = 1e-5
    for _ in range(6):
        mu = int(rng.integers(4))
        x = int(rng.integers(V))
        X = 1j * lattice.random_hermitian_traceless(rng, 1)[0]
        E = scipy.linalg.expm(eps * X)
        Em = scipy.linalg.expm(-eps * X)
        Up, Um = lattice.as_su3(U.copy()), lattice.as_su3(U.copy())
        Up[mu, x] = E @ Up[mu, x]
        Um[mu, x] = Em @ Um[mu, x]
        Sp = action(dims, Ls, mq, M0, b, c, bc, lattice.from_su3(Up), eta)
        Sm = action(dims, Ls, mq, M0, b, c, bc, lattice.from_su3(Um), eta)
        fd = (Sp - Sm) / (2 * eps)
        an = -np.real(np.trace(X @ Fm[mu, x]))
        assert abs(fd - an) <= 1e-6 * max(1.0, abs(an)), (mu, x, fd, an)
    # the force is anti-hermitian and traceless
    assert np.abs(Fm + np.conj(np.swapaxes(Fm, -1, -2))).max() < 1e-12
    assert np.abs(np.trace(Fm, axis1=-2, axis2=-1)).max() < 1e-12


@pytest.mark.parametrize("prm", PARAMS[:2])
def test_force_H_modes_agree_with_D_modes(prm):
    dims, Ls, mq, M0, b, c, bc = prm
    U = lattice.random_gauge(dims, 0.5, 6)
    eta = lattice.random_spinor5(dims, Ls, 7)
    op = OracleDWF(dims, Ls, mq, M0, b, c, bc=bc)
    op.set_config(U)
    FD = op.force_core(eta)
    h = op.mult("H", eta)
    FH = op.force_core1("H", h, eta) + op.force_core1("Hdag", eta, h)
    assert np.abs(FH - FD).max() < 1e-11 * np.abs(FD).max()
