"""Pins the oracle's stout force recursion by finite differences (CPU only).

The fermion force on smeared links (AForce_F_Smeared_alt::force_udiv,
aforce_F_Smeared_alt-tmpl.h:84-104) is the force_udiv of the operator on the
top-level links U_n, pulled back through every smearing level by
mult_jacobian (AForceSmear_APE::force_udiv, Morningstar-Peardon), then
mult_generator with the thin links U_0.  For S(U_0) = |D[U_n(U_0)] eta|^2 the
result F satisfies dS/deps = -Re tr(X F_mu(x)) for U_mu(x) -> exp(eps X) U_mu(x),
the same relation as for the unsmeared force.
"""
import numpy as np
import pytest
import scipy.linalg

from oracle import (OracleDWF, stout_smear, stout_mult_jacobian, mult_generator_gauge,
                    stout_force_udiv, exp_iQ)
from dwf_amd import lattice

DIMS, LS = (4, 2, 4, 2), 4
OPP = (0.1, 1.6, 1.5, 0.5)


def smear_levels(U, rho, n):
    lv = [U]
    for _ in range(n):
        lv.append(stout_smear(DIMS, rho, lv[-1]))
    return lv


def action(U, eta, rho, n):
    o = OracleDWF(DIMS, LS, *OPP)
    o.set_config(smear_levels(U, rho, n)[-1])
    v = o.mult("D", eta)
    return float(np.sum(v * v))


@pytest.mark.parametrize("rho,nsmear", [(0.1, 1), (0.12, 2)])
def test_smeared_force_matches_finite_differences(rho, nsmear):
    U = lattice.random_gauge(DIMS, 0.4, 51)
    eta = lattice.random_spinor5(DIMS, LS, 52)
    lv = smear_levels(U, rho, nsmear)
    o = OracleDWF(DIMS, LS, *OPP)
    o.set_config(lv[-1])
    Fu = o.force_udiv(eta)
    F = lattice.as_su3(mult_generator_gauge(DIMS, stout_mult_jacobian(DIMS, rho, lv[:-1], Fu), U))
    rng = np.random.default_rng(53)
    V = lattice.volume(DIMS)
    eps = 1e-5
    for _ in range(6):
        mu, x = int(rng.integers(4)), int(rng.integers(V))
        X = 1j * lattice.random_hermitian_traceless(rng, 1)[0]
        Up, Um = lattice.as_su3(U.copy()), lattice.as_su3(U.copy())
        Up[mu, x] = scipy.linalg.expm(eps * X) @ Up[mu, x]
        Um[mu, x] = scipy.linalg.expm(-eps * X) @ Um[mu, x]
        fd = (action(lattice.from_su3(Up), eta, rho, nsmear) -
              action(lattice.from_su3(Um), eta, rho, nsmear)) / (2 * eps)
        an = -np.real(np.trace(X @ F[mu, x]))
        assert abs(fd - an) <= 1e-6 * max(1.0, abs(an)), (mu, x, fd, an)


def test_stout_force_with_zero_rho_is_identity():
    # rho = 0: U' = U, so the pulled-back force equals the input
    U = lattice.random_gauge(DIMS, 0.4, 54)
    rng = np.random.default_rng(55)
    S = rng.standard_normal(U.shape)
    assert np.abs(stout_force_udiv(DIMS, 0.0, S, U) - S).max() < 1e-15


def test_stout_force_small_Q_branch():
    # unit links: every staple sum is proportional to 1, iQ = 0; the recursion
    # takes the |iQ|^2 <= 1e-10 branch: Xi = Sigma', iTheta = 0 -> Sigma = Sigma'
    U = lattice.unit_gauge(DIMS)
    rng = np.random.default_rng(56)
    S = rng.standard_normal(U.shape)
    assert np.abs(stout_force_udiv(DIMS, 0.1, S, U) - S).max() < 1e-15
