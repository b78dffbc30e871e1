"""Pins the rational pieces of the oracle (CPU only).

* The Math_Rational coefficient files (src/lib/Tools/math_Rational.cpp:129-199)
  approximate x^(n/d) on [x_min, x_max] to the error the file states; the
  reader's (a0, residue, pole) order is pinned by that and by the a0 of
  parameter_rational.1_4_16_0.00010000_010.000.
* AFopr_Rational::mult (afopr_Rational-tmpl.h:181-219) agrees with r(A^dag A)
  evaluated through the eigen-decomposition of the dense A^dag A.
* AForce_F_Rational_Ratio::force_udiv (aforce_F_Rational_Ratio-tmpl.h:345-447)
  followed by mult_generator is the derivative of the dense action
  S = <phi| r1(A^dag A) r2(B^dag B) r1(A^dag A) |phi>: for the link variation
  U_mu(x) -> exp(eps X) U_mu(x), dS/deps = +Re tr(X F_mu(x)) (the sign of a
  pseudofermion action with an inverse power: S decreases where |A eta|^2 grows).
"""
import glob
import os

import numpy as np
import pytest
import scipy.linalg

from oracle import OracleDWF, rational_mult, rational_ratio_force_udiv
from dwf_amd import lattice

HERE = os.path.dirname(os.path.abspath(__file__))
RAT = os.path.join(HERE, "golden", "rational")


def rat(name):
    d = lattice.read_rational(os.path.join(RAT, name))
    return d["a0"], d["res"], d["pole"]


def r_eval(coef, x):
    a0, bl, cl = coef
    return a0 + np.sum(bl[None, :] / (x[:, None] + cl[None, :]), axis=1)


@pytest.mark.parametrize("path", sorted(glob.glob(os.path.join(RAT, "parameter_rational.*"))))
def test_rational_files_approximate_the_power(path):
    d = lattice.read_rational(path)
    name = os.path.basename(path)
    assert name == lattice.rational_filename(d["n"], d["d"], d["np"], d["x_min"], d["x_max"])
    x = np.exp(np.linspace(np.log(d["x_min"]), np.log(d["x_max"]), 2001))
    want = x ** (d["n"] / d["d"])
    got = r_eval((d["a0"], d["res"], d["pole"]), x)
    err = np.abs(got / want - 1).max()
    # (the 24-pole files state errors at the level of double rounding)
    assert err < max(1.5 * d["error"], 5e-14), (name, err, d["error"])
    # poles are positive and sorted ascending, so pole[0] is the base shift
    assert np.all(d["pole"] > 0) and np.all(np.diff(d["pole"]) > 0)


def test_rational_known_a0():
    a0, bl, cl = rat("parameter_rational.1_4_16_0.00010000_010.000")
    assert abs(a0 - 4.579418) < 1e-6
    assert len(bl) == 16 and abs(cl[-1] - 127.5) / 127.5 < 1e-3


def dense(orc, mode, n, shape):
    cols = np.empty((n, n))
    e = np.zeros(n)
    for k in range(n):
        e[k] = 1.0
        cols[:, k] = orc.mult(mode, e.reshape(shape)).reshape(-1)
        e[k] = 0.0
    return cols


def r_matrix(coef, A):
    lam, Q = np.linalg.eigh(A.T @ A)
    return (Q * r_eval(coef, lam)[None, :]) @ Q.T


DIMS, LS = (4, 2, 2, 2), 4
OP1 = dict(mq=0.5, M0=1.6, b=1.5, c=0.5)
OP2 = dict(mq=0.1, M0=1.6, b=1.5, c=0.5)
NUM = "parameter_rational.1_4_16_0.00010000_010.000"
DEN = "parameter_rational.1_2inv_16_0.00010000_010.000"


def make(U, p):
    o = OracleDWF(DIMS, LS, p["mq"], p["M0"], p["b"], p["c"])
    o.set_config(U)
    return o


def test_rational_mult_matches_eigendecomposition():
    U = lattice.random_gauge(DIMS, 0.5, 21)
    phi = lattice.random_spinor5(DIMS, LS, 22)
    o = make(U, OP2)
    n = phi.size
    A = dense(o, "D", n, phi.shape)
    coef = rat(DEN)
    want = r_matrix(coef, A) @ phi.reshape(-1)
    got, _ = rational_mult(o, *coef, phi)
    assert np.linalg.norm(got.reshape(-1) - want) < 1e-9 * np.linalg.norm(want)
    # and r2 ~ (A^dag A)^(-1/2) on the spectrum of this operator
    lam = np.linalg.eigvalsh(A.T @ A)
    assert lam.min() > 1e-4 and lam.max() < 10.0


def action(U, phi, c1, c2):
    n = phi.size
    A = dense(make(U, OP1), "D", n, phi.shape)
    B = dense(make(U, OP2), "D", n, phi.shape)
    v1 = r_matrix(c1, A) @ phi.reshape(-1)
    v2 = r_matrix(c2, B) @ v1
    return float(v1 @ v2)


def test_rational_ratio_force_matches_finite_differences():
    U = lattice.random_gauge(DIMS, 0.5, 23)
    phi = lattice.random_spinor5(DIMS, LS, 24)
    c1, c2 = rat(NUM), rat(DEN)
    o1, o2 = make(U, OP1), make(U, OP2)
    Fu = rational_ratio_force_udiv(o1, o2, c1, c2, phi)
    F = lattice.as_su3(o1.mult_generator(Fu))
    rng = np.random.default_rng(25)
    V = lattice.volume(DIMS)
    eps = 1e-5
    for _ in range(3):
        mu, x = int(rng.integers(4)), int(rng.integers(V))
        X = 1j * lattice.random_hermitian_traceless(rng, 1)[0]
        Up, Um = lattice.as_su3(U.copy()), lattice.as_su3(U.copy())
        Up[mu, x] = scipy.linalg.expm(eps * X) @ Up[mu, x]
        Um[mu, x] = scipy.linalg.expm(-eps * X) @ Um[mu, x]
        fd = (action(lattice.from_su3(Up), phi, c1, c2) -
              action(lattice.from_su3(Um), phi, c1, c2)) / (2 * eps)
        an = np.real(np.trace(X @ F[mu, x]))
        assert abs(fd - an) <= 1e-6 * max(1.0, abs(an)), (mu, x, fd, an)
