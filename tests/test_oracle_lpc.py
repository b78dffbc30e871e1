"""Pins the oracle's learned polynomial preconditioner (CPU only; the
preconditioner has no counterpart in the reference).

* With W_k = 1 the polynomial is the truncated Neumann series: K D = 1 - H^(L+1).
* K^+ is the adjoint of K: <u, K v> = <K^+ u, v> for random weights.
* W+ = 1, W- = -1 is gamma_5 (the chirality convention of the weights).
* The least-squares fit leaves a residual K D v - v orthogonal to every
  feature H^k D v in each chiral component (the normal equations), and
  recovers weights that reproduce the target exactly when it is representable.
* The counter hash of the random fields equals a plain-integer splitmix64.
"""
import numpy as np
import pytest

from oracle import (OracleDWF, chiral_apply, lpc_K, lpc_Kdag, chiral_gram, hash_uniform,
                    uniform_field, lpc_train)
from dwf_amd import lattice

DIMS, LS = (4, 2, 2, 4), 6


def make(b=1.5, c=0.5, mq=0.05):
    U = lattice.random_gauge(DIMS, 0.4, 71)
    o = OracleDWF(DIMS, LS, mq, 1.6, b, c)
    o.set_config(U)
    return o


@pytest.mark.parametrize("bc", [(1.0, 0.0), (1.5, 0.5)])
def test_identity_weights_give_neumann_series(bc):
    o = make(*bc)
    L = 3
    W = np.array([np.stack([np.eye(LS), np.eye(LS)])] * (L + 1))
    v = lattice.random_spinor5(DIMS, LS, 72)
    KDv = lpc_K(o, W, o.mult("D", v))
    h = v
    for _ in range(L + 1):
        h = h - o.mult("D", h)
    assert np.abs(KDv + h - v).max() < 1e-12 * np.abs(v).max()


def test_kdag_is_adjoint():
    o = make()
    rng = np.random.default_rng(73)
    W = rng.standard_normal((3, 2, LS, LS))
    u = lattice.random_spinor5(DIMS, LS, 74)
    v = lattice.random_spinor5(DIMS, LS, 75)
    a = np.sum(u * lpc_K(o, W, v))
    b = np.sum(lpc_Kdag(o, W, u) * v)
    assert abs(a - b) < 1e-11 * abs(a)


def test_chirality_convention_is_gamma5():
    x = lattice.random_spinor5(DIMS, LS, 76)
    W = np.stack([np.eye(LS), -np.eye(LS)])
    g5x = np.concatenate([x[..., 12:], x[..., :12]], axis=-1)
    assert np.abs(chiral_apply(W, x) - g5x).max() < 1e-15


def features(o, y, L):
    z = [y]
    for _ in range(L):
        z.append(z[-1] - o.mult("D", z[-1]))
    return z


def test_training_solves_the_normal_equations():
    o = make()
    L = 2
    vs = [lattice.random_spinor5(DIMS, LS, 77 + j) for j in range(2)]
    W, loss = lpc_train(o, L, vs, ridge=0.0)
    assert 0 < loss < 1
    # gradient of sum_v |K D v - v|^2 vanishes: sum_v G(R_v, z_k) = 0
    grad = np.zeros((L + 1, 2, LS, LS))
    scale = 0.0
    for v in vs:
        z = features(o, o.mult("D", v), L)
        R = lpc_K(o, W, z[0]) - v
        for k in range(L + 1):
            grad[k] += chiral_gram([R, z[k]])[1]
            scale = max(scale, np.abs(chiral_gram([v, z[k]])[1]).max())
    assert np.abs(grad).max() < 1e-9 * scale


def test_normal_equations_recover_representable_weights():
    # target K* y with known weights K*: the least-squares solution is K*
    o = make()
    L = 2
    Wstar = np.random.default_rng(80).standard_normal((L + 1, 2, LS, LS))
    y = o.mult("D", lattice.random_spinor5(DIMS, LS, 81))
    z = features(o, y, L)
    target = lpc_K(o, Wstar, y)
    F = L + 2
    G = chiral_gram([target] + z)
    pid = lambda i, j: i * F - i * (i - 1) // 2 + (j - i)
    d = (L + 1) * LS
    for c in (0, 1):
        A = np.zeros((d, d))
        B = np.zeros((d, LS))
        for a in range(L + 1):
            for b in range(L + 1):
                i, j = 1 + a, 1 + b
                A[a * LS:(a + 1) * LS, b * LS:(b + 1) * LS] = (
                    G[pid(i, j)][c] if i <= j else G[pid(j, i)][c].T)
            B[a * LS:(a + 1) * LS] = G[pid(0, 1 + a)][c].T
        X = np.linalg.solve(A, B)
        for k in range(L + 1):
            assert np.abs(X[k * LS:(k + 1) * LS].T - Wstar[k, c]).max() < 1e-7


def test_hash_matches_integer_splitmix():
    M = (1 << 64) - 1

    def ref(seed, i):
        z = (seed + (i + 1) * 0x9E3779B97F4A7C15) & M
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & M
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & M
        z ^= z >> 31
        return (z >> 11) * 2.0 ** -52 - 1.0

    seed = 0xDEADBEEFCAFEF00D
    idx = np.array([0, 1, 2, 12345, 2 ** 40 + 7])
    got = hash_uniform(seed, idx)
    assert all(got[n] == ref(seed, int(i)) for n, i in enumerate(idx))
    f = uniform_field(5, 3, 4)
    assert f.shape == (4, 3, 24) and f[2, 1, 7] == ref(5, (1 * 24 + 7) * 4 + 2)
    assert -1 <= f.min() and f.max() < 1
