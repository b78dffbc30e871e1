"""The multishift CG's Krylov-block fold (the shifts' x and p updated once per
block of K iterations from the unshifted residuals, |p_j|^2 carried by the
recurrence |p_j|^2 <- alpha_j^2 |p_j|^2 + zeta_j^2 |r|^2 between folds) against
the reference iteration (ashiftsolver_CG-tmpl.h:216-286: every shift updated and
its |p_j|^2 computed every iteration), on a dense SPD matrix with condition
number 1e4 and 12 shifts: the same iterations, the same truncation iteration of
every shift, and the same solutions to round-off.  A planted mistake (the
recurrence without the alpha_j^2 factor) changes the truncation.

With A = M^T M (D^dag D), the unshifted system's <p0, (A + sigma0) p0> taken as
|M p0|^2 + sigma0 |p0|^2 (the first norm fused into the operator, the second
into the p0 update) and sigma0 p0 added in the residual update is the same
iteration up to rounding: on a problem that converges before CG loses
orthogonality, the iteration count within 2, solutions equal to the solver's
tolerance and every shift's true residual below it (a run past the loss of
orthogonality, 709 iterations at n = 400, took 719: rounding, not a different
method).  Dropping the sigma0 |p0|^2 term does not converge."""
import numpy as np
import pytest


def _problem(n=400, seed=3, with_m=False):
    rng = np.random.default_rng(seed)
    q, _ = np.linalg.qr(rng.standard_normal((n, n)))
    lam = np.logspace(-3, 1, n)
    a = (q * lam) @ q.T
    m = (q * np.sqrt(lam)).T   # a = m^T m
    out = (a, rng.standard_normal(n), np.logspace(-3, 1, 12))
    return out + (m,) if with_m else out


def reference(a, b, sig, stop):
    """AShiftsolver_CG::solve as the reference runs it."""
    n = len(sig)
    x = [np.zeros_like(b) for _ in sig]
    p = [b.copy() for _ in sig]
    r = b.copy()
    snorm = 1 / (b @ b)
    rr = r @ r
    z1, z2, csh = np.ones(n), np.ones(n), sig - sig[0]
    ap, bp, n2, pp, trunc = 0.0, 1.0, n, np.zeros(n), {}
    for it in range(10000):
        s = a @ p[0] + sig[0] * p[0]
        rrp, pa = rr, s @ p[0]
        beta = -rrp / pa
        x[0] -= beta * p[0]
        r = r + beta * s
        rr = r @ r
        alpha = rr / rrp
        p[0] = r + alpha * p[0]
        pp[0] = rr
        ah = 1 + ap * beta / bp
        for j in range(1, n2):
            z = 1 / ((ah - csh[j] * beta) / z1[j] + (1 - ah) / z2[j])
            zr = z / z1[j]
            x[j] -= beta * zr * p[j]
            p[j] = alpha * zr * zr * p[j] + z * r
            pp[j] = (p[j] @ p[j]) * snorm
            z2[j], z1[j] = z1[j], z
        ish1 = n2
        for j in range(n2 - 1, -1, -1):
            if pp[j] > stop:
                ish1 = j + 1
                break
        for j in range(ish1, n2):
            trunc[j] = it
        n2, ap, bp = ish1, alpha, beta
        if rr * snorm < stop:
            return it, trunc, x


def folded(a, b, sig, stop, K=16, planted=False, m_op=None, planted_fused=False):
    """The same iteration with the shifts folded once per block of K iterations;
    with m_op (a = m_op^T m_op): pa = |m_op p0|^2 + sigma0 |p0|^2 and sigma0 p0
    added in the residual update."""
    n = len(sig)
    X = [np.zeros_like(b) for _ in sig]
    P = [b.copy() for _ in sig]
    snorm = 1 / (b @ b)
    z1, z2, csh = np.ones(n), np.ones(n), sig - sig[0]
    ap, bp, n2, pp = 0.0, 1.0, n, np.zeros(n)
    ppe = np.full(n, b @ b)
    p0, x0, R = b.copy(), np.zeros_like(b), [b.copy()]
    ca, cb = np.ones(n), np.zeros(n)
    cc, cd = np.zeros((n, K)), np.zeros((n, K))
    rr, m, nfold, trunc = b @ b, 0, n, {}
    pp0 = b @ b   # |p0|^2
    for it in range(10000):
        rrp = rr
        if m_op is None:
            s = a @ p0 + sig[0] * p0
            pa = s @ p0
        else:
            u = m_op @ p0
            s = m_op.T @ u + sig[0] * p0
            pa = u @ u + (0.0 if planted_fused else sig[0]) * pp0
        beta = -rrp / pa
        x0 -= beta * p0
        rn = R[m] + beta * s
        rr = rn @ rn
        alpha = rr / rrp
        p0 = rn + alpha * p0
        pp0 = p0 @ p0
        pp[0] = rr
        m += 1
        R.append(rn)
        ah = 1 + ap * beta / bp
        for j in range(1, n2):
            z = 1 / ((ah - csh[j] * beta) / z1[j] + (1 - ah) / z2[j])
            zr = z / z1[j]
            bs, as_ = beta * zr, alpha * zr * zr
            cb[j] -= bs * ca[j]
            cd[j, :m - 1] -= bs * cc[j, :m - 1]
            ca[j] *= as_
            cc[j, :m - 1] *= as_
            cc[j, m - 1] = z
            ppe[j] = (1.0 if planted else as_ * as_) * ppe[j] + z * z * rr
            pp[j] = ppe[j] * snorm
            z2[j], z1[j] = z1[j], z
        ish1 = n2
        for j in range(n2 - 1, -1, -1):
            if pp[j] > stop:
                ish1 = j + 1
                break
        for j in range(ish1, n2):
            trunc[j] = it
        n2, ap, bp = ish1, alpha, beta
        done = rr * snorm < stop
        if m == K or done:
            for j in range(1, nfold):
                X[j] = X[j] + cb[j] * P[j] + sum(cd[j, l] * R[l + 1] for l in range(m))
                P[j] = ca[j] * P[j] + sum(cc[j, l] * R[l + 1] for l in range(m))
                ppe[j] = P[j] @ P[j]
            R = [R[m]]
            ca[:], cb[:], cc[:], cd[:] = 1, 0, 0, 0
            m, nfold = 0, n2
        if done:
            X[0] = x0
            return it, trunc, X


@pytest.mark.parametrize("K", [4, 16])
def test_fold_reproduces_the_reference_iteration(K):
    a, b, sig = _problem()
    it_r, tr_r, x_r = reference(a, b, sig, 1e-22)
    it_f, tr_f, x_f = folded(a, b, sig, 1e-22, K)
    assert it_f == it_r
    assert tr_f == tr_r and len(tr_r) == len(sig) - 1
    for xr, xf in zip(x_r, x_f):
        assert np.linalg.norm(xf - xr) / np.linalg.norm(xr) < 1e-12


def test_planted_recurrence_mistake_changes_truncation():
    a, b, sig = _problem()
    _, tr_r, _ = reference(a, b, sig, 1e-22)
    _, tr_p, _ = folded(a, b, sig, 1e-22, 16, planted=True)
    assert tr_p != tr_r


def _problem_mtm(n=600, seed=3):
    """A = M^T M with eigenvalues on [1e-2, 1] (CG converges in ~120 < n iterations)."""
    rng = np.random.default_rng(seed)
    q, _ = np.linalg.qr(rng.standard_normal((n, n)))
    lam = np.logspace(-2, 0, n)
    return (q * lam) @ q.T, rng.standard_normal(n), np.logspace(-3, 1, 12), (q * np.sqrt(lam)).T


def test_fused_unshifted_reductions_match_the_reference_iteration():
    a, b, sig, m = _problem_mtm()
    it_r, _, x_r = reference(a, b, sig, 1e-22)
    it_f, _, x_f = folded(a, b, sig, 1e-22, 16, m_op=m)
    assert abs(it_f - it_r) <= 2
    for s_, xr, xf in zip(sig, x_r, x_f):
        assert np.linalg.norm(xf - xr) / np.linalg.norm(xr) < 1e-10
        assert np.linalg.norm(a @ xf + s_ * xf - b) / np.linalg.norm(b) < 1e-10


def test_planted_missing_shift_term_in_pa_does_not_converge():
    a, b, sig, m = _problem_mtm()
    assert folded(a, b, sig, 1e-22, 16, m_op=m, planted_fused=True) is None
