"""CPU restatement (numpy, FP64) of the HMC gauge part.

TEST INFRASTRUCTURE ONLY: imported by tests/ as the checker of the GPU
kernels, never by the product path.

* rect_action / rect_force: Action_G_Rectangle::calcH and
  Force_G_Rectangle::force_core (src/lib/Action/Gauge/action_G_Rectangle.cpp:113-193,
  src/lib/Force/Gauge/force_G_Rectangle.cpp:84-216), written with whole-lattice
  shifts the way the reference writes them (ShiftField_lex::backward(v, w, mu):
  v(x) = w(x + mu); forward: v(x) = w(x - mu); Staple_lex::upper / lower,
  src/lib/Measurements/Gauge/staple_lex.cpp:260-301).
* exp_update: SU_N::mat_exp + reunit_SU3 (src/lib/Tools/mat_SU_N.h:637-655,
  mat_SU_N.cpp:18-86).
* set_iP: Langevin_Momentum::set_iP_SU3 (src/lib/HMC/langevin_Momentum.cpp:146-200).

Gauge fields are corelib Field_G float64 arrays (4, V, 18), lexical sites.
"""
from __future__ import annotations

import numpy as np


def _mats(U, dims):
    """(4, V, 18) -> complex (4, Nt, Nz, Ny, Nx, 3, 3)."""
    Nx, Ny, Nz, Nt = dims
    M = U.reshape(4, -1, 3, 3, 2).view(np.complex128)[..., 0]
    return M.reshape(4, Nt, Nz, Ny, Nx, 3, 3)


def _field(M):
    """complex (4, Nt, Nz, Ny, Nx, 3, 3) -> (4, V, 18)."""
    M = np.ascontiguousarray(M.reshape(4, -1, 3, 3), np.complex128)
    return M.view(np.float64).reshape(4, -1, 18).copy()


_AXIS = {0: 3, 1: 2, 2: 1, 3: 0}   # direction mu -> axis of a (Nt, Nz, Ny, Nx, 3, 3) array


def backward(A, mu):
    """v(x) = A(x + mu)"""
    return np.roll(A, -1, axis=_AXIS[mu])


def forward(A, mu):
    """v(x) = A(x - mu)"""
    return np.roll(A, 1, axis=_AXIS[mu])


def mnn(a, b):
    return a @ b


def mnd(a, b):
    return a @ np.conj(np.swapaxes(b, -1, -2))


def mdn(a, b):
    return np.conj(np.swapaxes(a, -1, -2)) @ b


def staple_upper(M, mu, nu):
    """U_nu(x) U_mu(x+nu) U_nu(x+mu)^+"""
    return mnn(M[nu], mnd(backward(M[mu], nu), backward(M[nu], mu)))


def staple_lower(M, mu, nu):
    """U_nu(x-nu)^+ U_mu(x-nu) U_nu(x-nu+mu)"""
    return forward(mdn(M[nu], mnn(M[mu], backward(M[nu], mu))), nu)


def _retr_nd(a, b):
    return np.sum((a * np.conj(b)).real)


def rect_action(U, dims, beta, c_plaq, c_rect):
    """(H, average plaquette) of the plaquette + rectangle gauge action."""
    M = _mats(U, dims)
    V = int(np.prod(dims))
    plaqF = rectF = 0.0
    for mu in range(4):
        for nu in range(mu + 1, 4):
            cup1 = staple_upper(M, mu, nu)
            plaqF += _retr_nd(M[mu], cup1)
            cup2 = staple_upper(M, nu, mu)
            v = backward(cup2, mu)
            c = backward(M[mu], nu)
            c = mnn(M[nu], mnd(c, v))
            rectF += _retr_nd(M[mu], c)
            v = backward(M[nu], mu)
            c = backward(cup1, nu)
            c = mnn(M[nu], mnd(c, v))
            rectF += _retr_nd(M[mu], c)
    H = beta * (c_plaq * 6 * V - c_plaq * plaqF / 3 + c_rect * 6 * V * 2 - c_rect * rectF / 3)
    return H, plaqF / 3 / V / 6


def at(m):
    """traceless anti-hermitian part (m - m^+)/2 - tr(m - m^+)/6"""
    a = 0.5 * (m - np.conj(np.swapaxes(m, -1, -2)))
    tr = np.trace(a, axis1=-2, axis2=-1) / 3.0
    return a - tr[..., None, None] * np.eye(3)


def rect_force(U, dims, beta, c_plaq, c_rect):
    """The gauge force F (4, V, 18): F_mu = -(beta/3) at(U_mu S_mu^+)."""
    M = _mats(U, dims)
    F = np.empty_like(M)
    for mu in range(4):
        S = np.zeros_like(M[0])
        for nu in range(4):
            if nu == mu:
                continue
            cup1 = staple_upper(M, mu, nu)
            cup2 = staple_upper(M, nu, mu)
            cdn1 = staple_lower(M, mu, nu)
            cdn2 = staple_lower(M, nu, mu)
            S += c_plaq * (cup1 + cdn1)
            Umu, Unu = M[mu], M[nu]
            # 2x1 up-right, 1x2 up, 2x1 up-left
            S += c_rect * mnn(Unu, mnd(backward(Umu, nu), backward(cup2, mu)))
            S += c_rect * mnn(Unu, mnd(backward(cup1, nu), backward(Unu, mu)))
            S += c_rect * mnn(cdn2, mnd(backward(Umu, nu), backward(Unu, mu)))
            # 2x1 down-right, 1x2 down, 2x1 down-left
            S += c_rect * forward(mdn(Unu, mnn(Umu, backward(cup2, mu))), nu)
            S += c_rect * forward(mdn(Unu, mnn(cdn1, backward(Unu, mu))), nu)
            S += c_rect * forward(mdn(cdn2, mnn(Umu, backward(Unu, mu))), nu)
        F[mu] = -(beta / 3.0) * at(mnd(M[mu], S))
    return _field(F)


def reunit(m):
    """Mat_SU_N::reunit_SU3 on (..., 3, 3): normalise row 0, orthonormalise
    row 1 against it, row 2 = (row0 x row1)^*."""
    r0 = m[..., 0, :] / np.linalg.norm(m[..., 0, :], axis=-1)[..., None]
    sp = np.sum(np.conj(r0) * m[..., 1, :], axis=-1)
    r1 = m[..., 1, :] - sp[..., None] * r0
    r1 = r1 / np.linalg.norm(r1, axis=-1)[..., None]
    r2 = np.conj(np.cross(r0, r1))
    return np.stack([r0, r1, r2], axis=-2)


def exp_update(U, alpha, iP, nprec):
    """exp(alpha iP) U by the nested Taylor form of order nprec, reunitarised."""
    u0 = U.reshape(4, -1, 3, 3, 2).view(np.complex128)[..., 0]
    iv = iP.reshape(4, -1, 3, 3, 2).view(np.complex128)[..., 0]
    tmp = u0.copy()
    for k in range(nprec):
        tmp = (iv @ tmp) * (alpha / (nprec - k)) + u0
    out = reunit(tmp)
    return np.ascontiguousarray(out).view(np.float64).reshape(4, -1, 18).copy()


def set_iP(gauss):
    """iP (4, V, 18) and H_kin = |iP|^2/2 from gauss (4, V, 8)."""
    sq3r = 1.0 / np.sqrt(3.0)
    h = [gauss[..., i] for i in range(8)]
    z = np.zeros_like(h[0])
    comps = [z, h[2] + h[7] * sq3r, h[1], h[0], h[4], h[3],
             -h[1], h[0], z, -h[2] + h[7] * sq3r, h[6], h[5],
             -h[4], h[3], -h[6], h[5], z, -h[7] * 2.0 * sq3r]
    P = np.stack(comps, axis=-1)
    return P, 0.5 * float(np.sum(P * P))
