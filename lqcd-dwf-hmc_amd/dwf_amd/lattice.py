"""Host-side lattice data: synthetic inputs, on-disk formats and layout helpers.

All arrays use the reference's corelib interchange layouts, which are the
layouts the C-ABI of libdwfhip accepts at its boundary:

* gauge field ``Field_G(18, V, 4)`` (src/lib/Field/field.h:64-66): a float64
  array of shape ``(4, V, 18)``; element ``ab = ri + 2*(b + 3*a)`` of the
  row-major 3x3 matrix ``U_mu(x)[a][b]`` (src/lib/Tools/mat_SU_N.h).
  ``as_su3(U)`` gives the complex view ``(4, V, 3, 3)``.
* 5D even-odd spinor ``Field(24, V/2, Ls)``: float64 array ``(Ls, V/2, 24)``;
  ``in = ri + 2*(c + 3*d)`` (colour c, spin d).  ``as_spinor`` gives the
  complex view ``(Ls, V/2, 4, 3)``.
* lexical site ``x + Nx*(y + Ny*(z + Nz*t))`` (src/lib/Field/index_lex.h:58);
  even-odd half site ``x//2 + Nx/2*(y + Ny*(z + Nz*t))`` with parity
  ``(x+y+z+t) % 2`` (src/lib/Field/index_eo.h:66, index_eo.cpp).
"""
from __future__ import annotations

import math
import os
import re

import numpy as np


def volume(dims):
    return int(dims[0]) * int(dims[1]) * int(dims[2]) * int(dims[3])


def as_su3(U):
    """Complex (4, V, 3, 3) view of a Field_G (4, V, 18) float64 array."""
    return U.reshape(U.shape[0], U.shape[1], 3, 3, 2).view(np.complex128)[..., 0]


def from_su3(M):
    """Field_G float64 (4, V, 18) array from a complex (4, V, 3, 3) array."""
    M = np.ascontiguousarray(M, dtype=np.complex128)
    return M.view(np.float64).reshape(M.shape[0], M.shape[1], 18).copy()


def as_spinor(psi):
    """Complex (Ls, Vh, 4, 3) view of a Field(24, Vh, Ls) float64 array."""
    return psi.reshape(psi.shape[0], psi.shape[1], 4, 3, 2).view(np.complex128)[..., 0]


def from_spinor(P):
    P = np.ascontiguousarray(P, dtype=np.complex128)
    return P.view(np.float64).reshape(P.shape[0], P.shape[1], 24).copy()


def random_hermitian_traceless(rng, n):
    """n Gaussian traceless Hermitian 3x3 matrices (entries of unit variance)."""
    a = rng.standard_normal((n, 3, 3)) + 1j * rng.standard_normal((n, 3, 3))
    h = 0.5 * (a + np.conj(np.swapaxes(a, -1, -2)))
    tr = np.trace(h, axis1=-2, axis2=-1).real / 3.0
    h -= tr[:, None, None] * np.eye(3)[None]
    return h


def expi_hermitian(h, eps):
    """exp(i eps h) for a batch of Hermitian matrices, by eigendecomposition."""
    w, v = np.linalg.eigh(h)
    ph = np.exp(1j * eps * w)
    return np.einsum("nij,nj,nkj->nik", v, ph, np.conj(v))


def random_gauge(dims, eps, seed):
    """Synthetic links U_mu(x) = exp(i eps H), H Gaussian traceless Hermitian.

    Returns a Field_G float64 (4, V, 18) array.  Deterministic in ``seed``.
    """
    V = volume(dims)
    rng = np.random.default_rng(seed)
    out = np.empty((4, V, 3, 3), np.complex128)
    chunk = 1 << 16
    for mu in range(4):
        for s0 in range(0, V, chunk):
            n = min(chunk, V - s0)
            out[mu, s0:s0 + n] = expi_hermitian(random_hermitian_traceless(rng, n), eps)
    return from_su3(out)


def unit_gauge(dims):
    V = volume(dims)
    M = np.broadcast_to(np.eye(3, dtype=np.complex128), (4, V, 3, 3))
    return from_su3(M)


def random_spinor5(dims, Ls, seed):
    """Gaussian complex 5D even-odd spinor, Field(24, V/2, Ls) float64."""
    Vh = volume(dims) // 2
    rng = np.random.default_rng(seed)
    return rng.standard_normal((Ls, Vh, 24))


def read_text_4x4x4x8(path, dims):
    """Gauge field in the reference's Text_4x4x4x8 format, tiled onto dims.

    Follows src/lib/IO/fieldIO_Text_4x4x4x8.cpp:26-160: the file lists, for
    each site of a 4x4x4x8 lattice in lexical order, the 4 directions times 18
    reals; the lattice is filled periodically from it.
    """
    vals = np.loadtxt(path, dtype=np.float64)
    F = (4, 4, 4, 8)
    VF = volume(F)
    if vals.size != VF * 4 * 18:
        raise ValueError(f"{path}: expected {VF*4*18} values, found {vals.size}")
    return tile_4x4x4x8(vals.reshape(VF, 4, 18), dims)


def tile_4x4x4x8(file_order, dims):
    """Periodic tiling of a (4*4*4*8, 4, 18) site-major array onto dims."""
    F = (4, 4, 4, 8)
    Nx, Ny, Nz, Nt = dims
    x = np.arange(Nx) % F[0]
    y = np.arange(Ny) % F[1]
    z = np.arange(Nz) % F[2]
    t = np.arange(Nt) % F[3]
    g = (x[None, None, None, :] + F[0] * (y[None, None, :, None] + F[1] * (
        z[None, :, None, None] + F[2] * t[:, None, None, None]))).reshape(-1)
    return np.ascontiguousarray(np.transpose(file_order[g], (1, 0, 2)))


def read_binary(path, dims):
    """GaugeConfig("Binary"): ILDG order, big-endian (libdwfhip's reader).
    Returns a Field_G float64 (4, V, 18) array."""
    import ctypes as C
    from . import _lib
    U = np.empty((4, volume(dims), 18))
    d = (C.c_int * 4)(*[int(v) for v in dims])
    rc = _lib.load().dwf_gauge_read_binary(os.fsencode(path), d,
                                            U.ctypes.data_as(C.POINTER(C.c_double)))
    if rc != 0:
        raise IOError(f"{path}: not a Binary gauge file of lattice {tuple(dims)}")
    return U


def write_binary(path, dims, U):
    """GaugeConfig("Binary").write_file: ILDG order, big-endian."""
    import ctypes as C
    from . import _lib
    U = np.ascontiguousarray(U, np.float64)
    assert U.shape == (4, volume(dims), 18)
    d = (C.c_int * 4)(*[int(v) for v in dims])
    rc = _lib.load().dwf_gauge_write_binary(os.fsencode(path), d,
                                             U.ctypes.data_as(C.POINTER(C.c_double)))
    if rc != 0:
        raise IOError(f"{path}: write failed")


# ---------------------------------------------------------------------------
# rational approximation files (src/lib/Tools/math_Rational.cpp, read_file)
# ---------------------------------------------------------------------------

def rational_filename(n_exp, d_exp, npoles, x_min, x_max):
    """parameter_rational.{|n|}_{d}{inv}_{Np:02d}_{xmin:010.8f}_{xmax:07.3f}."""
    inv = "inv" if n_exp < 0 else ""
    return "parameter_rational.%d_%d%s_%02d_%010.8f_%07.3f" % (
        abs(n_exp), d_exp, inv, npoles, x_min, x_max)


def read_rational(path):
    """Return dict(np, n, d, x_min, x_max, error, a0, res[Np], pole[Np])."""
    with open(path) as f:
        toks = f.read().split()
    np_, n, d = int(toks[0]), int(toks[1]), int(toks[2])
    x_min, x_max = float(toks[3]), float(toks[4])
    err = float(toks[5])
    a0 = float(toks[6])
    body = [float(v) for v in toks[7:7 + 2 * np_]]
    res = np.array(body[0::2])
    pole = np.array(body[1::2])
    return dict(np=np_, n=n, d=d, x_min=x_min, x_max=x_max, error=err,
                a0=a0, res=res, pole=pole)


def log_spaced_shifts(n, lo=1e-4, hi=10.0):
    """n shifts log-spaced on [lo, hi] (ascending; sigma_0 is the smallest)."""
    return np.exp(np.linspace(math.log(lo), math.log(hi), n))
