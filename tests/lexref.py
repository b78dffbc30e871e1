"""Independent dense formulation of the Moebius domain-wall operator for tests.

A second, differently written statement of the operator used only to pin the
oracle: full (not even-odd) lexical lattice, spin handled with explicit 4x4
gamma matrices, shifts with ``np.roll``.  Conventions:

* gamma matrices: the reference's Dirac representation
  (src/lib/Tools/gammaMatrixSet_Dirac.cpp:37-65), entered here as data;
* Wilson hop: sum_mu (1 - g_mu) U_mu(x) psi(x+mu) + (1 + g_mu) U_mu(x-mu)^+ psi(x-mu)
  (forward term with (1 - gamma_mu), src/lib/Fopr/Imp/fopr_Wilson_eo_impl.cpp:714-890);
* D_W = (4 - M0) - hop/2;
* Moebius: (D psi)_s = (b D_W + 1) psi_s + (c D_W - 1) (P- psi_{s+1} + P+ psi_{s-1}),
  with -mq at the walls (src/lib/Fopr/afopr_Domainwall_eo-tmpl.h:765-959).
"""
import numpy as np

I4 = np.eye(4, dtype=np.complex128)
G = np.zeros((5, 4, 4), np.complex128)   # G[0..3] = gamma_1..gamma_4, G[4] = gamma_5
# GAMMA1
G[0, 0, 3] = -1j; G[0, 1, 2] = -1j; G[0, 2, 1] = 1j; G[0, 3, 0] = 1j
# GAMMA2
G[1, 0, 3] = -1; G[1, 1, 2] = 1; G[1, 2, 1] = 1; G[1, 3, 0] = -1
# GAMMA3
G[2, 0, 2] = -1j; G[2, 1, 3] = 1j; G[2, 2, 0] = 1j; G[2, 3, 1] = -1j
# GAMMA4
G[3, 0, 0] = 1; G[3, 1, 1] = 1; G[3, 2, 2] = -1; G[3, 3, 3] = -1
# GAMMA5
G[4, 0, 2] = 1; G[4, 1, 3] = 1; G[4, 2, 0] = 1; G[4, 3, 1] = 1

PP = 0.5 * (I4 + G[4])
PM = 0.5 * (I4 - G[4])


def lex_gauge(U, dims):
    """(4, V, 18) Field_G -> complex (4, Nt, Nz, Ny, Nx, 3, 3)."""
    Nx, Ny, Nz, Nt = dims
    M = U.reshape(4, -1, 3, 3, 2).view(np.complex128)[..., 0]
    return M.reshape(4, Nt, Nz, Ny, Nx, 3, 3)


# array axis of direction mu in a (..., Nt, Nz, Ny, Nx, 4, 3) field
def _ax(mu, nd):
    return nd - 3 - mu


def _bc_mask(dims, mu, forward, bc, shape_lead):
    """Boundary factor field: bc at the slice whose neighbour wraps."""
    Nx, Ny, Nz, Nt = dims
    N = dims[mu]
    f = np.ones(N)
    f[N - 1 if forward else 0] = bc
    sh = [1] * 4
    sh[3 - mu] = N
    return f.reshape(sh)


def hop(psi, Ul, dims, bc):
    """Wilson hop on a (Nt, Nz, Ny, Nx, 4, 3) spinor (one 4D slice)."""
    out = np.zeros_like(psi)
    for mu in range(4):
        ax = 3 - mu
        U = Ul[mu]
        fwd = np.roll(psi, -1, axis=ax) * _bc_mask(dims, mu, True, bc[mu], None)[..., None, None]
        t = np.einsum("...ab,...db->...da", U, fwd)           # U_mu(x) psi(x+mu)
        out += t - np.einsum("de,...ea->...da", G[mu], t)      # (1 - g_mu)
        Ub = np.roll(U, 1, axis=ax)                            # U_mu(x-mu)
        bwd = np.roll(psi, 1, axis=ax) * _bc_mask(dims, mu, False, bc[mu], None)[..., None, None]
        t = np.einsum("...ba,...db->...da", np.conj(Ub), bwd)  # U^+ psi(x-mu)
        out += t + np.einsum("de,...ea->...da", G[mu], t)      # (1 + g_mu)
    return out


def spin(M, psi):
    return np.einsum("de,...ea->...da", M, psi)


def dwf_full(psi5, Ul, dims, Ls, mq, M0, b, c, bc):
    """Full 5D Moebius operator on (Ls, Nt, Nz, Ny, Nx, 4, 3)."""
    b = np.broadcast_to(np.asarray(b, float), (Ls,))
    c = np.broadcast_to(np.asarray(c, float), (Ls,))
    def DW(f):
        return (4.0 - M0) * f - 0.5 * hop(f, Ul, dims, bc)

    out = np.empty_like(psi5)
    for s in range(Ls):
        up = spin(PM, psi5[(s + 1) % Ls]) * (-mq if s == Ls - 1 else 1.0)
        dn = spin(PP, psi5[(s - 1) % Ls]) * (-mq if s == 0 else 1.0)
        mv = up + dn
        out[s] = b[s] * DW(psi5[s]) + psi5[s] + c[s] * DW(mv) - mv
    return out


def eo_maps(dims):
    """For parity p: lexical (t,z,y,x) coordinates of each half site."""
    Nx, Ny, Nz, Nt = dims
    Vh = Nx * Ny * Nz * Nt // 2
    maps = []
    s2 = np.arange(Vh)
    x2 = s2 % (Nx // 2)
    r = s2 // (Nx // 2)
    y = r % Ny
    r //= Ny
    z = r % Nz
    t = r // Nz
    for p in range(2):
        x = 2 * x2 + (y + z + t + p) % 2
        maps.append((t, z, y, x))
    return maps


def to_eo(psi_lex, dims, p):
    """(Ls, Nt, Nz, Ny, Nx, 4, 3) -> Field(24, Vh, Ls) float64 of parity p."""
    t, z, y, x = eo_maps(dims)[p]
    sub = psi_lex[:, t, z, y, x]              # (Ls, Vh, 4, 3)
    return np.ascontiguousarray(sub).view(np.float64).reshape(sub.shape[0], sub.shape[1], 24)


def from_eo(fe, fo, dims):
    """Two parity fields (Ls, Vh, 24) -> lexical (Ls, Nt, Nz, Ny, Nx, 4, 3)."""
    Nx, Ny, Nz, Nt = dims
    Ls = fe.shape[0]
    out = np.zeros((Ls, Nt, Nz, Ny, Nx, 4, 3), np.complex128)
    for p, f in enumerate((fe, fo)):
        t, z, y, x = eo_maps(dims)[p]
        out[:, t, z, y, x] = f.reshape(Ls, -1, 4, 3, 2).view(np.complex128)[..., 0]
    return out


def momentum_matrix(p, Ls, mq, M0, b, c):
    """Free-field Moebius operator at momentum p as a (4Ls x 4Ls) matrix.

    D_W(p) = (4 - M0 - sum cos p) + i sum gamma_mu sin p_mu.
    """
    b = np.broadcast_to(np.asarray(b, float), (Ls,))
    c = np.broadcast_to(np.asarray(c, float), (Ls,))
    DW = (4.0 - M0 - sum(np.cos(p))) * I4 + 1j * sum(G[m] * np.sin(p[m]) for m in range(4))
    D = np.zeros((Ls, 4, Ls, 4), np.complex128)
    for s in range(Ls):
        D[s, :, s, :] += b[s] * DW + I4
        su, sd = (s + 1) % Ls, (s - 1) % Ls
        fu = -mq if s == Ls - 1 else 1.0
        fd = -mq if s == 0 else 1.0
        D[s, :, su, :] += fu * (c[s] * DW - I4) @ PM
        D[s, :, sd, :] += fd * (c[s] * DW - I4) @ PP
    return D.reshape(4 * Ls, 4 * Ls)
