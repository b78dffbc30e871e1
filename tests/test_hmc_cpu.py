"""HMC pieces on the CPU, pinned by the reference's printed energies.

The 8^4 sample trajectory of the reference (sample_1gpu/run.log, HMC_DWF_Nf2p1_eo.yaml,
conf_04040408.txt tiled to 8^4, RandomNumbers_parallel<Mseries> reset to 200000):
* H_kin = |iP|^2/2 of the Langevin momenta pins the M-series generator, its
  parallel (z, t) streams, Box-Muller and the momentum formula;
* the three pseudofermion |xi|^2 (Ls = 12, even sites) pin the order of the
  draws after the momenta;
* H_Grectangle pins the plaquette + rectangle action (oracle);
* the oracle's gauge force is pinned by finite differences of that action and
  the Taylor link update by scipy's matrix exponential.
The random numbers come from the product (host code of libdwfhip, no GPU).
"""
import json
import os

import numpy as np
import pytest
from scipy.linalg import expm

from gauge_oracle import exp_update, rect_action, rect_force, set_iP
from dwf_amd import lattice
from dwf_amd.hmc import RandomNumbers_parallel_Mseries, even_part

HERE = os.path.dirname(os.path.abspath(__file__))
KNOWN = json.load(open(os.path.join(HERE, "golden", "known_answers.json")))
DIMS = (8, 8, 8, 8)
BETA, CP, CR = 4.17, 1.66666666666666666666, -0.08333333333333333333


def conf8():
    return lattice.tile_4x4x4x8(np.load(os.path.join(HERE, "golden", "conf_04040408.npy")), DIMS)


def test_rng_momenta_and_noise_match_reference():
    r = RandomNumbers_parallel_Mseries(DIMS, 1234567)
    r.reset(200000)
    g = r.gauss_lex_global(8, 4)
    P, H_kin = set_iP(g)
    assert abs(H_kin - KNOWN["H_kin_init"]["value"]) < 1e-7
    for key in ("H_Frational_init", "H_Fratio_Nf2_init", "H_Fratio_Nf2prec_init"):
        xi = even_part(r.gauss_lex_global(24, 12), DIMS)
        h = float(np.sum(xi * xi))
        assert abs(h - KNOWN[key]["value"]) / KNOWN[key]["value"] < 1e-12, (key, h)


def test_rng_matches_reference_on_32x16x16x24():
    """The same input on 32x16x16x24 ran on 4 ranks (grid [1, 1, 2, 2]): the
    parallel generator's streams belong to global (z, t) slices, so this single
    process must draw the numbers of that run (submit_4node.sh.o1137366:1515-1545)."""
    dims = (32, 16, 16, 24)
    r = RandomNumbers_parallel_Mseries(dims, 1234567)
    r.reset(200000)
    P, H_kin = set_iP(r.gauss_lex_global(8, 4))
    del P
    assert abs(H_kin - KNOWN["H_kin_init_32x16x16x24"]["value"]) < 1e-6
    for key in ("H_Frational_init", "H_Fratio_Nf2_init", "H_Fratio_Nf2prec_init"):
        xi = r.gauss_eo_global(24, 12, 0)
        h = float(np.sum(xi * xi))
        want = KNOWN[key + "_32x16x16x24"]["value"]
        # H_F = |xi|^2 to the solver's precision (the reference prints |phi^+ M^-1 phi|)
        assert abs(h - want) / want < 1e-9, (key, h, want)


def test_rng_even_odd_draws_equal_lexical_then_even_part():
    """gauss_eo_global(nin, nex, p) = even_part-style selection of gauss_lex_global
    (Index_eo::convertField) from the same stream positions, both parities, and
    the streams end in the same state."""
    dims = (8, 4, 4, 6)
    a = RandomNumbers_parallel_Mseries(dims, 77)
    b = RandomNumbers_parallel_Mseries(dims, 77)
    for parity in (0, 1):
        lex = a.gauss_lex_global(24, 3)
        want = even_part(lex, dims) if parity == 0 else _odd_part(lex, dims)
        got = b.gauss_eo_global(24, 3, parity)
        assert np.array_equal(got, want)
    assert np.array_equal(a.gauss_lex_global(8, 1), b.gauss_lex_global(8, 1))


def _odd_part(xi_lex, dims):
    Nx, Ny, Nz, Nt = dims
    T, Z, Y, X2 = np.meshgrid(np.arange(Nt), np.arange(Nz), np.arange(Ny), np.arange(Nx // 2),
                              indexing="ij")
    x = 2 * X2 + (Y + Z + T + 1) % 2
    lex = (x + Nx * (Y + Ny * (Z + Nz * T))).reshape(-1)
    return xi_lex[:, lex, :]


def test_rng_reset_restarts_the_streams():
    r = RandomNumbers_parallel_Mseries((4, 4, 4, 4), 1)
    r.reset(7)
    a = r.gauss_lex_global(8, 2)
    b = r.gauss_lex_global(8, 2)
    r.reset(7)
    assert np.array_equal(r.gauss_lex_global(8, 2), a)
    assert not np.array_equal(a, b)
    # variance 1/2 per number, mean 0
    big = RandomNumbers_parallel_Mseries((8, 8, 8, 8), 3).gauss_lex_global(24, 4)
    assert abs(big.mean()) < 5e-3 and abs(big.var() - 0.5) < 5e-3


def test_rect_action_matches_reference():
    H, plaq = rect_action(conf8(), DIMS, BETA, CP, CR)
    assert abs(H - KNOWN["H_Grectangle_init"]["value"]) < 1e-7
    assert abs(plaq - KNOWN["plaq_init_8x8x8x8_from_conf_04040408"]["value"]) < 1e-14


def _su3_generator_field(rng, shape):
    a = rng.standard_normal(shape + (3, 3)) + 1j * rng.standard_normal(shape + (3, 3))
    x = 0.5 * (a - np.conj(np.swapaxes(a, -1, -2)))
    x -= (np.trace(x, axis1=-2, axis2=-1) / 3.0)[..., None, None] * np.eye(3)
    return x


def test_rect_force_is_minus_gradient_of_action():
    """The momentum update iP += eps F conserves H = |iP|^2/2 + S only if
    dS/deps of S(exp(eps X) U) equals -sum Re tr(X^+ F) for anti-hermitian
    traceless X; checked by central differences."""
    dims = (4, 4, 4, 4)
    U = lattice.random_gauge(dims, 0.4, 5)
    F = rect_force(U, dims, BETA, CP, CR)
    rng = np.random.default_rng(6)
    X = _su3_generator_field(rng, (4, int(np.prod(dims))))
    Fc = F.reshape(4, -1, 3, 3, 2).view(np.complex128)[..., 0]
    want = -np.sum((np.conj(X) * Fc).real)

    def S(eps):
        Uc = U.reshape(4, -1, 3, 3, 2).view(np.complex128)[..., 0]
        E = np.stack([np.stack([expm(eps * X[m, s]) for s in range(X.shape[1])]) for m in range(4)])
        Un = np.ascontiguousarray(E @ Uc).view(np.float64).reshape(4, -1, 18)
        return rect_action(Un, dims, BETA, CP, CR)[0]

    h = 1e-4
    got = (S(h) - S(-h)) / (2 * h)
    assert abs(got - want) / abs(want) < 1e-6, (got, want)


def test_exp_update_is_taylor_of_expm_and_unitary():
    rng = np.random.default_rng(8)
    dims = (2, 2, 2, 2)
    U = lattice.random_gauge(dims, 0.7, 9)
    X = _su3_generator_field(rng, (4, 16)) * 0.3
    iP = np.ascontiguousarray(X).view(np.float64).reshape(4, 16, 18)
    W = exp_update(U, 0.1, iP, 8)
    Wc = W.reshape(4, -1, 3, 3, 2).view(np.complex128)[..., 0]
    Uc = U.reshape(4, -1, 3, 3, 2).view(np.complex128)[..., 0]
    ref = np.stack([np.stack([expm(0.1 * X[m, s]) @ Uc[m, s] for s in range(16)]) for m in range(4)])
    assert np.abs(Wc - ref).max() < 1e-13
    eye = np.eye(3)
    assert np.abs(Wc @ np.conj(np.swapaxes(Wc, -1, -2)) - eye).max() < 1e-14
    assert np.abs(np.linalg.det(Wc) - 1).max() < 1e-14


def test_binary_gauge_file_round_trip(tmp_path):
    """GaugeConfig("Binary"): the file holds, per lexical site, the 4 x 18 link
    reals big-endian; write then read is the identity, and the reference's
    plaquette of conf_04040408 tiled to 8^4 survives the round trip."""
    from oracle import plaquette
    U = conf8()
    path = str(tmp_path / "conf.bin")
    lattice.write_binary(path, DIMS, U)
    raw = np.fromfile(path, dtype=">f8")
    assert raw.size == U.size
    assert np.array_equal(raw.reshape(-1, 4, 18).transpose(1, 0, 2), U)
    back = lattice.read_binary(path, DIMS)
    assert np.array_equal(back, U)
    assert abs(plaquette(DIMS, back) - KNOWN["plaq_init_8x8x8x8_from_conf_04040408"]["value"]) < 1e-14
    with pytest.raises(IOError):
        lattice.read_binary(path, (8, 8, 8, 16))   # short file
