"""GPU DWF fermion force against the oracle (FP64, 1e-13 relative)."""
import numpy as np
import pytest

from oracle import OracleDWF
import dwf_amd
from dwf_amd import lattice

pytestmark = pytest.mark.gpu

PARAMS = [
    ((4, 4, 2, 4), 4, 0.02, 1.6, 1.0, 0.0, (1, 1, 1, -1)),
    ((4, 2, 4, 4), 6, 0.05, 1.0, 1.5, 0.5, (1, 1, 1, -1)),
    ((4, 4, 2, 2), 5, 0.3, 1.3, 1.2, 0.3, (-1, 1, -1, -1)),
    ((8, 8, 8, 8), 16, 0.01, 1.6, 1.0, 0.0, (1, 1, 1, -1)),
]


def rel(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


@pytest.mark.parametrize("prm", PARAMS)
def test_force_modes_match_oracle(prm):
    dims, Ls, mq, M0, b, c, bc = prm
    print("case", prm, flush=True)
    U = lattice.random_gauge(dims, 0.5, 8)
    orc = OracleDWF(dims, Ls, mq, M0, b, c, bc=bc)
    orc.set_config(U)
    gpu = dwf_amd.AFopr_Domainwall_5din_eo(dims, Ls, mq, M0, b, c, bc)
    gpu.set_config(U)
    force = dwf_amd.AForce_F_Domainwall_5din_eo(gpu)
    zeta = lattice.random_spinor5(dims, Ls, 9)
    eta = lattice.random_spinor5(dims, Ls, 10)
    fz, fe = gpu.new_field(), gpu.new_field()
    gpu.convert(fz, zeta)
    gpu.convert(fe, eta)
    F = force.new_force()
    for mode in ("D", "Ddag", "H", "Hdag"):
        force.set_mode(mode)
        force.force_udiv1(F, fz, fe)
        assert rel(F.to_host(), orc.force_udiv1(mode, zeta, eta)) < 1e-13, mode
        force.force_core1(F, fz, fe)
        assert rel(F.to_host(), orc.force_core1(mode, zeta, eta)) < 1e-13, mode
    force.force_udiv(F, fe)
    assert rel(F.to_host(), orc.force_udiv(eta)) < 1e-13
    force.force_core(F, fe)
    assert rel(F.to_host(), orc.force_core(eta)) < 1e-13


RAT_NUM = "parameter_rational.1_4_16_0.00010000_010.000"
RAT_DEN = "parameter_rational.1_2inv_16_0.00010000_010.000"


def _rational(name):
    import os
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "rational", name)
    return dwf_amd.Rational.from_file(path, maximum_number_of_iteration=20000,
                                      convergence_criterion_squared=1e-24)


@pytest.mark.parametrize("dims,Ls", [((4, 4, 2, 4), 4), ((8, 8, 8, 8), 8)])
def test_rational_ratio_force_matches_oracle(dims, Ls):
    from oracle import rational_mult, rational_ratio_force_udiv
    print("case", dims, Ls, flush=True)
    U = lattice.random_gauge(dims, 0.3, 41)
    phi = lattice.random_spinor5(dims, Ls, 42)
    p1 = (0.5, 1.6, 1.5, 0.5)
    p2 = (0.05, 1.6, 1.5, 0.5)
    o1 = OracleDWF(dims, Ls, *p1)
    o2 = OracleDWF(dims, Ls, *p2)
    g1 = dwf_amd.AFopr_Domainwall_5din_eo(dims, Ls, *p1)
    g2 = dwf_amd.AFopr_Domainwall_5din_eo(dims, Ls, *p2)
    for o in (o1, o2, g1, g2):
        o.set_config(U)
    r1, r2 = _rational(RAT_NUM), _rational(RAT_DEN)
    c1, c2 = (r1.a0, r1.bl, r1.cl), (r2.a0, r2.bl, r2.cl)
    # AFopr_Rational::mult
    fphi, fv = g2.new_field(), g2.new_field()
    g2.convert(fphi, phi)
    rm = dwf_amd.AFopr_Rational(g2, r2)
    rm.mult(fv, fphi)
    want, _ = rational_mult(o2, *c2, phi)
    assert rel(g2.reverse(fv), want) < 1e-10
    assert g2.get_mode() == "DdagD"
    # AForce_F_Rational_Ratio::force_udiv and force_core
    force = dwf_amd.AForce_F_Rational_Ratio(g1, g2, r1, r2)
    F = force.new_force()
    force.force_udiv(F, fphi)
    Fu = rational_ratio_force_udiv(o1, o2, c1, c2, phi)
    assert rel(F.to_host(), Fu) < 1e-9
    assert g1.get_mode() == "H"
    force.force_core(F, fphi)
    assert rel(F.to_host(), o1.mult_generator(Fu)) < 1e-9


@pytest.mark.parametrize("dims,rho,nsmear", [((4, 2, 4, 2), 0.12, 2), ((8, 8, 8, 8), 0.1, 2)])
def test_stout_force_matches_oracle(dims, rho, nsmear):
    from oracle import stout_smear, stout_force_udiv, stout_mult_jacobian, mult_generator_gauge
    print("case", dims, rho, nsmear, flush=True)
    U = lattice.random_gauge(dims, 0.4, 61)
    sm = dwf_amd.Director_alt_Smear(dims, rho, nsmear)
    sm.set_config(U)
    lv = [U]
    for _ in range(nsmear):
        lv.append(stout_smear(dims, rho, lv[-1]))
    rng = np.random.default_rng(62)
    S = rng.standard_normal(U.shape)
    dS = dwf_amd.DeviceArray(U.shape)
    dS.from_host(S)
    out = dwf_amd.DeviceArray(U.shape)
    # one level, out of place
    sm.force_udiv(out, dS, 1)
    assert rel(out.to_host(), stout_force_udiv(dims, rho, S, lv[1])) < 1e-12
    # the whole jacobian in place, then the generator with the thin links
    want = stout_mult_jacobian(dims, rho, lv[:-1], S)
    sm.mult_jacobian(dS)
    assert rel(dS.to_host(), want) < 1e-12
    sm.mult_generator(dS)
    assert rel(dS.to_host(), mult_generator_gauge(dims, want, U)) < 1e-12


def test_smeared_fermion_force_matches_oracle():
    from oracle import stout_smear, stout_mult_jacobian, mult_generator_gauge
    dims, Ls, rho, nsmear = (4, 4, 4, 4), 6, 0.1, 2
    U = lattice.random_gauge(dims, 0.4, 63)
    eta = lattice.random_spinor5(dims, Ls, 64)
    prm = (0.05, 1.6, 1.5, 0.5)
    lv = [U]
    for _ in range(nsmear):
        lv.append(stout_smear(dims, rho, lv[-1]))
    orc = OracleDWF(dims, Ls, *prm)
    orc.set_config(lv[-1])
    want = mult_generator_gauge(dims, stout_mult_jacobian(dims, rho, lv[:-1], orc.force_udiv(eta)), U)
    gpu = dwf_amd.AFopr_Domainwall_5din_eo(dims, Ls, *prm)
    sm = dwf_amd.Director_alt_Smear(dims, rho, nsmear)
    gpu.set_config_smeared(sm, U)
    fs = dwf_amd.AForce_F_Smeared_alt(dwf_amd.AForce_F_Domainwall_5din_eo(gpu), sm)
    fe = gpu.new_field()
    gpu.convert(fe, eta)
    F = fs.new_force()
    fs.force_core(F, fe)
    assert rel(F.to_host(), want) < 1e-12
