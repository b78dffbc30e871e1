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
