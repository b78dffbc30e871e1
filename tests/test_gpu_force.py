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
