"""The C++ host mirror (lqcd-dwf-hmc_amd/host/afopr_hip.h) driven like the
reference's benchmark driver: D_eo and a mixed-precision Richardson solve."""
import os
import subprocess

import numpy as np
import pytest

from dwf_amd import lattice

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "lqcd-dwf-hmc_amd", "host", "example_solve")


def test_example_builds_and_rejects_bad_usage():
    assert os.path.exists(EXE), "built by lqcd-dwf-hmc_amd/build.py"
    r = subprocess.run([EXE], capture_output=True, text=True)
    assert r.returncode == 2 and "usage" in r.stderr


@pytest.mark.gpu
def test_example_against_oracle(tmp_path):
    from oracle import OracleDWF
    dims, Ls, mq, M0, b, c = (4, 4, 4, 8), 8, 0.05, 1.6, 1.5, 0.5
    U = lattice.random_gauge(dims, 0.5, 31)
    src = lattice.random_spinor5(dims, Ls, 32)
    U.tofile(tmp_path / "U.bin")
    src.tofile(tmp_path / "src.bin")
    out = str(tmp_path / "out")
    r = subprocess.run([EXE, *map(str, dims), str(Ls), str(mq), str(M0), str(b), str(c),
                        str(tmp_path / "U.bin"), str(tmp_path / "src.bin"), out],
                       capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr)
    assert r.returncode == 0
    orc = OracleDWF(dims, Ls, mq, M0, b, c)
    orc.set_config(U)
    deo = np.fromfile(out + ".deo").reshape(src.shape)
    want = orc.mult("Deo", src)
    assert np.linalg.norm(deo - want) / np.linalg.norm(want) < 1e-13
    x = np.fromfile(out + ".x").reshape(src.shape)
    res = np.linalg.norm(orc.mult("DdagD", x) - src) / np.linalg.norm(src)
    assert res < 1e-13
