"""CPU-side checks of the C ABI: the library loads and exports exactly what
include/dwf_hip.h declares; without a GPU every entry point fails cleanly."""
import ctypes as C
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "dwf_hip.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(dwf_[A-Za-z0-9_]+)\s*\(", text)))


def test_header_declares_the_boundary():
    names = declared_functions()
    for n in ("dwf_fopr_create", "dwf_fopr_set_config", "dwf_fopr_set_mode", "dwf_fopr_mult",
              "dwf_fopr_mult_dag", "dwf_fopr_convert", "dwf_fopr_reverse",
              "dwf_solver_solve", "dwf_shiftsolver_cg_solve", "dwf_smear_set_config"):
        assert n in names


def test_library_exports_every_declared_symbol():
    from dwf_amd import _lib
    L = _lib.load()
    names = declared_functions()
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing
    # the binding covers the whole header
    assert sorted(_lib.exported_symbols()) == names


def test_no_device_fails_cleanly():
    """Without a GPU, create returns DWF_ERR_NO_DEVICE with a message (no exit)."""
    import torch  # noqa: F401  (only to know whether a GPU is visible)
    if torch.cuda.is_available():
        pytest.skip("a GPU is visible")
    from dwf_amd import _lib
    L = _lib.load()
    prm = _lib.FoprParams()
    for i in range(4):
        prm.lattice_size[i] = 4
        prm.boundary_condition[i] = 1
    prm.extent_of_5th_dimension = 4
    prm.domain_wall_height = 1.6
    prm.b_uniform = 1.0
    h = C.c_void_p()
    rc = L.dwf_fopr_create(C.byref(prm), C.byref(h))
    assert rc == 7
    assert b"device" in L.dwf_last_error()
