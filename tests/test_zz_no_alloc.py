"""dwf_fopr_mult_noalloc (the entry point that allocates nothing) against
dwf_fopr_mult, called through the C ABI: identical results on two calls in a
row, the second on a smaller operator and input, for every operator mode;
an operator whose intermediate fields exceed the static scratch is rejected."""
import ctypes as C

import numpy as np
import pytest

import dwf_amd
from dwf_amd import _lib, lattice

pytestmark = pytest.mark.gpu

MODES = ["D", "Ddag", "DdagD", "DDdag", "H", "Hdag", "Deo", "Doe", "Dee", "Doo",
         "Dee_inv", "Doo_inv"]


def _op(dims, Ls, precision="double", b=1.5, c=0.5, seed=3):
    U = lattice.random_gauge(dims, 0.5, seed)
    op = dwf_amd.AFopr_Domainwall_5din_eo(dims, Ls, 0.03, 1.6, b, c, (1, 1, 1, -1), precision)
    op.set_config(U)
    return op


def _check(op, x, alias=False):
    L = _lib.load()
    w, v_ref, v_na = op.new_field(), op.new_field(), op.new_field()
    op.convert(w, x)
    for mode in MODES:
        op.set_mode(mode)
        _lib.check("dwf_fopr_mult", L.dwf_fopr_mult(op._h, v_ref.ptr, w.ptr))
        if alias:   # v == w: the aliased input goes through the scratch too
            v_na.copy_from(w)
            rc = L.dwf_fopr_mult_noalloc(op._h, v_na.ptr, v_na.ptr)
        else:
            rc = L.dwf_fopr_mult_noalloc(op._h, v_na.ptr, w.ptr)
        _lib.check("dwf_fopr_mult_noalloc", rc)
        assert np.array_equal(op.reverse(v_na), op.reverse(v_ref)), mode


def test_noalloc_matches_mult_two_calls():
    big = _op((8, 8, 8, 8), 16)
    _check(big, lattice.random_spinor5((8, 8, 8, 8), 16, 4))
    small = _op((4, 4, 4, 8), 8, seed=5)
    _check(small, lattice.random_spinor5((4, 4, 4, 8), 8, 6), alias=True)
    # the first operator again after the smaller one (the scratch keeps old contents)
    _check(big, lattice.random_spinor5((8, 8, 8, 8), 16, 7))


def test_noalloc_reduced_precision():
    op = _op((4, 4, 4, 8), 16, "half", b=1.0, c=0.0)
    _check(op, lattice.random_spinor5((4, 4, 4, 8), 16, 8))


def test_noalloc_rejects_oversize():
    op = _op((16, 16, 16, 32), 16, "float", b=1.0, c=0.0)   # 5 x 100.7 MB > 256 MiB
    w, v = op.new_field(), op.new_field()
    op.set_mode("DdagD")
    rc = _lib.load().dwf_fopr_mult_noalloc(op._h, v.ptr, w.ptr)
    assert rc == 1
    assert b"maximum" in _lib.load().dwf_last_error()
