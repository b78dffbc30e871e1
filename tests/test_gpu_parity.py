"""GPU parity of libdwfhip against the CPU oracle (through the C ABI).

Bars: FP64 operator modes within 1e-13 relative (the reference's Dslash
tolerance); FP32 within 2e-6, FP16/BF16 storage within 4e-3 / 3e-2 relative
(storage rounding of spinors and links, FP32 arithmetic); solvers by the FP64
true residual and by agreement of the solutions with the oracle's.
"""
import numpy as np
import pytest

from oracle import OracleDWF, stout_smear as orc_stout, plaquette as orc_plaq, richardson
import dwf_amd
from dwf_amd import lattice

pytestmark = pytest.mark.gpu

PARAMS = [
    # dims, Ls, mq, M0, b, c, bc
    ((4, 4, 2, 4), 4, 0.02, 1.6, 1.0, 0.0, (1, 1, 1, -1)),
    ((4, 2, 4, 4), 6, 0.05, 1.0, 1.5, 0.5, (1, 1, 1, -1)),
    ((4, 4, 2, 2), 5, 0.3, 1.3, 1.2, 0.3, (-1, 1, -1, -1)),
    ((8, 8, 8, 8), 8, 0.02, 1.6, 1.0, 0.0, (1, 1, 1, -1)),
]
LS16 = [((4, 4, 4, 8), 16, 0.02, 1.6, 1.0, 0.0, (1, 1, 1, -1)),      # matrix-core epilogue
        ((4, 2, 4, 4), 16, 0.05, 1.0, 1.5, 0.5, (1, 1, 1, -1))]
# Ls = 12 (the reference benchmark's Ls): compile-time Ls; a 16-bit site has 8
# threads (6 used) and an FP32 site 16 (12 used)
LS12 = [((4, 4, 4, 8), 12, 0.02, 1.6, 1.5, 0.5, (1, 1, 1, -1))]
# the same with a partial last block: Vh = 216 half sites (blocks of 32 sites at
# FP32 and 16-bit storage, 16 at FP64)
LS12_PARTIAL = [((6, 6, 2, 6), 12, 0.04, 1.6, 1.0, 0.0, (1, 1, 1, -1))]
MODES = ["D", "Ddag", "DdagD", "DDdag", "H", "Hdag", "Deo", "Doe", "Dee", "Doo",
         "Dee_inv", "Doo_inv"]


def rel(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


def pair(prm, precision="double", seed=5, eps=0.5):
    dims, Ls, mq, M0, b, c, bc = prm
    U = lattice.random_gauge(dims, eps, seed)
    orc = OracleDWF(dims, Ls, mq, M0, b, c, bc=bc)
    orc.set_config(U)
    gpu = dwf_amd.AFopr_Domainwall_5din_eo(dims, Ls, mq, M0, b, c, bc, precision)
    gpu.set_config(U)
    return orc, gpu, U


@pytest.mark.parametrize("prm", PARAMS + LS16[1:] + LS12 + LS12_PARTIAL)
def test_modes_fp64(prm):
    print("case", prm, flush=True)
    orc, gpu, U = pair(prm)
    x = lattice.random_spinor5(prm[0], prm[1], 7)
    for mode in MODES:
        for dag in (False, True):
            want = orc.mult_dag(mode, x) if dag else orc.mult(mode, x)
            got = gpu.apply(mode, x, dagger=dag)
            assert rel(got, want) < 1e-13, (mode, dag, rel(got, want))


@pytest.mark.parametrize("prm", PARAMS[:2] + LS16 + LS12 + LS12_PARTIAL)
@pytest.mark.parametrize("precision,tol", [("float", 2e-6), ("half", 4e-3), ("bfloat16", 3e-2)])
def test_modes_reduced_precision(prm, precision, tol):
    orc, gpu, U = pair(prm, precision)
    x = lattice.random_spinor5(prm[0], prm[1], 8)
    for mode in ("D", "Ddag", "DdagD", "Deo", "Doe", "Dee_inv"):
        want = orc.mult(mode, x)
        got = gpu.apply(mode, x)
        assert rel(got, want) < tol, (precision, mode, rel(got, want))


@pytest.mark.parametrize("prm", [PARAMS[1], LS16[1]], ids=["ls6", "ls16"])
def test_modes_fp64_alpha(prm):
    """The Accel alpha (Neff) form, parameter_alpha = 1.3, against the oracle's
    f[0] = mq dm[0] / alpha construction, every mode incl. Dee_inv / Doo_inv."""
    dims, Ls, mq, M0, b, c, bc = prm
    U = lattice.random_gauge(dims, 0.5, 31)
    orc = OracleDWF(dims, Ls, mq, M0, b, c, alpha=1.3, bc=bc)
    orc.set_config(U)
    gpu = dwf_amd.AFopr_Domainwall_5din_eo(dims, Ls, mq, M0, b, c, bc, "double",
                                           parameter_alpha=1.3)
    gpu.set_config(U)
    x = lattice.random_spinor5(dims, Ls, 32)
    for mode in MODES:
        for dag in (False, True):
            want = orc.mult_dag(mode, x) if dag else orc.mult(mode, x)
            got = gpu.apply(mode, x, dagger=dag)
            assert rel(got, want) < 1e-13, (mode, dag, rel(got, want))


def test_richardson_fp32_inner_ls12():
    prm = LS12[0]
    orc, gpu, U = pair(prm)
    gf = dwf_amd.AFopr_Domainwall_5din_eo(*prm, "float")
    gf.set_config(U)
    b = lattice.random_spinor5(prm[0], prm[1], 33)
    gpu.set_mode("DdagD")
    fb, fx = gpu.new_field(), gpu.new_field()
    gpu.convert(fb, b)
    n, d = dwf_amd.ASolver_Richardson(gpu, gf, 100, 1e-24, 10000, 1e-8).solve(fx, fb)
    x = gpu.reverse(fx)
    true = np.linalg.norm(orc.mult("DdagD", x) - b) / np.linalg.norm(b)
    assert n >= 2 and d < 1e-12 and true < 1e-11, (n, d, true)


def test_gm5_and_reflection():
    prm = PARAMS[1]
    orc, gpu, U = pair(prm)
    x = lattice.random_spinor5(prm[0], prm[1], 9)
    w, v = gpu.new_field(), gpu.new_field()
    gpu.convert(w, x)
    gpu.mult_gm5(v, w)
    assert rel(gpu.reverse(v), orc.mult_gm5(x)) < 1e-15
    gpu.mult_gm5R(v, w)
    assert rel(gpu.reverse(v), orc.mult_gm5R(x)) < 1e-15


def test_convert_reverse_roundtrip_and_aliasing():
    prm = PARAMS[1]
    orc, gpu, U = pair(prm)
    x = lattice.random_spinor5(prm[0], prm[1], 10)
    w = gpu.new_field()
    gpu.convert(w, x)
    assert np.array_equal(gpu.reverse(w), x)
    gpu.set_mode("DdagD")
    v = gpu.new_field()
    gpu.mult(v, w)
    ref = gpu.reverse(v)
    gpu.mult(w, w)   # in place
    assert np.array_equal(gpu.reverse(w), ref)


def test_blas():
    prm = PARAMS[1]
    orc, gpu, U = pair(prm)
    x = lattice.random_spinor5(prm[0], prm[1], 11)
    y = lattice.random_spinor5(prm[0], prm[1], 12)
    fx, fy = gpu.new_field(), gpu.new_field()
    gpu.convert(fx, x)
    gpu.convert(fy, y)
    assert abs(fx.norm2() - np.sum(x * x)) / np.sum(x * x) < 1e-14
    assert abs(fx.dot(fy) - np.sum(x * y)) / abs(np.sum(x * y)) < 1e-12
    xc = x.reshape(-1, 2) @ np.array([1, 1j])
    yc = y.reshape(-1, 2) @ np.array([1, 1j])
    want = np.vdot(xc, yc)
    assert abs(fx.dotc(fy) - want) / abs(want) < 1e-12
    fy.axpy(0.37, fx)
    assert rel(gpu.reverse(fy), y + 0.37 * x) < 1e-15
    fy.aypx(-1.5, fx)
    assert rel(gpu.reverse(fy), -1.5 * (y + 0.37 * x) + x) < 1e-15
    fy.scal(2.0)
    fz = gpu.new_field()
    fz.copy_from(fy)
    assert np.array_equal(gpu.reverse(fz), gpu.reverse(fy))
    fz.set_zero()
    assert fz.norm2() == 0.0
    fz.set(0.25)   # AField::set(a): every element equals a
    assert np.array_equal(gpu.reverse(fz), np.full_like(x, 0.25))


def test_cg_fp64_matches_oracle():
    prm = ((4, 4, 4, 4), 6, 0.05, 1.6, 1.0, 0.0, (1, 1, 1, -1))
    orc, gpu, U = pair(prm)
    b = lattice.random_spinor5(prm[0], prm[1], 13)
    xo, no, do = orc.cg(b, "DdagD", niter=3000, stop_cond2=1e-24)
    gpu.set_mode("DdagD")
    fb, fx = gpu.new_field(), gpu.new_field()
    gpu.convert(fb, b)
    solver = dwf_amd.ASolver_CG(gpu, 3000, 1e-24)
    n, d = solver.solve(fx, fb)
    x = gpu.reverse(fx)
    assert n > 0 and abs(n - no) <= 2
    assert d < 1e-11
    true = np.linalg.norm(orc.mult("DdagD", x) - b) / np.linalg.norm(b)
    assert true < 1e-11
    assert rel(x, xo) < 1e-9


def test_cg_nonconvergence_reports_minus_one():
    prm = PARAMS[0]
    orc, gpu, U = pair(prm)
    b = lattice.random_spinor5(prm[0], prm[1], 14)
    gpu.set_mode("DdagD")
    fb, fx = gpu.new_field(), gpu.new_field()
    gpu.convert(fb, b)
    n, d = dwf_amd.ASolver_CG(gpu, 3, 1e-30).solve(fx, fb)
    assert n == -1 and d > 1e-4


def test_richardson_mixed_precision_matches_oracle():
    prm = ((4, 4, 4, 4), 6, 0.05, 1.6, 1.0, 0.0, (1, 1, 1, -1))
    orc, gpu, U = pair(prm)
    gf = dwf_amd.AFopr_Domainwall_5din_eo(prm[0], prm[1], prm[2], prm[3], prm[4], prm[5],
                                          prm[6], "float")
    gf.set_config(U)
    b = lattice.random_spinor5(prm[0], prm[1], 15)
    gpu.set_mode("DdagD")
    fb, fx = gpu.new_field(), gpu.new_field()
    gpu.convert(fb, b)
    solver = dwf_amd.ASolver_Richardson(gpu, gf, 100, 1e-28, 10000, 1e-8)
    n, d = solver.solve(fx, fb)
    x = gpu.reverse(fx)
    assert n >= 2 and d < 1e-13
    true = np.linalg.norm(orc.mult("DdagD", x) - b) / np.linalg.norm(b)
    assert true < 1e-13
    orf = OracleDWF(prm[0], prm[1], prm[2], prm[3], prm[4], prm[5], bc=prm[6], precision="float")
    orf.set_config(U)
    xo, no, do, inner = richardson(orc, orf, b, "DdagD", 100, 1e-28, 10000, 1e-8)
    assert rel(x, xo) < 1e-11
    assert abs(n - no) <= 1


@pytest.mark.parametrize("inner", ["half", "float", "bfloat16", "double"])
@pytest.mark.parametrize("prm", LS16 + LS12 + [((4, 4, 4, 8), 12, 0.02, 1.6, 1.0, 0.0, (1, 1, 1, -1))]
                         + LS12_PARTIAL,
                         ids=["ls16-shamir", "ls16-mobius", "ls12-mobius", "ls12-shamir",
                              "ls12-partial-block"])
def test_richardson_low_precision_inner_matrix_core(prm, inner):
    """Mixed solve with the sync-free inner CG at Ls = 16 and 12 (matrix-core
    epilogue, fused p update + Mp prologue): the FP64 residual reaches the
    stopping condition and the solution is the FP64 CG one."""
    orc, gpu, U = pair(prm)
    gf = dwf_amd.AFopr_Domainwall_5din_eo(prm[0], prm[1], prm[2], prm[3], prm[4], prm[5],
                                          prm[6], inner)
    gf.set_config(U)
    b = lattice.random_spinor5(prm[0], prm[1], 18)
    gpu.set_mode("DdagD")
    fb, fx, fy = gpu.new_field(), gpu.new_field(), gpu.new_field()
    gpu.convert(fb, b)
    n, d = dwf_amd.ASolver_Richardson(gpu, gf, 100, 1e-24, 10000, 1e-4).solve(fx, fb)
    assert n >= 2 and d < 1e-12
    x = gpu.reverse(fx)
    true = np.linalg.norm(orc.mult("DdagD", x) - b) / np.linalg.norm(b)
    assert true < 1e-11
    n2, d2 = dwf_amd.ASolver_CG(gpu, 10000, 1e-24).solve(fy, fb)
    assert rel(x, gpu.reverse(fy)) < 1e-7


def test_schedule_autotuner_picks_a_converging_fastest_candidate():
    """dwf_solver_tune_richardson: every candidate (inner storage x inner
    criterion) is timed; the returned solver is the fastest one that converged
    and solves the system to the FP64 stopping condition."""
    prm = LS16[1]
    orc, gpu, U = pair(prm)
    inner = []
    for prec in ("half", "float"):
        g = dwf_amd.AFopr_Domainwall_5din_eo(prm[0], prm[1], prm[2], prm[3], prm[4], prm[5],
                                             prm[6], prec)
        g.set_config(U)
        inner.append(g)
    gpu.set_mode("DdagD")
    b = lattice.random_spinor5(prm[0], prm[1], 23)
    fb, fx = gpu.new_field(), gpu.new_field()
    gpu.convert(fb, b)
    solver, times, (i, j) = dwf_amd.tune_richardson(gpu, inner, [1e-4, 1e-3], fb, fx,
                                                    100, 1e-24, 10000)
    print("  tuner times", times, "choice", (i, j), flush=True)
    assert times[i, j] > 0 and times[i, j] == times[times > 0].min()
    fy = gpu.new_field()
    n, d = solver.solve(fy, fb)
    assert n >= 1 and d < 1e-12
    x = gpu.reverse(fy)
    assert np.linalg.norm(orc.mult("DdagD", x) - b) / np.linalg.norm(b) < 1e-11


def test_multishift_matches_oracle():
    prm = ((4, 4, 4, 4), 6, 0.05, 1.6, 1.0, 0.0, (1, 1, 1, -1))
    orc, gpu, U = pair(prm)
    b = lattice.random_spinor5(prm[0], prm[1], 16)
    sigma = lattice.log_spaced_shifts(6, 1e-3, 10.0)
    xo, no, do, rc = orc.shift_cg(b, sigma, "DdagD", niter=3000, stop_cond2=1e-24)
    gpu.set_mode("DdagD")
    fb = gpu.new_field()
    gpu.convert(fb, b)
    xs = [gpu.new_field() for _ in sigma]
    n, d = dwf_amd.AShiftsolver_CG(gpu, 3000, 1e-24).solve(xs, sigma, fb)
    assert abs(n - no) <= 2
    assert d < 1e-10
    for s, fx, xr in zip(sigma, xs, xo):
        x = gpu.reverse(fx)
        r = orc.mult("DdagD", x) + s * x - b
        assert np.linalg.norm(r) / np.linalg.norm(b) < 1e-10
        assert rel(x, xr) < 1e-8


def test_multishift_ddagger_mode_matches_oracle():
    """A mode other than DdagD (here DDdag) keeps the reference's form of the
    unshifted update (s += sigma0 p0 and <p0, s> in one pass); DdagD takes
    <p0, s> from |D p0|^2 fused into the operator."""
    prm = ((4, 4, 4, 4), 6, 0.05, 1.6, 1.0, 0.0, (1, 1, 1, -1))
    print("case multishift DDdag", prm, flush=True)
    orc, gpu, U = pair(prm)
    b = lattice.random_spinor5(prm[0], prm[1], 17)
    sigma = lattice.log_spaced_shifts(6, 1e-3, 10.0)
    xo, no, do, rc = orc.shift_cg(b, sigma, "DDdag", niter=3000, stop_cond2=1e-24)
    gpu.set_mode("DDdag")
    fb = gpu.new_field()
    gpu.convert(fb, b)
    xs = [gpu.new_field() for _ in sigma]
    n, d = dwf_amd.AShiftsolver_CG(gpu, 3000, 1e-24).solve(xs, sigma, fb)
    assert abs(n - no) <= 2
    assert d < 1e-10
    for s, fx, xr in zip(sigma, xs, xo):
        x = gpu.reverse(fx)
        r = orc.mult("DDdag", x) + s * x - b
        assert np.linalg.norm(r) / np.linalg.norm(b) < 1e-10
        assert rel(x, xr) < 1e-8


def test_multishift_many_shifts():
    """More shifts than the one-pass update handles (32): the per-shift launch
    path; every shift meets the stopping condition."""
    prm = ((4, 4, 2, 4), 4, 0.1, 1.6, 1.0, 0.0, (1, 1, 1, -1))
    orc, gpu, U = pair(prm)
    b = lattice.random_spinor5(prm[0], prm[1], 20)
    sigma = lattice.log_spaced_shifts(36, 1e-2, 10.0)
    gpu.set_mode("DdagD")
    fb = gpu.new_field()
    gpu.convert(fb, b)
    xs = [gpu.new_field() for _ in sigma]
    n, d = dwf_amd.AShiftsolver_CG(gpu, 3000, 1e-24).solve(xs, sigma, fb)
    assert d < 1e-10
    for s_, fx in zip(sigma, xs):
        x = gpu.reverse(fx)
        r = orc.mult("DdagD", x) + s_ * x - b
        assert np.linalg.norm(r) / np.linalg.norm(b) < 1e-10


def test_stout_smearing_matches_oracle():
    dims = (4, 4, 4, 8)
    U = lattice.random_gauge(dims, 0.5, 17)
    sm = dwf_amd.Director_alt_Smear(dims, 0.1, 2)
    sm.set_config(U)
    ref = U
    for lev in (1, 2):
        ref = orc_stout(dims, 0.1, ref)
        assert np.abs(sm.get_config(lev) - ref).max() < 1e-13
    assert abs(dwf_amd.plaquette(dims, ref) - orc_plaq(dims, ref)) < 1e-13


def test_smeared_operator_matches_oracle():
    prm = PARAMS[1]
    dims, Ls, mq, M0, b, c, bc = prm
    U = lattice.random_gauge(dims, 0.5, 18)
    sm = dwf_amd.Director_alt_Smear(dims, 0.1, 2)
    gpu = dwf_amd.AFopr_Domainwall_5din_eo(dims, Ls, mq, M0, b, c, bc)
    gpu.set_config_smeared(sm, U)
    Us = orc_stout(dims, 0.1, orc_stout(dims, 0.1, U))
    orc = OracleDWF(dims, Ls, mq, M0, b, c, bc=bc)
    orc.set_config(Us)
    x = lattice.random_spinor5(dims, Ls, 19)
    assert rel(gpu.apply("DdagD", x), orc.mult("DdagD", x)) < 1e-13


def test_unknown_mode_is_an_error():
    prm = PARAMS[0]
    orc, gpu, U = pair(prm)
    with pytest.raises(dwf_amd.DwfError) as e:
        gpu.set_mode("DdagX")
    assert e.value.code == 2
    w, v = gpu.new_field(), gpu.new_field()
    with pytest.raises(dwf_amd.DwfError):
        gpu.mult(v, w, "D")      # mult(v, w, mode) only takes the block modes


def test_full_size_deo_and_ddagd_parity():
    """16^3 x 32, Ls = 16 (the mixed-precision solve workload) at full size."""
    dims, Ls = (16, 16, 16, 32), 16
    print("case full-size Deo/DdagD", dims, Ls, flush=True)
    U0 = lattice.random_gauge(dims, 0.3, 20)
    sm = dwf_amd.Director_alt_Smear(dims, 0.1, 2)
    gpu = dwf_amd.AFopr_Domainwall_5din_eo(dims, Ls, 0.01, 1.6, 1.0, 0.0)
    gpu.set_config_smeared(sm, U0)
    Us = sm.get_config(2)
    orc = OracleDWF(dims, Ls, 0.01, 1.6, 1.0, 0.0)
    orc.set_config(Us)
    x = lattice.random_spinor5(dims, Ls, 21)
    assert rel(gpu.apply("Deo", x), orc.mult("Deo", x)) < 1e-13
    assert rel(gpu.apply("DdagD", x), orc.mult("DdagD", x)) < 1e-13
    # smearing at full size against the oracle (one level)
    assert np.abs(sm.get_config(1) - orc_stout(dims, 0.1, U0)).max() < 1e-13


def test_full_size_adjointness():
    """<x, D y> = <D^+ x, y> on the GPU at 16^3 x 32, Ls = 16 (size-independent)."""
    dims, Ls = (16, 16, 16, 32), 16
    U = lattice.random_gauge(dims, 0.3, 22)
    gpu = dwf_amd.AFopr_Domainwall_5din_eo(dims, Ls, 0.01, 1.6, 1.0, 0.0)
    gpu.set_config(U)
    fx, fy, t = gpu.new_field(), gpu.new_field(), gpu.new_field()
    gpu.convert(fx, lattice.random_spinor5(dims, Ls, 23))
    gpu.convert(fy, lattice.random_spinor5(dims, Ls, 24))
    gpu.set_mode("D")
    gpu.mult(t, fy)
    lhs = fx.dotc(t)
    gpu.mult_dag(t, fx)
    rhs = t.dotc(fy)
    assert abs(lhs - rhs) / abs(lhs) < 1e-12
