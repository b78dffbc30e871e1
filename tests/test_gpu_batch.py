"""Batched multishift CG on the GPU against the oracle's AShiftsolver_CG.

Several right-hand sides on one operator (multi-RHS hops, padded to a power of
two), lattices of different volume in grouped launches, the captured-graph
iteration, and FP32 operators.  Every solution is compared with the oracle's
multishift solution for the same source, every true residual recomputed with
the oracle's FP64 D^+D; iteration counts agree with the oracle's within 2.
"""
import numpy as np
import pytest

from oracle import OracleDWF
import dwf_amd
from dwf_amd import lattice

pytestmark = pytest.mark.gpu


def rel(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


def make(dims, Ls, mq, M0, b, c, seed, precision="double", bc=(1, 1, 1, -1)):
    U = lattice.random_gauge(dims, 0.5, seed)
    orc = OracleDWF(dims, Ls, mq, M0, b, c, bc=bc)
    orc.set_config(U)
    gpu = dwf_amd.AFopr_Domainwall_5din_eo(dims, Ls, mq, M0, b, c, bc, precision)
    gpu.set_config(U)
    gpu.set_mode("DdagD")
    return orc, gpu


def check_system(orc, gpu, b, sigma, xs, n, stop, tol_res, tol_x):
    xo, no, do, rc = orc.shift_cg(b, sigma, "DdagD", niter=5000, stop_cond2=stop)
    assert abs(n - no) <= 2, (n, no)
    for s_, fx, xr in zip(sigma, xs, xo):
        x = gpu.reverse(fx)
        r = orc.mult("DdagD", x) + s_ * x - b
        assert np.linalg.norm(r) / np.linalg.norm(b) < tol_res, (s_, np.linalg.norm(r))
        assert rel(x, xr) < tol_x, (s_, rel(x, xr))


@pytest.mark.parametrize("prm,nrhs", [
    (((4, 4, 4, 4), 6, 0.05, 1.6, 1.0, 0.0), 3),      # padded to 4 right-hand sides
    (((4, 4, 4, 8), 16, 0.05, 1.6, 1.5, 0.5), 4),     # matrix-core epilogue, Moebius
    (((4, 4, 4, 8), 12, 0.05, 1.6, 1.0, 0.0), 2),
], ids=["ls6-3rhs", "ls16-mobius-4rhs", "ls12-2rhs"])
def test_batch_multi_rhs_matches_oracle(prm, nrhs):
    dims, Ls = prm[0], prm[1]
    print("case batch multi-rhs", prm, nrhs, flush=True)
    orc, gpu = make(*prm, seed=51)
    sigma = lattice.log_spaced_shifts(6, 1e-3, 10.0)
    bs = [lattice.random_spinor5(dims, Ls, 60 + k) for k in range(nrhs)]
    fbs = []
    for b in bs:
        f = gpu.new_field()
        gpu.convert(f, b)
        fbs.append(f)
    xs = [[gpu.new_field() for _ in sigma] for _ in bs]
    solver = dwf_amd.AShiftsolver_CG_Batch([gpu] * nrhs, len(sigma))
    ns, ds = solver.solve(xs, sigma, fbs, 5000, 1e-24)
    for k in range(nrhs):
        assert ds[k] < 1e-10
        check_system(orc, gpu, bs[k], sigma, xs[k], ns[k], 1e-24, 1e-10, 1e-8)


def test_batch_three_lattices_grouped_launch():
    """Three lattices of different volume, Ls = 12, one grouped launch per hop;
    one of them with two right-hand sides."""
    specs = [((4, 4, 4, 4), 0.05), ((4, 4, 4, 8), 0.04), ((8, 4, 4, 4), 0.06)]
    sigma = lattice.log_spaced_shifts(5, 1e-3, 10.0)
    ops, orcs, bs, fbs = [], [], [], []
    for i, (dims, mq) in enumerate(specs):
        orc, gpu = make(dims, 12, mq, 1.6, 1.0, 0.0, seed=70 + i)
        nr = 2 if i == 1 else 1
        for k in range(nr):
            b = lattice.random_spinor5(dims, 12, 80 + 10 * i + k)
            f = gpu.new_field()
            gpu.convert(f, b)
            ops.append(gpu)
            orcs.append(orc)
            bs.append(b)
            fbs.append(f)
    xs = [[g.new_field() for _ in sigma] for g in ops]
    solver = dwf_amd.AShiftsolver_CG_Batch(ops, len(sigma))
    ns, ds = solver.solve(xs, sigma, fbs, 5000, 1e-24)
    for k in range(len(ops)):
        print("  system", k, "nconv", ns[k], "diff", ds[k], flush=True)
        assert ds[k] < 1e-10
        check_system(orcs[k], ops[k], bs[k], sigma, xs[k], ns[k], 1e-24, 1e-10, 1e-8)


def test_batch_graph_replay_matches_direct():
    """The iteration replayed from a captured graph gives the same solutions
    (same kernels, same order) as the directly launched one, on two solves in
    a row (the second with other sources)."""
    prm = ((4, 4, 4, 8), 16, 0.05, 1.6, 1.0, 0.0)
    orc, gpu = make(*prm, seed=91)
    sigma = lattice.log_spaced_shifts(4, 1e-2, 1.0)
    direct = dwf_amd.AShiftsolver_CG_Batch([gpu] * 2, len(sigma))
    graph = dwf_amd.AShiftsolver_CG_Batch([gpu] * 2, len(sigma), graph=True)
    for rep in range(2):
        bs = [lattice.random_spinor5(prm[0], prm[1], 92 + 2 * rep + k) for k in range(2)]
        fbs = []
        for b in bs:
            f = gpu.new_field()
            gpu.convert(f, b)
            fbs.append(f)
        xa = [[gpu.new_field() for _ in sigma] for _ in bs]
        xb = [[gpu.new_field() for _ in sigma] for _ in bs]
        na, da = direct.solve(xa, sigma, fbs, 5000, 1e-22)
        nb, db = graph.solve(xb, sigma, fbs, 5000, 1e-22)
        assert na == nb
        for k in range(2):
            for j in range(len(sigma)):
                assert np.array_equal(gpu.reverse(xa[k][j]), gpu.reverse(xb[k][j]))
            check_system(orc, gpu, bs[k], sigma, xb[k], nb[k], 1e-22, 1e-9, 1e-7)


def test_batch_fp32_operator():
    prm = ((4, 4, 4, 8), 16, 0.05, 1.6, 1.0, 0.0)
    orc, gpu = make(*prm, seed=95, precision="float")
    sigma = lattice.log_spaced_shifts(4, 1e-2, 10.0)
    bs = [lattice.random_spinor5(prm[0], prm[1], 96 + k) for k in range(4)]
    fbs = []
    for b in bs:
        f = gpu.new_field()
        gpu.convert(f, b)
        fbs.append(f)
    xs = [[gpu.new_field() for _ in sigma] for _ in bs]
    ns, ds = dwf_amd.AShiftsolver_CG_Batch([gpu] * 4, len(sigma)).solve(xs, sigma, fbs, 5000,
                                                                        1e-12)
    for k in range(4):
        for s_, fx in zip(sigma, xs[k]):
            x = gpu.reverse(fx)
            r = orc.mult("DdagD", x) + s_ * x - bs[k]
            assert np.linalg.norm(r) / np.linalg.norm(bs[k]) < 2e-5


def test_batch_nonconvergence_reports_minus_one():
    prm = ((4, 4, 2, 4), 4, 0.02, 1.6, 1.0, 0.0)
    orc, gpu = make(*prm, seed=97)
    b = gpu.new_field()
    gpu.convert(b, lattice.random_spinor5(prm[0], prm[1], 98))
    sigma = [1e-3, 1e-1]
    xs = [[gpu.new_field() for _ in sigma]]
    ns, ds = dwf_amd.AShiftsolver_CG_Batch([gpu], 2).solve(xs, sigma, [b], 3, 1e-30,
                                                           raise_on_nonconvergence=False)
    assert ns == [-1] and ds[0] > 1e-4


def test_batch_mixed_precision_refined_to_fp64():
    """FP32 batched multishift, then every shift refined in FP64 (Richardson
    with a shifted FP32 CG): FP64 true residuals below 1e-10."""
    prm = ((4, 4, 4, 8), 16, 0.05, 1.6, 1.0, 0.0)
    dims, Ls = prm[0], prm[1]
    U = lattice.random_gauge(dims, 0.5, 99)
    orc = OracleDWF(dims, Ls, *prm[2:])
    orc.set_config(U)
    g64 = dwf_amd.AFopr_Domainwall_5din_eo(dims, Ls, *prm[2:], (1, 1, 1, -1), "double")
    g32 = dwf_amd.AFopr_Domainwall_5din_eo(dims, Ls, *prm[2:], (1, 1, 1, -1), "float")
    for g in (g64, g32):
        g.set_config(U)
        g.set_mode("DdagD")
    sigma = lattice.log_spaced_shifts(5, 1e-3, 10.0)
    bs = [lattice.random_spinor5(dims, Ls, 100 + k) for k in range(2)]
    fbs = []
    for b in bs:
        f = g64.new_field()
        g64.convert(f, b)
        fbs.append(f)
    xs = [[g64.new_field() for _ in sigma] for _ in bs]
    solver = dwf_amd.AShiftsolver_CG_Batch([g32] * 2, len(sigma))
    ns, ds = solver.solve_mixed([g64] * 2, xs, sigma, fbs, 5000, 1e-22)
    print("  refine iterations", solver.refine_iterations(), "diff", ds, flush=True)
    for k in range(2):
        assert ds[k] < 1e-11
        xo, no, do, rc = orc.shift_cg(bs[k], sigma, "DdagD", niter=5000, stop_cond2=1e-24)
        for s_, fx, xr in zip(sigma, xs[k], xo):
            x = g64.reverse(fx)
            r = orc.mult("DdagD", x) + s_ * x - bs[k]
            assert np.linalg.norm(r) / np.linalg.norm(bs[k]) < 1e-11
            assert rel(x, xr) < 1e-8
