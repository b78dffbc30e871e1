#!/usr/bin/env python3
"""Benchmark of the domain-wall hot path on one MI355X.

Workload: the mixed-precision solve on a 16^3 x 32 lattice, Ls = 16,
mq = 0.01, M0 = 1.6, Shamir (b = 1, c = 0), antiperiodic in t, links
U = exp(i 0.3 H) with H Gaussian traceless Hermitian, stout-smeared twice with
rho = 0.1 on the GPU; Gaussian complex source on the even sites; target
|r|/|b| < 1e-10 on the FP64 true residual of DdagD x = b.

One timed step = one solve by FP64 Richardson (the reference's
Fprop_alt_Standard_eo_Richardson outer solver) around a sync-free CG on a
reduced-precision operator; the inner storage type (FP16, FP32, BF16) and
inner |r|^2/|b|^2 criterion (1e-4, 2e-4, 5e-4) are chosen before the timed
steps by the library's schedule autotuner (one timed solve per candidate).  Also measured, outside the timed steps: the plain FP64 CG
solve of the same system, the reference's scheme (FP32 inner CG to 1e-8), the
Richardson with the learned polynomial preconditioner (training time apart),
and the FP64 even-odd Dslash D_eo (one fused hop kernel) timed with HIP events
on the library stream, reported as algorithmic HBM GB/s against the 8 TB/s
MI355X peak.

Prints one JSON line on stdout; progress goes to stderr.
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "lqcd-dwf-hmc_amd"))

METRIC = "time-to-|r|<1e-10, mixed+learned vs FP64 CG; Dslash HBM GB/s"
HBM_PEAK_GBS = 8000.0

DIMS = (16, 16, 16, 32)
LS = 16
MQ, M0, B, CC = 0.01, 1.6, 1.0, 0.0
EPS, RHO, NSMEAR = 0.3, 0.1, 2
TOL2 = 1e-20          # |r|^2 / |b|^2  <->  |r|/|b| < 1e-10
INNER_TOL2 = 1e-8     # fprop_alt_Standard_eo_Richardson-tmpl.h:60 (inner criterion)
INNER_TOL2_FAST = 2e-4   # inner criterion of the timed solver (fastest over 5 sources)


def log(*a):
    print(*a, file=sys.stderr, flush=True)


class HipEvents:
    """hipEvent timing on a given stream through the HIP runtime."""

    def __init__(self):
        for name in ("libamdhip64.so", "/opt/rocm/lib/libamdhip64.so"):
            try:
                self.L = C.CDLL(name)
                break
            except OSError:
                continue
        else:
            raise RuntimeError("libamdhip64.so not found")
        self.L.hipEventCreate.argtypes = [C.POINTER(C.c_void_p)]
        self.L.hipEventRecord.argtypes = [C.c_void_p, C.c_void_p]
        self.L.hipEventSynchronize.argtypes = [C.c_void_p]
        self.L.hipEventElapsedTime.argtypes = [C.POINTER(C.c_float), C.c_void_p, C.c_void_p]
        self.L.hipEventDestroy.argtypes = [C.c_void_p]

    def event(self):
        e = C.c_void_p()
        if self.L.hipEventCreate(C.byref(e)) != 0:
            raise RuntimeError("hipEventCreate failed")
        return e

    def record(self, e, stream):
        if self.L.hipEventRecord(e, C.c_void_p(stream)) != 0:
            raise RuntimeError("hipEventRecord failed")

    def elapsed_ms(self, a, b):
        self.L.hipEventSynchronize(b)
        ms = C.c_float()
        if self.L.hipEventElapsedTime(C.byref(ms), a, b) != 0:
            raise RuntimeError("hipEventElapsedTime failed")
        return ms.value


def deo_bytes(Vh, Ls, elem):
    """Algorithmic bytes of one D_eo: read the input spinor of the other parity,
    all links (8 per output site), write the output spinor."""
    spinor = Vh * Ls * 24 * elem
    links = Vh * 8 * 18 * elem
    return 2 * spinor + links


def cpu_baseline(U_smeared, x, seconds):
    """Oracle D_eo (FP64, corelib-style restatement) on the host cores."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    from oracle import OracleDWF, num_threads
    orc = OracleDWF(DIMS, LS, MQ, M0, B, CC)
    orc.set_config(U_smeared)
    orc.mult("Deo", x)  # warm
    n, t0 = 0, time.perf_counter()
    while True:
        orc.mult("Deo", x)
        n += 1
        if time.perf_counter() - t0 > seconds:
            break
    dt = (time.perf_counter() - t0) / n
    Vh = DIMS[0] * DIMS[1] * DIMS[2] * DIMS[3] // 2
    return {
        "value": round(deo_bytes(Vh, LS, 8) / dt / 1e9, 3),
        "unit": "GB/s",
        "cores": num_threads(),
        "kind": "port",
        "sample": f"{n} calls of the oracle's FP64 D_eo on 16^3x32, Ls=16 "
                  f"({dt * 1e3:.1f} ms/call); same algorithmic bytes as the GPU figure",
    }


def cpu_kernels_and_cfg1(dwf_amd, lattice, np, U_smeared, x, log):
    """Beside the D_eo figure: the oracle's FP64 DdagD at 16^3x32, Ls=16, and the
    reference's CPU-runnable case (8^4, Ls=8 Shamir, M0=1.6, mq=0.02, exp(0.5 iH)
    links with 2 stout steps, FP64 CG to |r|/|b| < 1e-14) solved by the oracle on
    the host and by libdwfhip on the GPU."""
    from oracle import OracleDWF
    orc = OracleDWF(DIMS, LS, MQ, M0, B, CC)
    orc.set_config(U_smeared)
    orc.mult("DdagD", x)
    t0 = time.perf_counter()
    for _ in range(3):
        orc.mult("DdagD", x)
    ddd_ms = (time.perf_counter() - t0) / 3 * 1e3
    dims1, ls1 = (8, 8, 8, 8), 8
    U1 = lattice.random_gauge(dims1, 0.5, 2024)
    sm1 = dwf_amd.Director_alt_Smear(dims1, 0.1, 2)
    sm1.set_config(U1)
    Us1 = sm1.get_config(2)
    b1 = lattice.random_spinor5(dims1, ls1, 7)
    o1 = OracleDWF(dims1, ls1, 0.02, 1.6, 1.0, 0.0)
    o1.set_config(Us1)
    t0 = time.perf_counter()
    xo, no, do = o1.cg(b1, "DdagD", 10000, 1e-28)
    t_cpu = time.perf_counter() - t0
    g1 = dwf_amd.AFopr_Domainwall_5din_eo(dims1, ls1, 0.02, 1.6, 1.0, 0.0, (1, 1, 1, -1), "double")
    g1.set_config_smeared(sm1, None)
    g1.set_mode("DdagD")
    fb, fx = g1.new_field(), g1.new_field()
    g1.convert(fb, b1)
    cg = dwf_amd.ASolver_CG(g1, 10000, 1e-28)
    cg.solve(fx, fb)
    dwf_amd.synchronize()
    t0 = time.perf_counter()
    ng, dg = cg.solve(fx, fb)
    dwf_amd.synchronize()
    t_gpu = time.perf_counter() - t0
    xg = g1.reverse(fx)
    out = {"oracle_ddagd_ms_16x16x16x32_Ls16": round(ddd_ms, 1),
           "cfg1_fp64_cg_8x8x8x8_Ls8": {
               "cpu_oracle_s": round(t_cpu, 4), "cpu_iterations": no,
               "gpu_s": round(t_gpu, 4), "gpu_iterations": ng,
               "gpu_true_residual": dg, "rel_diff_solutions":
                   float(np.linalg.norm(xg - xo) / np.linalg.norm(xo))}}
    log(f"CPU oracle DdagD {ddd_ms:.1f} ms; config 1 CG: CPU {t_cpu:.3f} s ({no} it), "
        f"GPU {t_gpu:.4f} s ({ng} it), solutions differ by "
        f"{out['cfg1_fp64_cg_8x8x8x8_Ls8']['rel_diff_solutions']:.1e}")
    return out


def bench_batch(dwf_amd, lattice, np, f64, f32, sm, log):
    """Batched multishift (16 shifts log-spaced on [1e-4, 10], |r|^2/|b|^2 < 1e-20):
    4 Gaussian right-hand sides on 16^3x32, Ls=16 batched against one multishift
    per right-hand side, with and without graph replay, and in mixed precision;
    then 3 lattices (8^4, 12^3x24, 16^3x32; Ls=12) in grouped launches against
    one solve per lattice."""
    sigma = np.logspace(-4, 1, 16)
    nrhs = 4
    out = {}

    def run(fn):
        fn()   # warm: allocations, graph capture
        dwf_amd.synchronize()
        t0 = time.perf_counter()
        r = fn()
        dwf_amd.synchronize()
        return time.perf_counter() - t0, r

    bs = []
    for k in range(nrhs):
        fb = f64.new_field()
        f64.convert(fb, lattice.random_spinor5(DIMS, LS, 300 + k))
        bs.append(fb)
    xs = [[f64.new_field() for _ in sigma] for _ in range(nrhs)]
    single = dwf_amd.AShiftsolver_CG(f64, 100000, TOL2)

    def seq():
        return [single.solve(xs[k], sigma, bs[k]) for k in range(nrhs)]
    t_seq, r_seq = run(seq)
    log(f"multishift 4 RHS sequential: {t_seq:.3f} s, iterations {[r[0] for r in r_seq]}")
    out["sequential_s"] = round(t_seq, 4)
    for name, graph in (("batched", False), ("batched_graph", True)):
        solver = dwf_amd.AShiftsolver_CG_Batch([f64] * nrhs, len(sigma), graph=graph)
        t_b, (ns, ds) = run(lambda: solver.solve(xs, sigma, bs, 100000, TOL2))
        if max(ds) >= 1e-10 or min(ns) < 0:
            raise RuntimeError(f"batched multishift missed the target: {ns} {ds}")
        log(f"multishift 4 RHS {name}: {t_b:.3f} s, iterations {ns}, max diff {max(ds):.2e}")
        out[name + "_s"] = round(t_b, 4)
        out[name + "_iterations"] = solver.iterations()
        out[name + "_ms_per_iteration"] = round(t_b / solver.iterations() * 1e3, 3)
        del solver
    out["speedup_batched_vs_sequential"] = round(t_seq / out["batched_s"], 3)
    out["speedup_graph_vs_sequential"] = round(t_seq / out["batched_graph_s"], 3)
    # mixed precision: FP32 batched multishift, every shift refined in FP64
    mixed = dwf_amd.AShiftsolver_CG_Batch([f32] * nrhs, len(sigma))
    t_m, (ns, ds) = run(lambda: mixed.solve_mixed([f64] * nrhs, xs, sigma, bs, 100000, TOL2,
                                                  low=1e-10, inner=1e-10))
    if max(ds) >= 1e-10:
        raise RuntimeError(f"mixed batched multishift missed the target: {ds}")
    log(f"multishift 4 RHS mixed FP32+FP64 refinement: {t_m:.3f} s, FP32 iterations {ns}, "
        f"refinement CG iterations {mixed.refine_iterations()}, max diff {max(ds):.2e}")
    out["mixed_fp32_refined_s"] = round(t_m, 4)
    out["mixed_refinement_cg_iterations"] = mixed.refine_iterations()
    out["max_diff_mixed"] = float(f"{max(ds):.3e}")
    del mixed, xs, bs, single

    # three lattices of uneven volume, Ls = 12, one right-hand side each
    lats = [(8, 8, 8, 8), (12, 12, 12, 24), (16, 16, 16, 32)]
    ops, bl, xl = [], [], []
    for i, d in enumerate(lats):
        U = lattice.random_gauge(d, EPS, 400 + i)
        smi = dwf_amd.Director_alt_Smear(d, RHO, NSMEAR)
        op = dwf_amd.AFopr_Domainwall_5din_eo(d, 12, MQ, M0, B, CC, (1, 1, 1, -1), "double")
        op.set_config_smeared(smi, U)
        op.set_mode("DdagD")
        fb = op.new_field()
        op.convert(fb, lattice.random_spinor5(d, 12, 410 + i))
        ops.append(op)
        bl.append(fb)
        xl.append([op.new_field() for _ in sigma])
    singles = [dwf_amd.AShiftsolver_CG(op, 100000, TOL2) for op in ops]
    t_s3, r3 = run(lambda: [singles[i].solve(xl[i], sigma, bl[i]) for i in range(3)])
    grouped = dwf_amd.AShiftsolver_CG_Batch(ops, len(sigma), graph=True)
    t_g3, (ns, ds) = run(lambda: grouped.solve(xl, sigma, bl, 100000, TOL2))
    if max(ds) >= 1e-10 or min(ns) < 0:
        raise RuntimeError(f"grouped multishift missed the target: {ns} {ds}")
    log(f"3 lattices Ls=12: sequential {t_s3:.3f} s, grouped+graph {t_g3:.3f} s, "
        f"iterations {ns}")
    out["three_lattices_ls12"] = {"sequential_s": round(t_s3, 4), "grouped_graph_s": round(t_g3, 4),
                                  "iterations": ns,
                                  "speedup": round(t_s3 / t_g3, 3)}
    return out


def load_traffic(key):
    path = os.path.join(ROOT, "profiles", "traffic.json")
    if not os.path.exists(path):
        return None
    try:
        data = json.load(open(path))
        v = data.get(key, {}).get("bytes_per_launch")
        return None if v is None else float(v)
    except Exception:
        return None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--dslash-reps", type=int, default=100)
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-fp64-cg", action="store_true")
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--no-ptc", action="store_true")
    ap.add_argument("--no-hmc", action="store_true")
    a = ap.parse_args()
    if a.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU only: there is no multi-GPU path (the lattice "
              "would shard by 4D domain decomposition with one halo exchange per hop).",
              file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import dwf_amd
    from dwf_amd import lattice

    t_setup = time.perf_counter()
    U0 = lattice.random_gauge(DIMS, EPS, 2024)
    b_host = lattice.random_spinor5(DIMS, LS, 7)
    sm = dwf_amd.Director_alt_Smear(DIMS, RHO, NSMEAR)
    f64 = dwf_amd.AFopr_Domainwall_5din_eo(DIMS, LS, MQ, M0, B, CC, (1, 1, 1, -1), "double")
    f32 = dwf_amd.AFopr_Domainwall_5din_eo(DIMS, LS, MQ, M0, B, CC, (1, 1, 1, -1), "float")
    f16 = dwf_amd.AFopr_Domainwall_5din_eo(DIMS, LS, MQ, M0, B, CC, (1, 1, 1, -1), "half")
    f64.set_config_smeared(sm, U0)
    f32.set_config_smeared(sm, None)
    f16.set_config_smeared(sm, None)
    for f in (f64, f32, f16):
        f.set_mode("DdagD")
    b = f64.new_field()
    x = f64.new_field()
    f64.convert(b, b_host)
    log(f"setup {time.perf_counter() - t_setup:.1f} s; lattice {DIMS} Ls={LS}")

    # ---- Dslash: FP64 D_eo, HIP events on the library stream ---------------
    ev = HipEvents()
    st = dwf_amd.stream()
    v = f64.new_field()
    for _ in range(5):
        f64.mult(v, b, "Deo")
    dwf_amd.synchronize()
    e0, e1 = ev.event(), ev.event()
    ev.record(e0, st)
    for _ in range(a.dslash_reps):
        f64.mult(v, b, "Deo")
    ev.record(e1, st)
    deo_ms = ev.elapsed_ms(e0, e1) / a.dslash_reps
    Vh = f64.Vh
    deo_b = deo_bytes(Vh, LS, 8)
    deo_gbs = deo_b / (deo_ms * 1e-3) / 1e9
    log(f"D_eo FP64: {deo_ms * 1e3:.1f} us/launch, {deo_gbs:.0f} GB/s algorithmic "
        f"({deo_b / 1e6:.1f} MB/launch), {deo_gbs / HBM_PEAK_GBS * 100:.1f}% of 8 TB/s")
    # FP32 and DdagD figures for the record
    v32, b32 = f32.new_field(), f32.new_field()
    b32.convert_precision_from(b)
    ev.record(e0, st)
    for _ in range(a.dslash_reps):
        f32.mult(v32, b32, "Deo")
    ev.record(e1, st)
    deo32_ms = ev.elapsed_ms(e0, e1) / a.dslash_reps
    v16, b16 = f16.new_field(), f16.new_field()
    b16.convert_precision_from(b)
    ev.record(e0, st)
    for _ in range(a.dslash_reps):
        f16.mult(v16, b16, "Deo")
    ev.record(e1, st)
    deo16_ms = ev.elapsed_ms(e0, e1) / a.dslash_reps
    ev.record(e0, st)
    for _ in range(20):
        f64.mult(v, b)
    ev.record(e1, st)
    ddd_ms = ev.elapsed_ms(e0, e1) / 20
    log(f"D_eo FP32: {deo32_ms * 1e3:.1f} us; DdagD FP64: {ddd_ms * 1e3:.1f} us")
    # error of the reduced-precision operators against FP64 on the same input:
    # |D^prec w - D^64 w| / |D^64 w| for D_eo and DdagD (w = the solve's source)
    op_err = {}
    for name, fp in (("f32", f32), ("f16", f16)):
        for mode in ("Deo", "DdagD"):
            wl, vl, vd = fp.new_field(), fp.new_field(), f64.new_field()
            wl.convert_precision_from(b)
            m = None if mode == "DdagD" else mode   # the operators' mode is DdagD
            fp.mult(vl, wl, m)
            f64.mult(v, b, m)
            vd.convert_precision_from(vl)
            ref2 = v.norm2()
            vd.axpy(-1.0, v)
            op_err[f"{mode}_{name}"] = float((vd.norm2() / ref2) ** 0.5)
    log(f"operator error vs FP64: {op_err}")

    # ---- timed steps: mixed-precision solves --------------------------------
    # FP64 Richardson (the reference's outer solver) around a sync-free CG; the
    # inner storage type and criterion come from the schedule autotuner (one timed
    # solve per candidate, outside the timed steps)
    fbf = dwf_amd.AFopr_Domainwall_5din_eo(DIMS, LS, MQ, M0, B, CC, (1, 1, 1, -1), "bfloat16")
    fbf.set_config_smeared(sm, None)
    fbf.set_mode("DdagD")
    t_tune = time.perf_counter()
    cand_ops, cand_names = [f16, f32, fbf], ["half", "float", "bfloat16"]
    cand_tols = [1e-4, INNER_TOL2_FAST, 5e-4]
    rich, tune_times, (ti, tj) = dwf_amd.tune_richardson(f64, cand_ops, cand_tols, b, x, 100,
                                                         TOL2, 100000)
    t_tune = time.perf_counter() - t_tune
    tuned = {"chosen": {"inner_storage": cand_names[ti], "inner_criterion": cand_tols[tj]},
             "tuning_s": round(t_tune, 2),
             "candidate_solve_ms": {f"{cand_names[a]}/{cand_tols[c]:g}":
                                    (round(float(tune_times[a, c]) * 1e3, 2) if tune_times[a, c] > 0 else None)
                                    for a in range(len(cand_names)) for c in range(len(cand_tols))}}
    log(f"schedule autotuner ({t_tune:.1f} s): {tuned['candidate_solve_ms']} -> "
        f"{cand_names[ti]} inner, {cand_tols[tj]:g}")
    for i in range(a.warmup):
        n, d = rich.solve(x, b)
        log(f"warmup mixed solve: nconv={n} |r|/|b|={d:.3e} inner={rich.inner_iterations()}")
    dwf_amd.synchronize()
    t0 = time.perf_counter()
    results = []
    for i in range(a.steps):
        n, d = rich.solve(x, b)
        results.append((n, d, rich.inner_iterations()))
    dwf_amd.synchronize()
    t_mixed = (time.perf_counter() - t0) / a.steps
    for n, d, inner in results:
        if not (n >= 0 and d < 1e-10):
            raise RuntimeError(f"mixed solve missed the target: nconv={n} |r|/|b|={d}")
    log(f"mixed FP64/{cand_names[ti]}-storage solve: {t_mixed:.3f} s, outer={results[-1][0]}, "
        f"inner CG iterations={results[-1][2]}, |r|/|b|={results[-1][1]:.3e}")

    # ---- comparisons (not in the timed steps) --------------------------------
    def timed(solver):
        solver.solve(x, b)   # warm (allocations)
        dwf_amd.synchronize()
        t0 = time.perf_counter()
        n, d = solver.solve(x, b)
        dwf_amd.synchronize()
        dt = time.perf_counter() - t0
        if not (n >= 0 and d < 1e-10):
            raise RuntimeError(f"solve missed the target: nconv={n} |r|/|b|={d}")
        return dt, n, d, solver.inner_iterations()

    t_fp64 = fp64_iters = fp64_diff = None
    if not a.no_fp64_cg:
        cg = dwf_amd.ASolver_CG(f64, 100000, TOL2, "ZERO")
        t_fp64, fp64_iters, fp64_diff, _ = timed(cg)
        log(f"plain FP64 CG: {t_fp64:.3f} s, {fp64_iters} iterations, |r|/|b|={fp64_diff:.3e}")
    # the same sync-free fused CG iteration in FP64 (one outer step of the FP64
    # Richardson with an FP64 inner operator solving to the full tolerance): the
    # FP64 baseline without host round trips
    t_fp64f = n_fp64f = i_fp64f = None
    if not a.no_fp64_cg:
        f64b = dwf_amd.AFopr_Domainwall_5din_eo(DIMS, LS, MQ, M0, B, CC, (1, 1, 1, -1), "double")
        f64b.set_config_smeared(sm, None)
        f64b.set_mode("DdagD")
        rich64 = dwf_amd.ASolver_Richardson(f64, f64b, 100, TOL2, 100000, TOL2)
        t_fp64f, n_fp64f, _, i_fp64f = timed(rich64)
        log(f"FP64 sync-free fused CG: {t_fp64f:.3f} s, outer={n_fp64f}, inner={i_fp64f}")
    ref = dwf_amd.ASolver_Richardson(f64, f32, 100, TOL2, 100000, INNER_TOL2)
    t_ref, n_ref, d_ref, i_ref = timed(ref)
    log(f"reference scheme FP64/FP32 (inner 1e-8): {t_ref:.3f} s, outer={n_ref}, inner={i_ref}")
    # learned polynomial preconditioner (degree 1, FP16 storage, fitted with FP32)
    lpc = dwf_amd.APrecond_LearnedPolynomial(f16, 1)
    dwf_amd.synchronize()
    t0 = time.perf_counter()
    lpc_loss = lpc.train(f32, nvec=1, seed=11)
    dwf_amd.synchronize()
    t_train = time.perf_counter() - t0
    richl = dwf_amd.ASolver_Richardson_LPC(f64, f32, lpc, 100, TOL2, 100000, INNER_TOL2_FAST)
    t_lpc, n_lpc, d_lpc, i_lpc = timed(richl)
    log(f"learned preconditioner: training {t_train:.3f} s (loss {lpc_loss:.4f}); "
        f"solve {t_lpc:.3f} s, outer={n_lpc}, inner={i_lpc}")

    # learned parallel-transport network (4 layers, 8 channels, bfloat16 weights):
    # trained on the GPU on an 8^3 x 16 lattice with the same action (2048
    # (source, solution) pairs b = D^+D x), weights moved to the 16^3 x 32 network
    ptc = None
    if not a.no_ptc:
        from dwf_amd.ptc import (ASolver_Richardson_PTC, ParallelTransportNet, operator_pairs,
                                 train)
        small = (8, 8, 8, 16)
        sms = dwf_amd.Director_alt_Smear(small, RHO, NSMEAR)
        fs = dwf_amd.AFopr_Domainwall_5din_eo(small, LS, MQ, M0, B, CC, (1, 1, 1, -1), "float")
        fs.set_config_smeared(sms, lattice.random_gauge(small, EPS, 77))
        fs.set_mode("DdagD")
        ch = (1, 8, 8, 8, 1)
        nets = ParallelTransportNet(fs, ch)
        nets.set_weights(nets.identity_weights(scale=0.5, noise=1e-3, seed=1))
        dwf_amd.synchronize()
        t0 = time.perf_counter()
        pairs = operator_pairs(fs, 2048, seed=5)
        hist = train(nets, pairs, steps=500, batch=4, lr=2e-3, seed=3)
        dwf_amd.synchronize()
        t_ptc_train = time.perf_counter() - t0
        del pairs
        netf = ParallelTransportNet(f32, ch)
        netf.set_weights(nets.get_weights())
        # one application at full size, HIP events
        w32, z32 = f32.new_field(), f32.new_field()
        w32.convert_precision_from(b)
        netf.apply(z32, w32)
        ev.record(e0, st)
        for _ in range(3):
            netf.apply(z32, w32)
        ev.record(e1, st)
        ptc_apply_ms = ev.elapsed_ms(e0, e1) / 3
        richp = ASolver_Richardson_PTC(f64, f32, netf, 100, TOL2, 100000, INNER_TOL2_FAST)
        t_p, n_p, d_p, i_p = timed(richp)
        log(f"learned PTC network: training {t_ptc_train:.1f} s (loss {hist[0]:.4f} -> "
            f"{np.mean(hist[-20:]):.4f}); one application {ptc_apply_ms:.2f} ms; solve "
            f"{t_p:.3f} s, outer={n_p}, inner={i_p}")
        # the same Richardson / FCG with the identity network (= CG) for the iteration gain
        netf.set_weights(nets.identity_weights(scale=0.5, noise=0.0))
        t_p0, n_p0, d_p0, i_p0 = timed(ASolver_Richardson_PTC(f64, f32, netf, 100, TOL2, 100000,
                                                               INNER_TOL2_FAST))
        log(f"  same solver, identity network: {t_p0:.3f} s, outer={n_p0}, inner={i_p0}")
        ptc = {"train_s": round(t_ptc_train, 2), "train_pairs": 2048, "train_steps": 500,
               "train_lattice": list(small), "loss_first": round(hist[0], 5),
               "loss_last20": round(float(np.mean(hist[-20:])), 5),
               "apply_ms_full_size": round(ptc_apply_ms, 3),
               "solve_s": round(t_p, 4), "outer_inner": [n_p, i_p],
               "identity_network_solve_s": round(t_p0, 4), "identity_outer_inner": [n_p0, i_p0],
               "inner_iteration_ratio": round(i_p / max(i_p0, 1), 3)}

    # multishift CG (the rational part): 16 shifts log-spaced on [1e-4, 10], FP64
    sigma = np.logspace(-4, 1, 16)
    xs = [f64.new_field() for _ in sigma]
    msolver = dwf_amd.AShiftsolver_CG(f64, 100000, TOL2)
    msolver.solve(xs, sigma, b)
    dwf_amd.synchronize()
    t0 = time.perf_counter()
    n_ms, d_ms = msolver.solve(xs, sigma, b)
    dwf_amd.synchronize()
    t_ms = time.perf_counter() - t0
    del xs
    log(f"multishift CG, 16 shifts: {t_ms:.3f} s, {n_ms} iterations, max rel residual {d_ms:.2e}")

    batch = None
    if not a.no_batch:
        batch = bench_batch(dwf_amd, lattice, np, f64, f32, sm, log)

    # the reference's 2+1-flavour trajectory on 32x16x16x24 (HMC_DWF_Nf2p1_eo.yaml),
    # whose energies and time the reference printed for 4 GPUs
    traj = None
    if not a.no_hmc:
        from dwf_amd.benchmark_hmc_dwf import run_trajectory
        known = json.load(open(os.path.join(ROOT, "tests", "golden", "known_answers.json")))
        ref_s = known["trajectory_seconds_32x16x16x24"]["value"]
        ref_dh = known["H_diff_32x16x16x24"]["value"]
        def leg(**kw):
            o = run_trajectory((32, 16, 16, 24), **kw)
            return {"seconds": round(o["seconds"], 3), "H_diff": o["H_diff"],
                    "reference_H_diff": ref_dh, "abs_H_diff_error": abs(o["H_diff"] - ref_dh),
                    "reference_seconds_4_gpus": ref_s,
                    "speedup_vs_reference_4_gpus": round(ref_s / o["seconds"], 3)}
        # the reference's scheme (Richardson solves with an FP32 inner CG to 1e-8),
        # then the same trajectory with the inner CG in FP16 storage to 2e-4
        traj = leg()
        traj["fp16_inner"] = leg(inner_precision="half", inner_stop=2e-4)
        for name, t in (("FP32 inner", traj), ("FP16 inner", traj["fp16_inner"])):
            log(f"HMC trajectory 32x16x16x24 Ls=12, {name}: {t['seconds']:.2f} s, H(diff) = "
                f"{t['H_diff']:.8f} (reference {ref_dh}, {ref_s} s on 4 GPUs)")


    cpu = None
    if not a.no_cpu_baseline:
        Us = sm.get_config(NSMEAR)
        cpu = cpu_baseline(Us, b_host, a.cpu_seconds)
        cpu.update(cpu_kernels_and_cfg1(dwf_amd, lattice, np, Us, b_host, log))
        log(f"CPU baseline: {cpu}")

    traffic = load_traffic("k_dwf_hop_f64_16x16x16x32_Ls16")
    out = {
        "metric": METRIC,
        "value": round(deo_gbs, 2),
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(t_mixed * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": "mixed-precision solve DdagD x = b to |r|/|b|<1e-10, 16^3x32, Ls=16, "
                        "mq=0.01, M0=1.6, Shamir, exp(0.3 iH) links + 2 stout(0.1) steps",
            "lattice": list(DIMS), "Ls": LS,
            "step": "one FP64-Richardson / low-precision-CG solve, inner storage and "
                    "|r|^2/|b|^2 criterion chosen by the schedule autotuner",
            "value": "FP64 D_eo algorithmic HBM GB/s",
        },
        "time_to_solution_s": {
            "fp64_cg": None if t_fp64 is None else round(t_fp64, 4),
            "fp64_cg_iterations": fp64_iters,
            "fp64_cg_sync_free_fused": None if t_fp64f is None else round(t_fp64f, 4),
            "fp64_cg_sync_free_fused_iterations": i_fp64f,
            "mixed_fp64_tuned_inner": round(t_mixed, 4),
            "mixed_fp64_tuned_inner_outer_inner": [results[-1][0], results[-1][2]],
            "reference_scheme_fp64_fp32_inner_1e-8": round(t_ref, 4),
            "mixed_learned": round(t_lpc, 4),
            "learned_training": round(t_train, 4),
            "learned_outer_inner": [n_lpc, i_lpc],
            "speedup_mixed_vs_fp64": None if t_fp64 is None else round(t_fp64 / t_mixed, 3),
            "speedup_mixed_vs_fp64_sync_free_fused":
                None if t_fp64f is None else round(t_fp64f / t_mixed, 3),
            "speedup_reference_scheme_vs_fp64": None if t_fp64 is None else round(t_fp64 / t_ref, 3),
            "speedup_learned_vs_fp64_incl_training":
                None if t_fp64 is None else round(t_fp64 / (t_lpc + t_train), 3),
        },
        "schedule_autotuner": tuned,
        "multishift_16_shifts_fp64": {"time_s": round(t_ms, 4), "iterations": n_ms,
                                      "ms_per_iteration": round(t_ms / max(n_ms, 1) * 1e3, 3),
                                      "max_rel_residual": d_ms},
        "multishift_16_shifts_4_rhs_batched": batch,
        "learned_ptc_network": ptc,
        "hmc_trajectory_32x16x16x24_Ls12": traj,
        "operator_rel_error_vs_fp64": {k: float(f"{e:.3e}") for k, e in op_err.items()},
        # the hop the timed solve spends its time in: FP16 storage D_eo, algorithmic
        # bytes (16-bit spinors and links) per launch over its HIP-event time
        "inner_hop_fp16_deo": {
            "us": round(deo16_ms * 1e3, 2),
            "bytes_per_launch": deo_bytes(Vh, LS, 2),
            "achieved_GBs": round(deo_bytes(Vh, LS, 2) / (deo16_ms * 1e-3) / 1e9, 1),
            "frac_of_hbm_peak": round(deo_bytes(Vh, LS, 2) / (deo16_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
        },
        "kernels_us": {"deo_f64": round(deo_ms * 1e3, 2), "deo_f32": round(deo32_ms * 1e3, 2),
                       "ddagd_f64": round(ddd_ms * 1e3, 2)},
        "roofline": {
            "bound": "hbm",
            "achieved": round(deo_gbs, 2),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(deo_gbs / HBM_PEAK_GBS, 4),
            "traffic": traffic,
            "kernel": "k_dwf hop FP64 (D_eo, output parity even)",
            "bytes_per_launch": deo_b,
        },
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
