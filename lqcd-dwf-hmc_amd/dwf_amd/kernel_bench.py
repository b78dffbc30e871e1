"""Micro-benchmark of single operator kernels of libdwfhip (and one plain copy
of known size, the calibration for hardware byte counters).

    python3 lqcd-dwf-hmc_amd/dwf_amd/kernel_bench.py [--mode Deo] [--prec double]
                                                     [--reps 50] [--dims 16 16 16 32] [--Ls 16]
                                                     [--lpc-train DEGREE]
                                                     [--ptc 1 8 8 8 1]

Runs <reps> launches of the operator mode (and <reps> field copies) on the
library stream, timed with HIP events, and prints one line per measurement.
The workload is the benchmark lattice: links exp(0.3 i H), Shamir b = 1, c = 0,
mq = 0.01, M0 = 1.6, antiperiodic in t.  --lpc-train also trains the learned
polynomial preconditioner of that degree once on the operator (its Gram
matrices run on the matrix cores for an FP32 operator at Ls <= 16).  --ptc
CHANNELS also times one application of the parallel-transport network with
those channel counts (FP32 operator; identity weights plus noise 0.05).
"""
from __future__ import annotations

import argparse
import ctypes as C
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))

import dwf_amd  # noqa: E402
from dwf_amd import lattice  # noqa: E402


class _Events:
    def __init__(self):
        self.L = C.CDLL("libamdhip64.so")
        self.L.hipEventCreate.argtypes = [C.POINTER(C.c_void_p)]
        self.L.hipEventRecord.argtypes = [C.c_void_p, C.c_void_p]
        self.L.hipEventSynchronize.argtypes = [C.c_void_p]
        self.L.hipEventElapsedTime.argtypes = [C.POINTER(C.c_float), C.c_void_p, C.c_void_p]

    def event(self):
        e = C.c_void_p()
        assert self.L.hipEventCreate(C.byref(e)) == 0
        return e

    def time(self, fn, reps, stream):
        a, b = self.event(), self.event()
        fn()
        self.L.hipEventRecord(a, C.c_void_p(stream))
        for _ in range(reps):
            fn()
        self.L.hipEventRecord(b, C.c_void_p(stream))
        self.L.hipEventSynchronize(b)
        ms = C.c_float()
        self.L.hipEventElapsedTime(C.byref(ms), a, b)
        return ms.value * 1e3 / reps


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--mode", default="Deo")
    ap.add_argument("--dag", action="store_true")
    ap.add_argument("--prec", default="double")
    ap.add_argument("--reps", type=int, default=50)
    ap.add_argument("--dims", type=int, nargs=4, default=[16, 16, 16, 32])
    ap.add_argument("--Ls", type=int, default=16)
    ap.add_argument("--lpc-train", type=int, default=0)
    ap.add_argument("--ptc", type=int, nargs="+", default=None)
    a = ap.parse_args()
    dims = tuple(a.dims)
    op = dwf_amd.AFopr_Domainwall_5din_eo(dims, a.Ls, 0.01, 1.6, 1.0, 0.0, (1, 1, 1, -1), a.prec)
    op.set_config(lattice.random_gauge(dims, 0.3, 1))
    w, v = op.new_field(), op.new_field()
    op.convert(w, lattice.random_spinor5(dims, a.Ls, 2))
    ev = _Events()
    st = dwf_amd.stream()
    elem = {"double": 8, "float": 4}.get(a.prec, 2)
    field_bytes = op.Vh * a.Ls * 24 * elem
    if a.mode in ("Deo", "Doe", "Dee", "Doo", "Dee_inv", "Doo_inv"):
        run = (lambda: op.mult_dag(v, w, a.mode)) if a.dag else (lambda: op.mult(v, w, a.mode))
    else:   # D, Ddag, DdagD, H, ...: the operator's current mode
        op.set_mode(a.mode)
        run = (lambda: op.mult_dag(v, w)) if a.dag else (lambda: op.mult(v, w))
    us = ev.time(run, a.reps, st)
    print(f"mode {a.mode}{'^dag' if a.dag else ''} prec {a.prec} dims {dims} Ls {a.Ls}: "
          f"{us:.2f} us/launch", flush=True)
    if a.lpc_train:
        lpc = dwf_amd.APrecond_LearnedPolynomial(op, a.lpc_train)
        loss = lpc.train(op, nvec=2)
        dwf_amd.synchronize()
        print(f"learned polynomial degree {a.lpc_train} trained: loss {loss:.4f}", flush=True)
    if a.ptc:
        from dwf_amd.ptc import ParallelTransportNet
        net = ParallelTransportNet(op, a.ptc)
        net.set_weights(net.identity_weights(scale=0.5, noise=0.05, seed=1))
        us = ev.time(lambda: net.apply(v, w), a.reps, st)
        print(f"network {tuple(a.ptc)}: {us / 1e3:.2f} ms/application", flush=True)
    us = ev.time(lambda: v.copy_from(w), a.reps, st)
    print(f"copy of one field ({field_bytes} bytes read, {field_bytes} written): "
          f"{us:.2f} us/launch, {2 * field_bytes / us / 1e3:.0f} GB/s", flush=True)


if __name__ == "__main__":
    main()
