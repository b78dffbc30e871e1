"""Runs the HMC-side kernels of libdwfhip a few times each on one lattice, for
kernel traces and hardware counters, and prints the algorithmic bytes of one
launch of each (what the kernel must read and write once):

* k_stout               one stout level: read U, write U' (gauge field = 4 V 18 doubles)
* k_stout_force_link    per level: read U, Sigma'; write Sigma, iTheta
* k_stout_force_staple  per level: read U, iTheta, Sigma; write Sigma
* k_force_tensor        two 5D half spinors, read and write the force rows of one parity
* k_shift_update_all    16 shifts: shifts 1..15 (x_j, p_j read and written, r read once;
                        shift 0 is the unshifted CG's own x, p)
* k_rect_force          read U, write F
* k_exp_update          read U, iP, write U
* k_copy<double>        one 5D half spinor copied (the calibration of the byte counters)

    python3 lqcd-dwf-hmc_amd/dwf_amd/hmc_kernel_bench.py [--dims 32 16 16 24] [--Ls 12] [--reps 3]

Links exp(0.3 i H), 3 stout levels with rho = 0.1, Mobius b = 1.5, c = 0.5,
M0 = 1, mq = 0.01, antiperiodic in t.
"""
from __future__ import annotations

import argparse
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))

import numpy as np  # noqa: E402

import dwf_amd  # noqa: E402
from dwf_amd import hmc, lattice  # noqa: E402


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--dims", type=int, nargs=4, default=[32, 16, 16, 24])
    ap.add_argument("--Ls", type=int, default=12)
    ap.add_argument("--reps", type=int, default=3)
    a = ap.parse_args()
    dims, Ls, reps = tuple(a.dims), a.Ls, a.reps
    V = int(np.prod(dims))
    G = 4 * V * 18 * 8
    half = V // 2 * 24 * Ls * 8
    print(f"lattice {dims} Ls {Ls}: gauge field {G} B, 5D half spinor {half} B", flush=True)

    U0 = lattice.random_gauge(dims, 0.3, 1)
    sm = dwf_amd.Director_alt_Smear(dims, 0.1, 3)
    for _ in range(reps):
        sm.set_config(U0)
    dwf_amd.synchronize()
    print(f"k_stout: {2 * G} B per launch (3 per set_config)", flush=True)

    op = dwf_amd.AFopr_Domainwall_5din_eo(dims, Ls, 0.01, 1.0, 1.5, 0.5, (1, 1, 1, -1), "double")
    op.set_config_smeared(sm, None)
    eta, zeta = op.new_field(), op.new_field()
    op.convert(eta, lattice.random_spinor5(dims, Ls, 2))
    op.convert(zeta, lattice.random_spinor5(dims, Ls, 3))
    force = dwf_amd.AForce_F_Domainwall_5din_eo(op)
    F = force.new_force()
    force.set_mode("H")
    for _ in range(reps):
        force.force_udiv1(F, zeta, eta)
    dwf_amd.synchronize()
    print(f"k_force_tensor: {2 * half + G} B per launch (2 per force_udiv1)", flush=True)
    for _ in range(reps):
        for level in (2, 1, 0):
            sm.force_udiv(F, F, level)
    dwf_amd.synchronize()
    print(f"k_stout_force_link: {4 * G} B, k_stout_force_staple: {4 * G} B per launch", flush=True)

    op.set_mode("DdagD")
    sigma = np.logspace(-4, 1, 16)
    xs = [op.new_field() for _ in sigma]
    nit = 10
    ms = dwf_amd.AShiftsolver_CG(op, nit, 1e-40)
    ms.solve(xs, sigma, eta, raise_on_nonconvergence=False)
    dwf_amd.synchronize()
    print(f"k_shift_update_all: {(4 * 15 + 1) * half} B per launch ({nit} launches, shifts 1..15 "
          f"active)", flush=True)
    del xs

    Ug = hmc.GaugeField(dims)
    Ug.from_host(U0)
    rect = hmc.Action_G_Rectangle(dims, 4.17, 5.0 / 3.0, -1.0 / 12.0)
    rect.set_config(Ug)
    Fg, P = hmc.GaugeField(dims), hmc.GaugeField(dims)
    for _ in range(reps):
        rect.force(Fg)
    dwf_amd.synchronize()
    print(f"k_rect_force: {2 * G} B per launch", flush=True)
    P.from_host(np.zeros((4, V, 18)))
    P.axpy(1e-3, Fg)
    upd = hmc.Integrator_UpdateU([], 8)
    for _ in range(reps):
        upd.evolve(0.01, P, Ug)
    dwf_amd.synchronize()
    print(f"k_exp_update: {3 * G} B per launch", flush=True)

    v = op.new_field()
    for _ in range(reps):
        v.copy_from(eta)
    dwf_amd.synchronize()
    print(f"k_copy<double>: {2 * half} B per launch ({half} read, {half} written)", flush=True)


if __name__ == "__main__":
    main()
