"""HMC on the GPU: the gauge kernels against the oracle, and the reference's
own 8^4 trajectory against the energies it printed.

The trajectory is benchmark_hmc_dwf with HMC_DWF_Nf2p1_eo.yaml (2+1 flavours:
rational ratio PV/Nf1, Hasenbusch ratios Nf2_prec/Nf2 and PV/Nf2_prec, Ls = 12,
b = 1.5, c = 0.5, M0 = 1, 3 stout levels, rectangle gauge action, Omelyan
[2, 1, 3], tau = 0.1) from conf_04040408.txt tiled to 8^4 with
RandomNumbers_parallel<Mseries> reset to 200000: sample_1gpu/run.log:1512-1994.
Bars: the Langevin energies to 1e-10 relative (they are sums of squares of the
reference's random numbers), the trial energies to 1e-6 absolute and
H(diff) = 0.00083995 to 1e-6 absolute (the reference prints 8 decimals).
"""
import json
import os

import numpy as np
import pytest

import gauge_oracle as go
import dwf_amd
from dwf_amd import hmc, lattice

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
KNOWN = json.load(open(os.path.join(HERE, "golden", "known_answers.json")))
RATIONAL = os.path.join(HERE, "golden", "rational")
DIMS8 = (8, 8, 8, 8)
BETA, CP, CR = 4.17, 1.66666666666666666666, -0.08333333333333333333


def conf8():
    return lattice.tile_4x4x4x8(np.load(os.path.join(HERE, "golden", "conf_04040408.npy")), DIMS8)


def dev(U, dims):
    g = hmc.GaugeField(dims)
    g.from_host(np.ascontiguousarray(U))
    return g


def test_rect_action_and_force_match_oracle():
    for dims, U in ((DIMS8, conf8()), ((4, 4, 4, 6), lattice.random_gauge((4, 4, 4, 6), 0.6, 3))):
        a = hmc.Action_G_Rectangle(dims, BETA, CP, CR)
        a.set_config(dev(U, dims))
        H = a.calcH()
        Ho, plo = go.rect_action(U, dims, BETA, CP, CR)
        assert abs(H - Ho) / Ho < 1e-13 and abs(a.plaquette - plo) < 1e-14
        F = hmc.GaugeField(dims)
        a.force(F)
        Fo = go.rect_force(U, dims, BETA, CP, CR)
        assert np.abs(F.to_host() - Fo).max() / np.abs(Fo).max() < 1e-13


def test_rect_action_reference_value():
    a = hmc.Action_G_Rectangle(DIMS8, BETA, CP, CR)
    a.set_config(dev(conf8(), DIMS8))
    assert abs(a.calcH() - KNOWN["H_Grectangle_init"]["value"]) < 1e-7


def test_momenta_and_exp_update_match_oracle():
    r = hmc.RandomNumbers_parallel_Mseries(DIMS8, 1234567)
    r.reset(200000)
    g = r.gauss_lex_global(8, 4)
    P = hmc.GaugeField(DIMS8)
    Hk = np.zeros(1)
    dwf_amd._lib.check("dwf_hmc_set_iP", dwf_amd._lib.load().dwf_hmc_set_iP(
        hmc._dims(DIMS8), P.ptr, hmc._ptr(np.ascontiguousarray(g)), hmc._ptr(Hk)))
    Po, Hko = go.set_iP(g)
    assert np.abs(P.to_host() - Po).max() < 1e-15
    assert abs(Hk[0] - KNOWN["H_kin_init"]["value"]) < 1e-7
    U = dev(conf8(), DIMS8)
    upd = hmc.Integrator_UpdateU([], 8)
    upd.evolve(0.0123, P, U)
    Uo = go.exp_update(conf8(), 0.0123, Po, 8)
    assert np.abs(U.to_host() - Uo).max() < 1e-14


@pytest.fixture(scope="module")
def trajectory():
    hmcobj, rng = hmc.benchmark_nf2p1(DIMS8, RATIONAL)
    rng.reset(200000)
    U = dev(conf8(), DIMS8)
    print("case: reference 8^4 trajectory (HMC_DWF_Nf2p1_eo.yaml)", flush=True)
    out = hmcobj.update(U)
    print("  init", out["init"], "\n  trial", out["trial"], "\n  H_diff", out["H_diff"],
          "plaq", out["plaq_trial"], flush=True)
    return out


def test_trajectory_langevin_energies(trajectory):
    h = trajectory["init"]
    pairs = [("H_kin", "H_kin_init"), ("H_0_Frational", "H_Frational_init"),
             ("H_1_Fratio_Nf2", "H_Fratio_Nf2_init"), ("H_2_Fratio_Nf2prec", "H_Fratio_Nf2prec_init"),
             ("H_3_gauge", "H_Grectangle_init")]
    for mine, ref in pairs:
        want = KNOWN[ref]["value"]
        assert abs(h[mine] - want) / want < 1e-10, (mine, h[mine], want)
    assert abs(trajectory["plaq_init"] - KNOWN["plaq_init_8x8x8x8_sample_1gpu"]["value"]) < 1e-13


def test_trajectory_trial_energies_and_delta_H(trajectory):
    h = trajectory["trial"]
    pairs = [("H_kin", "H_kin_trial"), ("H_0_Frational", "H_Frational_trial"),
             ("H_1_Fratio_Nf2", "H_Fratio_Nf2_trial"), ("H_2_Fratio_Nf2prec", "H_Fratio_Nf2prec_trial"),
             ("H_3_gauge", "H_Grectangle_trial")]
    for mine, ref in pairs:
        assert abs(h[mine] - KNOWN[ref]["value"]) < 1e-6, (mine, h[mine], KNOWN[ref]["value"])
    assert abs(trajectory["plaq_trial"] - KNOWN["plaq_trial"]["value"]) < 1e-11
    assert abs(trajectory["H_diff"] - KNOWN["H_diff"]["value"]) < 1e-6


# Config-1 style setup: Shamir DWF (b = 1, c = 0), M0 = 1.6, Ls = 8, links
# exp(0.5 iH) seeded, 2 stout steps; masses (Nf1, Nf2, Nf2_prec, PV).
CFG1 = dict(Ls=8, M0=1.6, b=1.0, c=0.0, masses=(0.04, 0.02, 0.2, 1.0), nsmear=2)


def _setup(dims, Nsteps, tau, seed=11):
    h, rng = hmc.benchmark_nf2p1(dims, RATIONAL, Nsteps=Nsteps, trajectory_length=tau, **CFG1)
    U0 = lattice.random_gauge(dims, 0.5, seed)
    return h, rng, U0


def test_reversibility():
    """Forward for tau, flip the momenta, forward for tau: the links and the
    energy return (Omelyan is time-reversible)."""
    dims = (4, 4, 4, 8)
    h, rng, U0 = _setup(dims, (2, 1, 3), 0.2)
    rng.reset(5)
    U = dev(U0, dims)
    for a in h.actions:
        a.set_config(U)
    for d in h.directors:
        d.set_config(U)
    iP = hmc.GaugeField(dims)
    e0 = sum(h.langevin(iP, U).values())
    h.integrator.evolve(h.tau, iP, U)
    e1 = sum(h.calc_Hamiltonian(iP, U).values())
    iP.axpy(-2.0, iP)
    h.integrator.evolve(h.tau, iP, U)
    e2 = sum(h.calc_Hamiltonian(iP, U).values())
    dU = np.abs(U.to_host() - U0).max()
    print(f"  reversibility: max |U - U0| = {dU:.3e}, H0 = {e0:.10f}, H1 - H0 = {e1 - e0:.3e}, "
          f"H2 - H0 = {e2 - e0:.3e}", flush=True)
    assert dU < 1e-10
    assert abs(e2 - e0) < 1e-8 * abs(e0)


def test_delta_H_falls_as_step_squared():
    """|dH| with N and 2N steps on the outermost level (the fermion forces; the
    inner levels fixed at 4 steps; same momenta, noise and links): Omelyan is
    second order, so the ratio is about 4 once the steps are small enough
    (measured: 0.159, 0.0330, 0.00815 for N = 2, 4, 8 at tau = 0.5)."""
    dims = (4, 4, 4, 8)
    dH = []
    for n in (4, 8):
        h, rng, U0 = _setup(dims, (n, 4, 4), 0.5)
        rng.reset(9)
        out = h.update(dev(U0, dims))
        dH.append(abs(out["H_diff"]))
    ratio = dH[0] / dH[1]
    print(f"  dH(N) = {dH[0]:.4e}, dH(2N) = {dH[1]:.4e}, ratio {ratio:.2f}", flush=True)
    assert 3.0 < ratio < 5.5


def test_config1_trajectory_8x8x8x8():
    """One 2+1-flavour trajectory, tau = 1, steps [2, 1, 3], on 8^4 with Ls = 8
    Shamir and 2 stout steps on exp(0.5 iH) links: the energy is conserved to
    |dH| / H < 1e-4 (measured dH = -2.42 for H = 2.4e5: the outer step of 0.5
    is coarse for these rough links; the 1/N^2 test above shows the error is
    the integrator's)."""
    dims = (8, 8, 8, 8)
    h, rng, U0 = _setup(dims, (2, 1, 3), 1.0, seed=2024)
    rng.reset(1)
    out = h.update(dev(U0, dims))
    print(f"  config-1 trajectory: H0 = {out['H_total_init']:.8f}, dH = {out['H_diff']:.6e}, "
          f"plaq {out['plaq_init']:.10f} -> {out['plaq_trial']:.10f}", flush=True)
    assert np.isfinite(out["H_diff"])
    assert abs(out["H_diff"]) < 1e-4 * abs(out["H_total_init"])


def test_reference_trajectory_32x16x16x24():
    """The same input file on 32x16x16x24, run by the reference on 4 ranks
    (submit_4node.sh.o1137366:1515-1997; the parallel generator's streams are per
    global (z, t), so one GPU draws the same numbers): the Langevin energies to
    1e-10 relative, the trial energies to 1e-11 relative (about 1e-4 absolute on
    terms of 1.4e7, the precision the stopping rule 1e-28 leaves in |phi^+ M^-1 phi|)
    and H(diff) = 0.05636821 to 1e-6 absolute."""
    import time
    dims = (32, 16, 16, 24)
    sfx = "_32x16x16x24"
    hmcobj, rng = hmc.benchmark_nf2p1(dims, RATIONAL)
    rng.reset(200000)
    U0 = lattice.tile_4x4x4x8(np.load(os.path.join(HERE, "golden", "conf_04040408.npy")), dims)
    U = dev(U0, dims)
    print("case: reference 32x16x16x24 trajectory (HMC_DWF_Nf2p1_eo.yaml)", flush=True)
    t0 = time.perf_counter()
    out = hmcobj.update(U)
    dt = time.perf_counter() - t0
    print(f"  {dt:.2f} s (reference, 4 ranks: {KNOWN['trajectory_seconds' + sfx]['value']} s)"
          f"\n  init {out['init']}\n  trial {out['trial']}\n  H_diff {out['H_diff']:.10f}",
          flush=True)
    names = [("H_kin", "H_kin"), ("H_0_Frational", "H_Frational"),
             ("H_1_Fratio_Nf2", "H_Fratio_Nf2"), ("H_2_Fratio_Nf2prec", "H_Fratio_Nf2prec"),
             ("H_3_gauge", "H_Grectangle")]
    for mine, ref in names:
        want = KNOWN[ref + "_init" + sfx]["value"]
        assert abs(out["init"][mine] - want) / want < 1e-10, (mine, out["init"][mine], want)
        want = KNOWN[ref + "_trial" + sfx]["value"]
        assert abs(out["trial"][mine] - want) / want < 1e-11, (mine, out["trial"][mine], want)
    assert abs(out["plaq_init"] - KNOWN["plaq_init" + sfx]["value"]) < 1e-12
    assert abs(out["plaq_trial"] - KNOWN["plaq_trial" + sfx]["value"]) < 1e-11
    assert abs(out["H_diff"] - KNOWN["H_diff" + sfx]["value"]) < 1e-6


def test_reference_trajectory_fp16_inner():
    """The 8^4 reference trajectory with the Richardson solves' inner CG on
    FP16-storage operators (inner |r|^2/|b|^2 < 2e-4, more outer steps to the
    same FP64 stopping condition 1e-28): the energies and H(diff) are those of
    the reference to 1e-6, as with its FP32 inner CG."""
    hmcobj, rng = hmc.benchmark_nf2p1(DIMS8, RATIONAL, inner_precision="half", inner_stop=2e-4)
    rng.reset(200000)
    out = hmcobj.update(dev(conf8(), DIMS8))
    print(f"  FP16 inner: H_diff {out['H_diff']:.10f}", flush=True)
    pairs = [("H_kin", "H_kin_trial"), ("H_0_Frational", "H_Frational_trial"),
             ("H_1_Fratio_Nf2", "H_Fratio_Nf2_trial"), ("H_2_Fratio_Nf2prec", "H_Fratio_Nf2prec_trial"),
             ("H_3_gauge", "H_Grectangle_trial")]
    for mine, ref in pairs:
        assert abs(out["trial"][mine] - KNOWN[ref]["value"]) < 1e-6, (mine, out["trial"][mine])
    assert abs(out["H_diff"] - KNOWN["H_diff"]["value"]) < 1e-6


def test_config5_parameters_mixed_matches_fp64():
    """The 2+1-flavour trajectory with the light mass 0.01 and Ls = 16 (the
    24^3 x 48 workload's parameters, here on 16^3 x 32): the Richardson solves'
    inner CG in FP16 storage (2e-4) and in FP64 (1e-8) stop on the same FP64
    residual, so H(diff) agrees to 1e-6 (measured at 24^3 x 48: 0.19950747 and
    0.19950745)."""
    dims = (16, 16, 16, 32)
    U0 = lattice.tile_4x4x4x8(np.load(os.path.join(HERE, "golden", "conf_04040408.npy")), dims)
    res = {}
    for inner, stop in (("half", 2e-4), ("double", 1e-8)):
        h, rng = hmc.benchmark_nf2p1(dims, RATIONAL, Ls=16, masses=(0.04, 0.01, 0.1, 1.0),
                                     inner_precision=inner, inner_stop=stop)
        rng.reset(200000)
        res[inner] = h.update(dev(U0, dims))
        print(f"  {inner} inner: H_diff {res[inner]['H_diff']:.10f}", flush=True)
    assert abs(res["half"]["H_diff"] - res["double"]["H_diff"]) < 1e-6
    for k, v in res["double"]["trial"].items():
        assert abs(res["half"]["trial"][k] - v) < 1e-6 * max(1.0, abs(v) * 1e-8), k
