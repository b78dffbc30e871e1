"""HMC on the GPU: the host-side driver mirroring the reference's HMC classes.

The classes follow src/lib/HMC (HMC_General, Integrator_Omelyan,
Integrator_UpdateP / UpdateU, Langevin_Momentum), src/lib/Action/Gauge
(Action_G_Rectangle) and src/lib_alt/Action/Fermion (Action_F_alt_Ratio_eo,
Action_F_alt_Rational_Ratio) with the same names, parameters and call order,
so that the random numbers are drawn in the reference's order and the
energies agree with the reference's.  Every field operation runs in
libdwfhip: the links, momenta, forces and pseudofermions stay in device
memory; the random numbers are generated on the host (as in the reference)
and uploaded.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from . import (AFopr_Domainwall_5din_eo, AForce_F_Domainwall_5din_eo, AForce_F_Rational_Ratio,
               AForce_F_Smeared_alt, AFopr_Rational, ASolver_Richardson, DeviceArray,
               Director_alt_Smear, Rational)


def _ptr(a):
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _dims(d):
    return (C.c_int * 4)(*[int(v) for v in d])


class RandomNumbers_parallel_Mseries:
    """RandomNumbers_parallel<RandomNumbers_Mseries> on one process: an
    M-series master generator and one per (z, t) slice (host)."""

    def __init__(self, lattice_size, seed):
        self.lattice_size = tuple(int(v) for v in lattice_size)
        h = C.c_void_p()
        _lib.check("dwf_rng_create",
                   _lib.load().dwf_rng_create(_dims(self.lattice_size), C.c_ulong(int(seed)),
                                              C.byref(h)))
        self._h = h.value

    def reset(self, seed):
        _lib.check("dwf_rng_reset", _lib.load().dwf_rng_reset(self._h, C.c_ulong(int(seed))))

    def get(self):
        v = np.zeros(1)
        _lib.check("dwf_rng_get", _lib.load().dwf_rng_get(self._h, _ptr(v)))
        return float(v[0])

    def gauss_lex_global(self, nin, nex):
        """Field(nin, V, nex) of Gaussian numbers (variance 1/2), as (nex, V, nin)."""
        V = int(np.prod(self.lattice_size))
        out = np.empty((nex, V, nin))
        _lib.check("dwf_rng_gauss_lex_global",
                   _lib.load().dwf_rng_gauss_lex_global(self._h, int(nin), int(nex), _ptr(out)))
        return out

    def gauss_eo_global(self, nin, nex, parity=0):
        """gauss_lex_global followed by Index_eo::convertField(., ., parity): the
        numbers of the sites of one parity, (nex, V/2, nin) in even-odd order."""
        Vh = int(np.prod(self.lattice_size)) // 2
        out = np.empty((nex, Vh, nin))
        _lib.check("dwf_rng_gauss_eo_global", _lib.load().dwf_rng_gauss_eo_global(
            self._h, int(nin), int(nex), int(parity), _ptr(out)))
        return out

    def __del__(self):
        if getattr(self, "_h", None):
            try:
                _lib.load().dwf_rng_destroy(self._h)
            except Exception:
                pass
            self._h = None


def even_part(xi_lex, lattice_size):
    """Index_eo::convertField(xi, xi_lex, 0): the even sites of a lexical
    Field(nin, V, nex) (nex, V, nin) array, as (nex, V/2, nin) in even-odd order
    (src/lib/Field/index_eo.cpp:125-156)."""
    Nx, Ny, Nz, Nt = lattice_size
    x2 = np.arange(Nx // 2)
    y = np.arange(Ny)
    z = np.arange(Nz)
    t = np.arange(Nt)
    T, Z, Y, X2 = np.meshgrid(t, z, y, x2, indexing="ij")
    x = 2 * X2 + (Y + Z + T) % 2
    lex = (x + Nx * (Y + Ny * (Z + Nz * T))).reshape(-1)
    return np.ascontiguousarray(xi_lex[:, lex, :])


# ---------------------------------------------------------------------------
# gauge field in device memory
# ---------------------------------------------------------------------------

class GaugeField(DeviceArray):
    """A device Field_G (4, V, 18) float64 array."""

    def __init__(self, lattice_size):
        self.lattice_size = tuple(int(v) for v in lattice_size)
        super().__init__((4, int(np.prod(self.lattice_size)), 18))

    @property
    def n(self):
        return int(np.prod(self.shape))

    def copy_from(self, other):
        _lib.check("dwf_device_copy", _lib.load().dwf_device_copy(self.ptr, other.ptr, self.n))

    def axpy(self, a, x):
        _lib.check("dwf_device_axpy", _lib.load().dwf_device_axpy(self.ptr, float(a), x.ptr, self.n))

    def norm2(self):
        v = np.zeros(1)
        _lib.check("dwf_device_norm2", _lib.load().dwf_device_norm2(self.ptr, self.n, _ptr(v)))
        return float(v[0])

    def set_zero(self):
        self.from_host(np.zeros(self.shape))

    def plaquette(self):
        v = np.zeros(1)
        _lib.check("dwf_plaquette_device",
                   _lib.load().dwf_plaquette_device(_dims(self.lattice_size), self.ptr, _ptr(v)))
        return float(v[0])


# ---------------------------------------------------------------------------
# smearing director and smeared operators
# ---------------------------------------------------------------------------

class Director:
    """Director_alt_Smear (src/lib_alt/Smear/director_alt_Smear-tmpl.h): the
    smeared levels of the current thin links, recomputed on first use after
    notify_linkv()."""

    def __init__(self, lattice_size, rho, nsmear):
        self.sm = Director_alt_Smear(lattice_size, rho, nsmear)
        self.U = None
        self.valid = False
        self.version = 0

    def set_config(self, U):
        self.U = U
        self.valid = False

    def notify_linkv(self):
        self.valid = False

    def ensure(self):
        if not self.valid:
            _lib.check("dwf_smear_set_config_device",
                       _lib.load().dwf_smear_set_config_device(self.sm._h, self.U.ptr))
            self.valid = True
            self.version += 1


class SmearedFopr:
    """AFopr_Smeared_alt: a domain-wall operator on the top smearing level."""

    def __init__(self, director, lattice_size, Ls, mq, M0, b, c, bc=(1, 1, 1, -1),
                 precision="double"):
        self.director = director
        self.op = AFopr_Domainwall_5din_eo(lattice_size, Ls, mq, M0, b, c, bc, precision)
        self._version = -1

    def refresh(self):
        self.director.ensure()
        if self._version != self.director.version:
            self.op.set_config_smeared(self.director.sm, None)
            self._version = self.director.version
        return self.op


class Fprop_alt_Standard_eo_Richardson:
    """Fprop_alt_Standard_eo_Richardson in mode DdagD_even
    (src/lib_alt/Measurements/Fermion/fprop_alt_Standard_eo_Richardson-tmpl.h:27-102,
    460-481): FP64 Richardson to convergence_criterion_squared around an FP32
    CG to 1e-8 with max_iter * max_restart iterations (the reference's scheme;
    fopr_f may be an FP16-storage operator with a looser inner_stop)."""

    def __init__(self, fopr_d, fopr_f, max_iter=100, max_restart=40, stop_cond=1e-28,
                 inner_stop=1e-8):
        self.fopr_d, self.fopr_f = fopr_d, fopr_f
        self.max_iter, self.inner_max = int(max_iter), int(max_iter * max_restart)
        self.stop = float(stop_cond)
        self.inner_stop = float(inner_stop)
        self._solver = None
        self.nconv = self.diff = None

    def invert(self, x, b):
        d, f = self.fopr_d.refresh(), self.fopr_f.refresh()
        d.set_mode("DdagD")
        f.set_mode("DdagD")
        if self._solver is None:
            self._solver = ASolver_Richardson(d, f, self.max_iter, self.stop, self.inner_max,
                                              self.inner_stop)
        self.nconv, self.diff = self._solver.solve(x, b)
        return self.nconv, self.diff


# ---------------------------------------------------------------------------
# actions
# ---------------------------------------------------------------------------

class Action_G_Rectangle:
    """Plaquette + rectangle gauge action (src/lib/Action/Gauge/action_G_Rectangle.cpp,
    src/lib/Force/Gauge/force_G_Rectangle.cpp)."""

    def __init__(self, lattice_size, beta, c_plaq, c_rect):
        self.lattice_size = tuple(int(v) for v in lattice_size)
        self.beta, self.c_plaq, self.c_rect = float(beta), float(c_plaq), float(c_rect)
        self.U = None
        self.label = "gauge"

    def set_config(self, U):
        self.U = U

    def langevin(self, rng):
        return self.calcH()

    def calcH(self):
        H, plaq = np.zeros(1), np.zeros(1)
        _lib.check("dwf_hmc_rect_action", _lib.load().dwf_hmc_rect_action(
            _dims(self.lattice_size), self.U.ptr, self.beta, self.c_plaq, self.c_rect,
            _ptr(H), _ptr(plaq)))
        self.plaquette = float(plaq[0])
        return float(H[0])

    def force(self, F):
        _lib.check("dwf_hmc_rect_force", _lib.load().dwf_hmc_rect_force(
            _dims(self.lattice_size), F.ptr, self.U.ptr, self.beta, self.c_plaq, self.c_rect))


def _noise_field(rng, fopr, Ls):
    """Gaussian pseudofermion noise: gauss_lex_global on Field(24, V, Ls), even
    sites, as a device field of fopr; returns (field, |xi|^2)."""
    xi = rng.gauss_eo_global(24, Ls, 0)
    f = fopr.new_field()
    fopr.convert(f, xi)
    return f, float(np.sum(xi * xi))


class Action_F_alt_Ratio_eo:
    """Ratio of two domain-wall determinants (Hasenbusch), even-odd:
    S = phi^+ H1 (D2^+ D2)^-1 H1^+ phi
    (src/lib_alt/Action/Fermion/action_F_alt_Ratio_eo-tmpl.h:66-315; force of
    src/lib_alt/Force/Fermion/aforce_F_Ratio_eo-tmpl.h:108-157 under
    AForce_F_Smeared_alt)."""

    def __init__(self, fopr1, fopr2, fprop1_H, fprop2_MD, fprop2_H, label=""):
        self.fopr1, self.fopr2 = fopr1, fopr2
        self.fprop1_H, self.fprop2_MD, self.fprop2_H = fprop1_H, fprop2_MD, fprop2_H
        self.label = label
        self.psf = None
        self.force1 = AForce_F_Domainwall_5din_eo(fopr1.op)
        self.force2 = AForce_F_Domainwall_5din_eo(fopr2.op)
        self.U = None

    def set_config(self, U):
        self.U = U

    def langevin(self, rng):
        o1, o2 = self.fopr1.refresh(), self.fopr2.refresh()
        xi, h = _noise_field(rng, o1, o1.Ls)
        v1, v2 = o1.new_field(), o1.new_field()
        o2.set_mode("H")
        o2.mult_dag(v2, xi)
        self.fprop1_H.invert(v1, v2)
        self.psf = o1.new_field()
        o1.set_mode("H")
        o1.mult(self.psf, v1)
        self.H_langevin = h
        self.calcH()
        return h

    def calcH(self):
        o1 = self.fopr1.refresh()
        v1, v2 = o1.new_field(), o1.new_field()
        o1.set_mode("H")
        o1.mult_dag(v1, self.psf)
        self.fprop2_H.invert(v2, v1)
        return v1.dot(v2)

    def force(self, F):
        o1 = self.fopr1.refresh()
        self.fopr2.refresh()
        v1, v2 = o1.new_field(), o1.new_field()
        o1.set_mode("H")
        o1.mult_dag(v1, self.psf)
        self.fprop2_MD.invert(v2, v1)
        # force_core1(zeta = v2, eta = psf) of the smeared ratio force
        self._udiv1(F, v2, self.psf)
        sm = self.fopr1.director.sm
        sm.mult_jacobian(F)
        sm.mult_generator(F)

    def _udiv1(self, F, zeta, eta):
        tmp = GaugeField(self.fopr1.op.lattice_size)
        self.force2.force_udiv(F, zeta)
        self.force1.set_mode("Hdag")
        self.force1.force_udiv1(tmp, zeta, eta)
        F.axpy(-1.0, tmp)
        self.force1.set_mode("H")
        self.force1.force_udiv1(tmp, eta, zeta)
        F.axpy(-1.0, tmp)


class Action_F_alt_Rational_Ratio:
    """One-flavour rational ratio S = phi^+ r1(A^+A) r2(B^+B) r1(A^+A) phi,
    A = fopr1 (Pauli-Villars), B = fopr2 (src/lib_alt/Action/Fermion/
    action_F_alt_Rational_Ratio-tmpl.h:37-260): the pseudofermion from
    (A^+A)^(-1/4) (B^+B)^(1/4) xi, the energy <r1_H(A^+A) phi, r2_H(B^+B) r1_H(A^+A) phi>,
    the force of AForce_F_Rational_Ratio with the MD rationals under
    AForce_F_Smeared_alt.  rationals: dict with Rational objects under
    "num_langevin", "den_langevin", "num_H", "den_H", "num_MD", "den_MD"."""

    def __init__(self, fopr1, fopr2, rationals, label=""):
        self.fopr1, self.fopr2 = fopr1, fopr2
        self.r = rationals
        self.label = label
        self.force_MD = AForce_F_Rational_Ratio(fopr1.op, fopr2.op, rationals["num_MD"],
                                                rationals["den_MD"])
        self.psf = None
        self.U = None
        self.nconv = []

    def set_config(self, U):
        self.U = U

    def langevin(self, rng):
        o1, o2 = self.fopr1.refresh(), self.fopr2.refresh()
        xi, h = _noise_field(rng, o1, o1.Ls)
        v1 = o1.new_field()
        self.psf = o1.new_field()
        r2 = AFopr_Rational(o2, self.r["den_langevin"])
        r2.mult(v1, xi)                                  # (B^+B)^(1/4) xi
        r1 = AFopr_Rational(o1, self.r["num_langevin"])
        r1.mult(self.psf, v1)                            # (A^+A)^(-1/4)
        self.nconv = [r2.nconv, r1.nconv]
        return h

    def calcH(self):
        o1, o2 = self.fopr1.refresh(), self.fopr2.refresh()
        v1, v2 = o1.new_field(), o1.new_field()
        AFopr_Rational(o1, self.r["num_H"]).mult(v1, self.psf)   # (A^+A)^(1/4)
        AFopr_Rational(o2, self.r["den_H"]).mult(v2, v1)         # (B^+B)^(-1/2)
        return v1.dot(v2)

    def force(self, F):
        self.fopr1.refresh()
        self.fopr2.refresh()
        self.force_MD.force_udiv(F, self.psf)
        sm = self.fopr1.director.sm
        sm.mult_jacobian(F)
        sm.mult_generator(F)


# ---------------------------------------------------------------------------
# integrators and the HMC update
# ---------------------------------------------------------------------------

class Integrator_UpdateP:
    """iP += step * (sum of the forces of this level's actions), the force
    cached until the links change (src/lib/HMC/integrator_UpdateP.cpp:31-47)."""

    def __init__(self, actions, lattice_size):
        self.actions = list(actions)
        self.force = GaugeField(lattice_size)
        self.tmp = GaugeField(lattice_size)
        self.valid = False
        self.evaluations = 0

    def invalidate_cache(self):
        self.valid = False

    def evolve(self, step, iP, U):
        if not self.valid:
            self.force.set_zero()
            for a in self.actions:
                a.force(self.tmp)
                self.force.axpy(1.0, self.tmp)
            self.valid = True
            self.evaluations += 1
        iP.axpy(step, self.force)


class Integrator_UpdateU:
    """U <- exp(step iP) U with a Taylor expansion of order Nprec, then every
    force cache and smearing director is told the links changed
    (src/lib/HMC/integrator_UpdateU.cpp)."""

    def __init__(self, directors, Nprec=8):
        self.directors = list(directors)
        self.Nprec = int(Nprec)
        self.notify = []

    def append_notify(self, client):
        self.notify.append(client)

    def invalidate_cache(self):
        for c in self.notify:
            c.invalidate_cache()

    def evolve(self, step, iP, U):
        _lib.check("dwf_hmc_exp_update", _lib.load().dwf_hmc_exp_update(
            _dims(U.lattice_size), U.ptr, float(step), iP.ptr, self.Nprec))
        self.invalidate_cache()
        for d in self.directors:
            d.notify_linkv()


class Integrator_Omelyan:
    """Second-order minimum-norm (Omelyan) integrator of one level
    (src/lib/HMC/integrator_Omelyan.cpp:112-153): lambda P, then per step
    U/2, (1 - 2 lambda) P, U/2, and 2 lambda P between steps."""

    def __init__(self, update_p, update_U, level, Nstep, lambda_omelyan=0.1931833275):
        self.update_p, self.update_U = update_p, update_U
        self.level, self.Nstep, self.lam = int(level), int(Nstep), float(lambda_omelyan)
        self.on_step = None   # optional callback(level, step) after each step

    def invalidate_cache(self):
        self.update_p.invalidate_cache()
        self.update_U.invalidate_cache()

    def evolve(self, step_size, iP, U):
        if self.Nstep <= 0:
            return
        e = step_size / self.Nstep
        self.update_p.evolve(e * self.lam, iP, U)
        for istep in range(1, self.Nstep + 1):
            self.update_U.evolve(e * 0.5, iP, U)
            self.update_p.evolve(e * (1.0 - 2.0 * self.lam), iP, U)
            self.update_U.evolve(e * 0.5, iP, U)
            if istep < self.Nstep:
                self.update_p.evolve(e * self.lam * 2.0, iP, U)
            if self.on_step:
                self.on_step(self.level, istep)
        self.update_p.evolve(e * self.lam, iP, U)


def build_omelyan(action_levels, Nsteps, directors, lattice_size, Nprec=8,
                  lambda_omelyan=0.1931833275):
    """Builder_Integrator::build_omelyan (src/lib/HMC/builder_Integrator.cpp:226-257):
    nested levels, the lowest one updating the links."""
    update_U = Integrator_UpdateU(directors, Nprec)
    integ = update_U
    for lv in range(len(Nsteps) - 1, -1, -1):
        up = Integrator_UpdateP(action_levels[lv], lattice_size)
        update_U.append_notify(up)
        integ = Integrator_Omelyan(up, integ, lv, Nsteps[lv], lambda_omelyan)
    return integ


class HMC_General:
    """One trajectory (src/lib/HMC/hmc_General.cpp:196-345): Langevin step
    (momenta, then every action's pseudofermion in order), molecular dynamics,
    the trial Hamiltonian and, if asked, the Metropolis test."""

    def __init__(self, actions, directors, integrator, rng, trajectory_length,
                 metropolis_test=False):
        self.actions = list(actions)
        self.directors = list(directors)
        self.integrator = integrator
        self.rng = rng
        self.tau = float(trajectory_length)
        self.metropolis = bool(metropolis_test)

    def langevin(self, iP, U):
        self.integrator.invalidate_cache()
        for d in self.directors:
            d.notify_linkv()
        g = self.rng.gauss_lex_global(8, 4)
        Hk = np.zeros(1)
        _lib.check("dwf_hmc_set_iP", _lib.load().dwf_hmc_set_iP(
            _dims(U.lattice_size), iP.ptr, _ptr(np.ascontiguousarray(g)), _ptr(Hk)))
        parts = {"H_kin": float(Hk[0])}
        for i, a in enumerate(self.actions):
            parts["H_%d_%s" % (i, a.label)] = a.langevin(self.rng)
        return parts

    def calc_Hamiltonian(self, iP, U):
        parts = {"H_kin": 0.5 * iP.norm2()}
        for i, a in enumerate(self.actions):
            parts["H_%d_%s" % (i, a.label)] = a.calcH()
        return parts

    def update(self, U):
        """Runs one trajectory from the device links U (updated in place when
        accepted); returns the energies and plaquettes."""
        Ucur = GaugeField(U.lattice_size)
        Ucur.copy_from(U)
        for a in self.actions:
            a.set_config(Ucur)
        for d in self.directors:
            d.set_config(Ucur)
        iP = GaugeField(U.lattice_size)
        h0 = self.langevin(iP, Ucur)
        plaq0 = Ucur.plaquette()
        self.integrator.evolve(self.tau, iP, Ucur)
        h1 = self.calc_Hamiltonian(iP, Ucur)
        plaq1 = Ucur.plaquette()
        H0, H1 = sum(h0.values()), sum(h1.values())
        dH = H1 - H0
        rand = self.rng.get() if self.metropolis else -1.0
        accepted = rand <= np.exp(-dH)
        if accepted:
            U.copy_from(Ucur)
        return {"init": h0, "trial": h1, "H_total_init": H0, "H_total_trial": H1, "H_diff": dH,
                "plaq_init": plaq0, "plaq_trial": plaq1, "accepted": bool(accepted),
                "U_trial": Ucur}


def rational(rational_dir, n_exp, d_exp, npoles, lower, upper, max_iter, stop_cond):
    """A Fopr_Rational parameter block: the Math_Rational coefficient file of
    x^(n/d) on [lower^2, upper^2] with npoles poles, and the shift solver's
    limits (src/lib/Tools/math_Rational.cpp:129-199)."""
    import os
    from .lattice import rational_filename
    name = rational_filename(n_exp, d_exp, npoles, lower * lower, upper * upper)
    return Rational.from_file(os.path.join(rational_dir, name),
                              maximum_number_of_iteration=max_iter,
                              convergence_criterion_squared=stop_cond)


def benchmark_nf2p1(lattice_size, rational_dir, seed=1234567, Nsteps=(2, 1, 3),
                    trajectory_length=0.1, Ls=12, M0=1.0, b=1.5, c=0.5,
                    masses=(0.02, 0.0005, 0.002, 1.0), beta=4.17,
                    c_plaq=1.66666666666666666666, c_rect=-0.08333333333333333333,
                    rho=0.1, nsmear=3, npoles_MD=16, npoles_H=24, lower=1.0e-2,
                    upper=3.16227766, stop_MD=1e-26, stop_H=1e-28, stop_solver=1e-28,
                    inner_precision="float", inner_stop=1e-8):
    """The 2+1-flavour domain-wall HMC of the reference benchmark
    (benchmark_hmc_dwf/hmc_dwf_eo.cpp:180-360 with HMC_DWF_Nf2p1_eo.yaml): a
    one-flavour rational ratio (Pauli-Villars / Nf1) and two Hasenbusch ratios
    (Nf2_prec / Nf2 and PV / Nf2_prec) on 3-level stout links, the
    plaquette + rectangle gauge action, Omelyan on 3 levels.  masses: (Nf1,
    Nf2, Nf2_prec, Pauli-Villars).  inner_precision / inner_stop: the inner CG of
    the Richardson solves ("float" and 1e-8 as in the reference; "half" storage
    with a looser criterion runs more outer steps, each cheaper, to the same
    FP64 stopping condition).  Returns (hmc, rng)."""
    dims = tuple(int(v) for v in lattice_size)
    director = Director(dims, rho, nsmear)
    m_nf1, m_nf2, m_prec, m_pv = masses

    def fopr(mq, precision="double"):
        return SmearedFopr(director, dims, Ls, mq, M0, b, c, (1, 1, 1, -1), precision)

    pv, nf1, nf2, prec = fopr(m_pv), fopr(m_nf1), fopr(m_nf2), fopr(m_prec)
    pv_f, nf2_f, prec_f = (fopr(m_pv, inner_precision), fopr(m_nf2, inner_precision),
                           fopr(m_prec, inner_precision))

    def fprop(d, f):
        return Fprop_alt_Standard_eo_Richardson(d, f, 100, 40, stop_solver, inner_stop)

    r = {
        "num_MD": rational(rational_dir, 1, 4, npoles_MD, lower, upper, 1000, stop_MD),
        "den_MD": rational(rational_dir, -1, 2, npoles_MD, lower, upper, 1000, stop_MD),
        "num_H": rational(rational_dir, 1, 4, npoles_H, lower, upper, 1000, stop_H),
        "den_H": rational(rational_dir, -1, 2, npoles_H, lower, upper, 1000, stop_H),
        "num_langevin": rational(rational_dir, -1, 4, npoles_H, lower, upper, 1000, stop_H),
        "den_langevin": rational(rational_dir, 1, 4, npoles_H, lower, upper, 1000, stop_H),
    }
    a_nf1 = Action_F_alt_Rational_Ratio(pv, nf1, r, label="Frational")
    a_nf2 = Action_F_alt_Ratio_eo(prec, nf2, fprop(prec, prec_f), fprop(nf2, nf2_f),
                                  fprop(nf2, nf2_f), label="Fratio_Nf2")
    a_prec = Action_F_alt_Ratio_eo(pv, prec, fprop(pv, pv_f), fprop(prec, prec_f),
                                   fprop(prec, prec_f), label="Fratio_Nf2prec")
    a_g = Action_G_Rectangle(dims, beta, c_plaq, c_rect)
    levels = [[a_nf1, a_nf2], [a_prec], [a_g]]
    integ = build_omelyan(levels, list(Nsteps), [director], dims)
    rng = RandomNumbers_parallel_Mseries(dims, seed)
    hmc = HMC_General([a_nf1, a_nf2, a_prec, a_g], [director], integ, rng, trajectory_length)
    return hmc, rng
