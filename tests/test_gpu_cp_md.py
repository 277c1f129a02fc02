"""Car-Parrinello steps in mixed precision against FP64 on the device (the
error contract of the mixed path, at a size where the split-bf16 contractions
and FP32 transforms are all active): the same state evaluated in both
precisions (|dE_KS| < 1e-6 Ha/atom, |dF_ion| < 1e-5 Ha/bohr), and the end
states of a few CP steps from the same start, both evaluated in FP64."""
import numpy as np
import pytest

from cpmd_amd import system as S
from cpmd_amd.ks import KSDevice

pytestmark = pytest.mark.gpu

ION_MASS = 12.0 * 1822.888486
DT, EMASS = 1.0, 400.0


def test_cp_steps_mixed_vs_fp64(gpu):
    n = 64
    sysm = S.make_system(24.0, 64, 70.0, nstate=n, with_density=True)
    assert 2 * sysm.gset.ngw >= 16384          # the split-bf16 products are used
    ks = KSDevice(sysm)
    ctx = ks.ctx
    f = sysm.f
    buf = {k: ctx.alloc_wfn(n) for k in ("c0", "cm", "cp", "c2")}
    # a partly optimised FP64 start (steepest descent, Loewdin)
    ctx.set_precision("fp64")
    ctx.upload(S.initial_coefficients(sysm), buf["c0"])
    ks.set_ions(sysm.tau)
    for _ in range(10):
        ks.energy(buf["c0"], f, gradient=buf["c2"])
        ctx.lagrange_dev(buf["c0"], buf["c2"], f, n)
        ctx.axpy_wfn(0.02, buf["c2"], buf["c0"], n)
        ctx.lowdin_dev(buf["c0"], n)
    c_start = ctx.download(buf["c0"], n)
    nat = sysm.nat

    def evaluate(c, tau, mode):
        ctx.set_precision(mode)
        ctx.upload(c, buf["c0"])
        return ks.energy(buf["c0"], f, tau=tau, gradient=buf["c2"], forces=True)

    # the same state in both precisions
    e64 = evaluate(c_start, sysm.tau, "fp64")
    emx = evaluate(c_start, sysm.tau, "mixed")
    assert emx["etot"] != e64["etot"]          # the mixed path is in use
    assert abs(emx["etot"] - e64["etot"]) / nat < 1e-6
    assert np.abs(emx["fion"] - e64["fion"]).max() < 1e-5

    def trajectory(mode, steps=5):
        ctx.set_precision(mode)
        ctx.upload(c_start, buf["c0"])
        ctx.axpby_wfn(0.0, buf["c0"], 0.0, buf["cm"], n)
        tau = sysm.tau.copy()
        vion = np.zeros_like(tau)
        for _ in range(steps):
            r = ks.energy(buf["c0"], f, tau=tau, gradient=buf["c2"], forces=True)
            vion = vion + DT / ION_MASS * r["fion"].T
            tau = tau + DT * vion
            ctx.posupa_dev(buf["c0"], buf["c2"], buf["cm"], buf["cp"], n, DT / EMASS, DT, epsog=1e-8, maxit=60)
            buf["c0"], buf["cp"] = buf["cp"], buf["c0"]
        return ctx.download(buf["c0"], n), tau

    c64, tau64 = trajectory("fp64")
    cmx, taumx = trajectory("mixed")
    end64 = evaluate(c64, tau64, "fp64")
    endmx = evaluate(cmx, taumx, "fp64")
    # two trajectories drift apart; 5 steps keep the gap within the energy
    # bound and within 5x the single-evaluation force bound
    assert abs(endmx["etot"] - end64["etot"]) / nat < 1e-6
    assert np.abs(endmx["fion"] - end64["fion"]).max() < 5e-5
    # the mixed trajectory stays orthonormal to FP32 level
    ctx.set_precision("fp64")
    ctx.upload(cmx, buf["c0"])
    ov = ctx.ovlap_dev(buf["c0"], buf["c0"], n)
    assert np.abs(ov - np.eye(n)).max() < 1e-6


def test_cp_energy_conservation_from_ground_state(gpu):
    """CP dynamics as production runs it: ODIIS ground state (mixed then
    FP64), then 10 CP steps at dt 4 a.u., emass 400 a.u. in each precision.
    CPMD's constant of motion E_ham = E_KS + ionic kinetic energy + emass
    sum_i dotp(cm_i, cm_i) (md_driver, rekine) with full-step velocities
    stays within 1e-5 Ha/atom, and the two precisions end within 1e-6 Ha/atom."""
    sysm = S.make_system(12.0, 8, 70.0, nstate=16, seed=5, with_density=True)
    ks = KSDevice(sysm)
    ctx = ks.ctx
    f = sysm.f
    n = len(f)
    nat = sysm.nat
    dt = 4.0
    buf = {k: ctx.alloc_wfn(n) for k in ("c0", "cm", "cp", "c2")}
    ks.set_ions(sysm.tau)
    ctx.upload(S.initial_coefficients(sysm, seed=5), buf["c0"])
    r = ks.optimize(buf["c0"], f, buf["c2"], tol=1e-6, method="diis", maxiter=400)
    assert r["converged"]
    ctx.set_precision("fp64")
    c_gs = ctx.download(buf["c0"], n)

    def e_ham(tau, vion):
        ctx.set_precision("fp64")
        e = ks.energy(buf["c0"], f, tau=tau, gradient=buf["c2"], forces=True)
        ctx.lagrange_dev(buf["c0"], buf["c2"], f, n)
        ctx.axpy_wfn(0.5 * dt / EMASS, buf["c2"], buf["cm"], n)
        ekinc = EMASS * float(ctx.dotp_wfn(buf["cm"], buf["cm"], n).sum())
        v = vion + 0.5 * dt / ION_MASS * e["fion"].T
        return e["etot"] + 0.5 * ION_MASS * float(np.sum(v * v)) + ekinc

    out = {}
    for mode in ("fp64", "mixed"):
        ctx.set_precision("fp64")
        ctx.upload(c_gs, buf["c0"])
        ctx.axpby_wfn(0.0, buf["c0"], 0.0, buf["cm"], n)
        tau = sysm.tau.copy()
        vion = np.zeros_like(tau)
        e0 = ks.energy(buf["c0"], f, tau=tau)["etot"]
        ctx.set_precision(mode)
        for _ in range(10):
            rr = ks.energy(buf["c0"], f, tau=tau, gradient=buf["c2"], forces=True)
            vion = vion + dt / ION_MASS * rr["fion"].T
            tau = tau + dt * vion
            ctx.posupa_dev(buf["c0"], buf["c2"], buf["cm"], buf["cp"], n, dt / EMASS, dt, epsog=1e-8, maxit=60)
            buf["c0"], buf["cp"] = buf["cp"], buf["c0"]
        out[mode] = (e0, e_ham(tau, vion), float(np.abs(tau - sysm.tau).max()))
    for mode, (e0, e1, moved) in out.items():
        print(f"{mode}: E_ham drift {(e1 - e0) / nat:.3e} Ha/atom, ions moved up to {moved:.2e} bohr", flush=True)
        assert moved > 1e-4                          # the ions do move
        assert abs(e1 - e0) / nat < 1e-5
    assert abs(out["mixed"][1] - out["fp64"][1]) / nat < 1e-6
