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
