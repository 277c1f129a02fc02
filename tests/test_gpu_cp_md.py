"""Car-Parrinello steps on the device.

* Mixed precision against FP64 (the error contract of the mixed path, at a
  size where the split-bf16 contractions and FP32 transforms are all active):
  the same state evaluated in both precisions (|dE_KS| < 1e-6 Ha/atom,
  |dF_ion| < 1e-5 Ha/bohr), and the end states of a few CP steps from the
  same start, both evaluated in FP64.
* rortv (the velocity constraint) against the oracle, and CPMD's velocity
  Verlet step against the leapfrog it is equivalent to.
* Energy conservation of CPMD's step from the ground state."""
import numpy as np
import pytest

import oracle
from cpmd_amd import ops
from cpmd_amd import system as S
from cpmd_amd.ks import KSDevice
from util import small_system

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


@pytest.fixture(scope="module")
def ground16(gpu):
    """12 bohr cell, 8 atoms, 16 states at 70 Ry: the ODIIS ground state
    (mixed then FP64, gemax < 1e-6)."""
    sysm = S.make_system(12.0, 8, 70.0, nstate=16, seed=5, with_density=True)
    ks = KSDevice(sysm)
    ctx = ks.ctx
    n = len(sysm.f)
    bufs = {k: ctx.alloc_wfn(n) for k in ("c0", "cm", "cp", "c2")}
    ks.set_ions(sysm.tau)
    ctx.upload(S.initial_coefficients(sysm, seed=5), bufs["c0"])
    r = ks.optimize(bufs["c0"], sysm.f, bufs["c2"], tol=1e-6, method="diis", maxiter=400)
    assert r["converged"]
    ctx.set_precision("fp64")
    return sysm, ks, bufs, ctx.download(bufs["c0"], n)


def test_rortv_parity_and_constraint(gpu):
    """mc_rortv against the oracle (FP64, 1e-12) and the constraint
    <c0|cm> + <cm|c0> = 0 after it, in FP64 and in mixed precision."""
    cell, gs, c0 = small_system("cubic", 6, seed=4)
    rng = np.random.default_rng(3)
    cm = np.asfortranarray((rng.standard_normal(c0.shape) + 1j * rng.standard_normal(c0.shape)) * 1e-2)
    cm[0] = cm[0].real
    ctx = ops.Context(cell, gs)
    out, gam = ctx.rortv(c0, cm, want_gam=True)
    ref, gref = oracle.rortv(c0, cm)
    assert np.abs(out - ref).max() < 1e-12 * np.abs(ref).max()
    assert np.abs(gam - gref).max() < 1e-12 * np.abs(gref).max()
    a = S.dotp_gamma(c0, out)
    assert np.abs(a + a.T).max() < 1e-13
    # mixed precision at a size where the split-bf16 products run
    sysm = S.make_system(24.0, 64, 70.0, nstate=64)
    assert 2 * sysm.gset.ngw >= 16384
    ctx2 = ops.Context(sysm.cell, sysm.gset)
    c = S.initial_coefficients(sysm)
    v = np.asfortranarray((rng.standard_normal(c.shape) + 1j * rng.standard_normal(c.shape)) * 1e-3)
    v[0] = v[0].real
    d0, dm = ctx2.upload(c), ctx2.upload(v)
    ctx2.set_precision("mixed")
    ekinc, _ = ctx2.cp_velupa_rortv_dev(d0, None, dm, 64, 0.0, 400.0)
    ctx2.set_precision("fp64")
    a = ctx2.ovlap_dev(d0, dm, 64)
    scale = np.abs(ctx2.ovlap_dev(dm, dm, 64)).max() ** 0.5
    assert np.abs(a + a.T).max() < 1e-6 * scale
    assert ekinc == pytest.approx(400.0 * ctx2.dotp_wfn(dm, dm, 64).sum(), rel=1e-12)


def test_velocity_verlet_matches_leapfrog(ground16):
    """CPMD's velocity Verlet step (velupa + posupa/rortog, forces, velupa +
    rortv) and the leapfrog with a half first kick give the same positions:
    rortv adds c0 gamy with gamy symmetric to the velocity, which the next
    rortog absorbs.  Four steps from the ground state at dt 4 a.u., both in
    FP64, agree to the rortog tolerance; after each VV step the velocity
    constraint holds and ekinc = emass sum_i dotp(cm_i, cm_i)."""
    sysm, ks, bufs, c_gs = ground16
    ctx = ks.ctx
    f = sysm.f
    n = len(f)
    dt = 4.0
    mass = ION_MASS
    ctx.set_precision("fp64")

    ctx.upload(c_gs, bufs["c0"])
    ctx.axpby_wfn(0.0, bufs["c0"], 0.0, bufs["cm"], n)
    st = ks.cp_init(bufs["c0"], bufs["cm"], bufs["c2"], f, sysm.tau, emass=EMASS, ion_mass=mass)
    for _ in range(4):
        ks.cp_step(st, bufs, f, dt, epsog=1e-12, maxit=100)
        a = ctx.ovlap_dev(bufs["c0"], bufs["cm"], n)
        assert np.abs(a + a.T).max() < 1e-12
        assert st["ekinc"] == pytest.approx(EMASS * ctx.dotp_wfn(bufs["cm"], bufs["cm"], n).sum(), rel=1e-13)
    c_vv, tau_vv = ctx.download(bufs["c0"], n), st["tau"].copy()

    ctx.upload(c_gs, bufs["c0"])
    ctx.axpby_wfn(0.0, bufs["c0"], 0.0, bufs["cm"], n)
    tau = sysm.tau.copy()
    vion = np.zeros_like(tau)
    for k in range(4):
        r = ks.energy(bufs["c0"], f, tau=tau, gradient=bufs["c2"], forces=True)
        kick = 0.5 if k == 0 else 1.0
        vion = vion + kick * dt / mass * r["fion"].T
        tau = tau + dt * vion
        ks.set_ions(tau)
        ctx.posupa_dev(bufs["c0"], bufs["c2"], bufs["cm"], bufs["cp"], n, kick * dt / EMASS, dt, epsog=1e-12,
                       maxit=100)
        bufs["c0"], bufs["cp"] = bufs["cp"], bufs["c0"]
    c_lf = ctx.download(bufs["c0"], n)
    assert np.abs(tau - sysm.tau).max() > 1e-5              # the ions moved
    assert np.abs(tau_vv - tau).max() < 1e-11
    assert np.abs(c_vv - c_lf).max() < 1e-9 * np.abs(c_lf).max()


def test_cp_energy_conservation_from_ground_state(ground16):
    """CP dynamics as production runs it: from the ODIIS ground state, 10
    velocity Verlet CP steps (CPMD's order, half kicks from rest) at dt 4
    a.u., emass 400 a.u. in each precision.  CPMD's constant of motion
    E_ham = E_KS + ekinp + ekinc (md_driver; ekinc from rekine after rortv)
    stays within the bound below, and the two precisions end within 1e-6
    Ha/atom."""
    sysm, ks, bufs, c_gs = ground16
    ctx = ks.ctx
    f = sysm.f
    n = len(f)
    nat = sysm.nat
    dt = 4.0
    out = {}
    for mode in ("fp64", "mixed"):
        ctx.set_precision("fp64")
        ctx.upload(c_gs, bufs["c0"])
        ctx.axpby_wfn(0.0, bufs["c0"], 0.0, bufs["cm"], n)
        st = ks.cp_init(bufs["c0"], bufs["cm"], bufs["c2"], f, sysm.tau, emass=EMASS, ion_mass=ION_MASS)
        e0 = st["eham"]
        ctx.set_precision(mode)
        hist = []
        for _ in range(10):
            ks.cp_step(st, bufs, f, dt)
            hist.append(st["eham"])
        ctx.set_precision("fp64")
        st_end = ks.cp_init(bufs["c0"], bufs["cm"], bufs["c2"], f, st["tau"], vion=st["vion"], emass=EMASS,
                            ion_mass=ION_MASS)
        out[mode] = (e0, st_end["eham"], float(np.abs(st["tau"] - sysm.tau).max()), hist)
    for mode, (e0, e1, moved, hist) in out.items():
        print(f"{mode}: E_ham drift {(e1 - e0) / nat:.3e} Ha/atom, max excursion "
              f"{max(abs(h - e0) for h in hist) / nat:.3e}, ions moved up to {moved:.2e} bohr", flush=True)
        assert moved > 1e-4                          # the ions do move
        assert abs(e1 - e0) / nat < 1e-5
    assert abs(out["mixed"][1] - out["fp64"][1]) / nat < 1e-6
