"""RESTART files from and to device wavefunctions (mc_restart_write_dev /
mc_restart_read_dev), and a CP run resumed from one.

* device C0 / CM -> RESTART -> other device buffers: bit-exact, and the file
  equals the host writer's bytes for the same arrays;
* 6 velocity Verlet CP steps in one go against 3 steps, a RESTART written
  (coordinates, ionic velocities, C0, CM), a new context built from the file
  alone, and 3 more steps: the same positions and coefficients to 1e-12
  (the restart repeats rortv on the read velocities, as CPMD's start does)."""
import numpy as np
import pytest

from cpmd_amd import ops
from cpmd_amd import restart as R
from cpmd_amd import system as S
from cpmd_amd.ks import KSDevice
from oracle import restart as OR

pytestmark = pytest.mark.gpu

ION_MASS = 12.0 * 1822.888486


def test_device_round_trip(gpu, tmp_path):
    sysm = S.make_system(12.0, 8, 40.0, nstate=10, with_density=True)
    ctx = ops.Context(sysm.cell, sysm.gset)
    rng = np.random.default_rng(4)
    c0 = S.initial_coefficients(sysm)
    cm = np.asfortranarray(1e-3 * (rng.standard_normal(c0.shape) + 1j * rng.standard_normal(c0.shape)))
    d0, dm = ctx.upload(c0), ctx.upload(cm)
    h = R.header_for(sysm)
    p = tmp_path / "RESTART.1"
    R.write_restart_dev(ctx, p, h, sysm.tau, c0_dev=d0, cm_dev=dm, f=sysm.f)
    assert p.read_bytes() == OR.wv30_bytes(h, sysm.tau, c0=c0, cm=cm, f=sysm.f)
    e0, em = ctx.alloc_wfn(10), ctx.alloc_wfn(10)
    r = R.read_restart_dev(ctx, p, c0_dev=e0, cm_dev=em)
    assert r["sections"] == R.C0 | R.CM | R.OCC and np.array_equal(r["tau0"], sysm.tau)
    assert np.array_equal(ctx.download(e0, 10), c0) and np.array_equal(ctx.download(em, 10), cm)
    nd = 2 * ctx.ld * 10
    assert np.array_equal(e0.to_host(nd), d0.to_host(nd))     # device layout, padding rows included


def test_cp_run_resumed_from_restart(gpu, tmp_path):
    sysm = S.make_system(12.0, 8, 40.0, nstate=16, seed=5, with_density=True)
    f = sysm.f
    n, dt, emass = 16, 4.0, 400.0
    ks = KSDevice(sysm)
    ctx = ks.ctx
    bufs = {k: ctx.alloc_wfn(n) for k in ("c0", "cm", "cp", "c2")}
    ks.set_ions(sysm.tau)
    ctx.upload(S.initial_coefficients(sysm, seed=5), bufs["c0"])
    assert ks.optimize(bufs["c0"], f, bufs["c2"], tol=1e-6, mixed=False, maxiter=400)["converged"]
    c_gs = ctx.download(bufs["c0"], n)

    def start(c0, tau, vion, cm=None):
        ctx.upload(c0, bufs["c0"])
        if cm is None:
            ctx.axpby_wfn(0.0, bufs["c0"], 0.0, bufs["cm"], n)
        else:
            ctx.upload(cm, bufs["cm"])
        return ks.cp_init(bufs["c0"], bufs["cm"], bufs["c2"], f, tau, vion=vion, emass=emass, ion_mass=ION_MASS)

    st = start(c_gs, sysm.tau, None)
    for _ in range(6):
        ks.cp_step(st, bufs, f, dt, epsog=1e-12, maxit=100)
    c_ref, tau_ref = ctx.download(bufs["c0"], n), st["tau"].copy()

    st = start(c_gs, sysm.tau, None)
    for _ in range(3):
        ks.cp_step(st, bufs, f, dt, epsog=1e-12, maxit=100)
    p = tmp_path / "RESTART.1"
    R.write_restart_dev(ctx, p, R.header_for(sysm), st["tau"], c0_dev=bufs["c0"], cm_dev=bufs["cm"], velp=st["vion"],
                        taui=sysm.tau, f=f)
    ks.ctx.close()

    # a new context, everything from the file
    r = R.read_restart(p)
    assert r["sections"] == R.VELP | R.TAUI | R.C0 | R.CM | R.OCC
    ks2 = KSDevice(sysm)
    ctx = ks2.ctx
    bufs = {k: ctx.alloc_wfn(n) for k in ("c0", "cm", "cp", "c2")}
    R.read_restart_dev(ctx, p, c0_dev=bufs["c0"], cm_dev=bufs["cm"])
    st = ks2.cp_init(bufs["c0"], bufs["cm"], bufs["c2"], r["f"], r["tau0"], vion=r["velp"], emass=emass,
                     ion_mass=ION_MASS)
    for _ in range(3):
        ks2.cp_step(st, bufs, r["f"], dt, epsog=1e-12, maxit=100)
    c_res = ctx.download(bufs["c0"], n)
    assert np.abs(st["tau"] - tau_ref).max() < 1e-12
    assert np.abs(c_res - c_ref).max() < 1e-12 * np.abs(c_ref).max()
