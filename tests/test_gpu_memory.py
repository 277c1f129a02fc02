"""The device memory budget (mc_set_mem_limit): the plain FP64 E_KS, c2 and
F_ion of the 512-state / 192^3 system (cfg4) with the library's device memory
capped at 16 GiB and at 7 GiB (which leaves room for a few dozen state pairs
per transform batch instead of 128) equal the uncapped evaluation to 1e-12,
and the library's peak device memory stays within each cap.  psi(r) is never
stored for all states: vpsi recomputes it from C per batch."""
import numpy as np
import pytest

from cpmd_amd import ops
from cpmd_amd import system as S
from cpmd_amd.ks import KSDevice
from cpmd_amd._lib import McError

pytestmark = pytest.mark.gpu

GIB = 1 << 30


def test_cfg4_fp64_under_memory_caps(gpu):
    cfg = S.CONFIGS["cfg4"]
    sysm = S.make_system(cfg["a"], cfg["nat"], cfg["ecut"], nstate=cfg["nstate"], with_density=True)
    n = sysm.nstate
    c0 = S.initial_coefficients(sysm)
    ops.set_mem_limit(0)
    try:
        results = {}
        d0 = d2 = None
        for cap in (0, 16 * GIB, 7 * GIB):
            ops.set_mem_limit(cap)
            ks = KSDevice(sysm)
            if d0 is None:
                d0, d2 = ks.ctx.upload(c0), ks.ctx.alloc_wfn(n)
            ks.set_ions(sysm.tau)
            ops.mem_reset_peak()
            print(f"cfg4 fp64 energy + gradient + forces, memory cap {cap / GIB:.0f} GiB", flush=True)
            r = ks.energy(d0, sysm.f, gradient=d2, forces=True)
            ks.ctx.synchronize()
            cur, peak, lim = ops.mem_stats()
            results[cap] = (r, ks.ctx.download(d2, n), peak)
            print(f"  peak {peak / GIB:.2f} GiB, E_KS {r['etot']:.12f}", flush=True)
            if cap:
                assert lim == cap and peak <= cap
            ks.ctx.close()
        r0, g0, peak0 = results[0]
        assert peak0 > 7 * GIB                      # the tight cap does bind
        for cap in (16 * GIB, 7 * GIB):
            r, g, _ = results[cap]
            for k in ("ekin", "ehep", "epseu", "exc", "enl", "etot"):
                assert r[k] == pytest.approx(r0[k], rel=1e-12, abs=1e-12), (cap, k)
            assert np.abs(g - g0).max() <= 1e-12 * np.abs(g0).max()
            assert np.abs(r["fion"] - r0["fion"]).max() <= 1e-12 * np.abs(r0["fion"]).max()
    finally:
        ops.set_mem_limit(0)


def test_limit_refuses_what_does_not_fit(gpu):
    sysm = S.make_system(12.0, 8, 40.0, nstate=16, with_density=True)
    ctx = ops.Context(sysm.cell, sysm.gset)
    cur, _, _ = ops.mem_stats()
    try:
        ops.set_mem_limit(cur + (1 << 20))
        with pytest.raises(McError, match="memory limit"):
            ops.DeviceArray(8 << 20)
        c = S.initial_coefficients(sysm)
        with pytest.raises(McError, match="memory limit"):
            ctx.rhoofr(c, sysm.f)                    # the wavefunction staging alone exceeds the cap
    finally:
        ops.set_mem_limit(0)
    rho, *_ = ctx.rhoofr(c, sysm.f)                  # and works again without it
    assert rho.sum() > 0
