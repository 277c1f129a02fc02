"""Adjoint checks on the GPU (the cfg1 system: a = 12 bohr, 8 atoms, 40 Ry,
16 states, 54^3 mesh):

* E_KS, c2 = -f H c and the ionic forces from the device kernels equal the
  oracle's on the same seeded inputs (FP64);
* the H psi adjoint (vpsi + fnonloc) is the gradient of the device E_KS along
  8 seeded random tangent directions (central differences);
* the ionic forces (potfor + Ewald + rnlfor) are -dE_KS/dR for 6 seeded atoms
  x 3 directions with an ionic step of 1e-4 bohr.

vpsi recomputes psi(r) from C for every state pair and never stores psi(r)
for all states (the recompute-psi backward).  Every check runs twice: with
no device memory limit, and with the library's device memory capped
(mc_set_mem_limit) so that the transforms work on two state pairs per batch
instead of all eight."""
import gc

import numpy as np
import pytest

import oracle
from cpmd_amd import ops
from cpmd_amd import system as S
from cpmd_amd.ks import KSDevice
from test_oracle_energy import ks_system, tangent_directions

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", params=["uncapped", "capped"])
def cfg3(request):
    sysm, K, c = ks_system(12.0, 8, 40.0, 16, seed=1234)
    ks = KSDevice(sysm)
    ks.set_ions(sysm.tau)
    d_c = ks.ctx.upload(c)
    n = c.shape[1]
    d_c2, d_p = ks.ctx.alloc_wfn(n), ks.ctx.alloc_wfn(n)
    if request.param == "capped":
        # one full evaluation allocates every buffer; the cap then leaves the
        # transform workspaces room for two of the eight state pairs
        ks.energy(d_c, sysm.f, gradient=d_c2, forces=True)
        ks.ctx.synchronize()
        gc.collect()                       # buffers of earlier tests are returned first
        s, cols, _ = ops.layout_stats(sysm.cell, sysm.gset)
        n1, n2, n3 = sysm.cell.nr
        per_pair = 16 * (s * n3 + cols * n2 * n3)
        cur, _, _ = ops.mem_stats()
        cap = cur - 6 * per_pair + (256 << 20)   # the library keeps 256 MiB for what a step allocates later
        ops.set_mem_limit(cap)
        print(f"cfg3 with the device memory capped at {cap} bytes", flush=True)
    yield sysm, K, c, ks, d_c, d_c2, d_p
    ops.set_mem_limit(0)
    ks.ctx.close()


def test_device_ks_matches_oracle(gpu, cfg3):
    sysm, K, c, ks, d_c, d_c2, _ = cfg3
    f = sysm.f
    n = c.shape[1]
    r = ks.energy(d_c, f, gradient=d_c2, forces=True)
    o = oracle.ks_energy(K, c, f, sysm.tau, forces=True, gradient=True)
    for k in ("ekin", "ehep", "epseu", "exc", "esr", "eself", "enl", "etot"):
        assert r[k] == pytest.approx(o[k], rel=1e-11, abs=1e-11), k
    c2 = ks.ctx.download(d_c2, n)
    assert np.abs(c2 - o["c2"]).max() < 1e-11 * np.abs(o["c2"]).max()
    assert np.abs(r["fion"] - o["fion"]).max() < 1e-9 * max(1.0, np.abs(o["fion"]).max())


def test_h_psi_is_gradient(gpu, cfg3):
    sysm, K, c, ks, d_c, d_c2, d_p = cfg3
    f = sysm.f
    n = c.shape[1]
    ks.energy(d_c, f, gradient=d_c2)
    c2 = ks.ctx.download(d_c2, n)
    h = 1e-4
    for d in tangent_directions(c, 8, seed=77):
        ks.ctx.upload(np.asfortranarray(c + h * d), d_p)
        ep = ks.energy(d_p, f)["etot"]
        ks.ctx.upload(np.asfortranarray(c - h * d), d_p)
        em = ks.energy(d_p, f)["etot"]
        g = -2.0 * np.trace(S.dotp_gamma(d, c2))
        assert (ep - em) / (2 * h) == pytest.approx(g, rel=1e-6)


def test_forces_are_position_derivatives(gpu, cfg3):
    sysm, K, c, ks, d_c, _, _ = cfg3
    f = sysm.f
    tau = sysm.tau
    fion = ks.energy(d_c, f, tau=tau, forces=True)["fion"]
    rng = np.random.default_rng(5)
    atoms = rng.choice(8, size=6, replace=False)
    h = 1e-4
    for ia in atoms:
        for k in range(3):
            tp, tm = tau.copy(), tau.copy()
            tp[ia, k] += h
            tm[ia, k] -= h
            fd = -(ks.energy(d_c, f, tau=tp)["etot"] - ks.energy(d_c, f, tau=tm)["etot"]) / (2 * h)
            assert fion[k, ia] == pytest.approx(fd, rel=1e-6, abs=1e-7)
    ks.set_ions(tau)
