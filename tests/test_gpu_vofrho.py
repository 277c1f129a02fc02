"""GPU parity of the density -> potential step (vofrhoa + vofrhob, Pade LDA)
and of the Ewald sum against the oracle on the same seeded inputs.  The
oracle's functions are pinned in test_oracle_energy.py (Madelung constant,
independence of the Gaussian split, finite differences of the energies)."""
import math

import numpy as np
import pytest

import oracle
from cpmd_amd import ops, _lib
from cpmd_amd import system as S

pytestmark = pytest.mark.gpu


def rel(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


def _case(a, nat, ecut, nstate, seed, raggio=1.2):
    sysm = S.make_system(a, nat, ecut, nstate=nstate, seed=seed, with_density=True)
    gs = sysm.gset
    c = S.initial_coefficients(sysm, seed=seed)
    ctx = ops.Context(sysm.cell, gs)
    rho, _, _, _ = ctx.rhoofr(c, sysm.f)
    zv, rg, vps, rhops = S.local_pp_set(sysm, gs.hg, raggio=raggio)
    nh = np.array([n // 2 + 1 for n in sysm.cell.nr])
    inyh = gs.miller + nh
    gk = gs.miller @ sysm.cell.bvec
    ctx.set_density_gset(inyh, gs.hg, gk)
    ctx.set_local_pp([nat], zv, rg, vps, rhops)
    ctx.set_ions(sysm.tau, sysm.cell.avec)
    frac = sysm.tau @ np.linalg.inv(sysm.cell.avec)
    L = oracle.Layout(sysm.cell, gs)
    return sysm, ctx, L, rho, inyh, gk, zv, rg, vps, rhops, frac


@pytest.mark.parametrize("cfg", ["small", "cfg1", "cfg2"])
def test_vofrho_parity(gpu, cfg):
    a, nat, ecut, nstate = {"small": (9.0, 3, 25.0, 6), "cfg1": (12.0, 8, 40.0, 16),
                            "cfg2": (24.0, 64, 70.0, 16)}[cfg]
    sysm, ctx, L, rho, inyh, gk, zv, rg, vps, rhops, frac = _case(a, nat, ecut, nstate, seed=7)
    gs = sysm.gset
    v, en, f = ctx.vofrho(rho)
    ov, oen, of, _ = oracle.vofrho(L, inyh, gs.hg, gk, [nat], vps, rhops, frac, rho)
    assert rel(v, ov) < 1e-12
    # sums over up to 5e5 G vectors / 2e6 mesh points in a different order
    for k in ("ehep", "ehii", "ehee", "epseu", "exc", "vxc"):
        assert en[k] == pytest.approx(oen[k], rel=1e-11), k
    assert en["vploc"] == pytest.approx(oen["vploc"], rel=1e-12, abs=1e-15)
    assert rel(f, of) < 1e-10
    # padding of v(r) is zero and the device-resident entry point agrees
    nr, kr = sysm.cell.nr, sysm.cell.kr
    vv = v.reshape(kr[2], kr[1], kr[0])
    assert np.all(vv[:, :, nr[0]:] == 0) and np.all(vv[:, nr[1]:, :] == 0) and np.all(vv[nr[2]:] == 0)
    d_rho = ops.DeviceArray(8 * rho.size).from_host(rho)
    d_v = ops.DeviceArray(8 * rho.size)
    en2, _ = ctx.vofrho_dev(d_rho, d_v)
    assert np.array_equal(d_v.to_host(rho.size), v)
    assert en2 == en


@pytest.mark.parametrize("na", [(5, 3), (5, 0, 3)])
def test_vofrho_parity_species(gpu, na):
    """Several species (one of them without atoms): the species blocks of the
    stick structure factors and of the potfor contraction."""
    sysm, ctx, L, rho, inyh, gk, _, _, _, _, frac = _case(12.0, 8, 40.0, 16, seed=5)
    gs = sysm.gset
    sc = (1.0, 0.6, 1.3)
    spp = [S.local_pp_set(sysm, gs.hg, raggio=r) for r in (1.2, 0.9, 1.05)[:len(na)]]
    zv = np.concatenate([p[0] * sc[i] for i, p in enumerate(spp)])
    rg = np.concatenate([p[1] for p in spp])
    vps = np.asfortranarray(np.vstack([p[2] * sc[i] for i, p in enumerate(spp)]))
    rhops = np.asfortranarray(np.vstack([p[3] * sc[i] for i, p in enumerate(spp)]))
    ctx.set_local_pp(list(na), zv, rg, vps, rhops)
    ctx.set_ions(sysm.tau, sysm.cell.avec)
    v, en, f = ctx.vofrho(rho)
    ov, oen, of, _ = oracle.vofrho(L, inyh, gs.hg, gk, list(na), vps, rhops, frac, rho)
    assert rel(v, ov) < 1e-12
    for k in ("ehep", "ehii", "ehee", "epseu", "exc", "vxc"):
        assert en[k] == pytest.approx(oen[k], rel=1e-11), k
    assert rel(f, of) < 1e-10
    assert np.abs(of).max() > 0.0


def test_ewald_parity(gpu):
    sysm, ctx, L, rho, inyh, gk, zv, rg, vps, rhops, frac = _case(12.0, 8, 40.0, 16, seed=3)
    esr, eself, f = ctx.ewald(iesr=2)
    oesr, oeself, of = oracle.ewald([8], zv, rg, sysm.tau, sysm.cell.avec, iesr=2)
    assert esr == pytest.approx(oesr, rel=1e-12)
    assert eself == pytest.approx(oeself, rel=1e-14)
    assert rel(f, of) < 1e-11
    # a triclinic cell (Ewald only: the density set is not used)
    ht = np.array([[12.0, 0.0, 0.0], [3.0, 11.0, 0.0], [1.0, 2.0, 10.5]])
    tau = np.random.default_rng(4).uniform(0, 10, size=(8, 3))
    ctx.set_ions(tau, ht)
    esr, _, f = ctx.ewald(iesr=3)
    oesr, _, of = oracle.ewald([8], zv, rg, tau, ht, iesr=3)
    assert esr == pytest.approx(oesr, rel=1e-12)
    assert rel(f, of) < 1e-11


def test_vofrho_requires_setup(gpu):
    sysm = S.make_system(9.0, 3, 25.0, nstate=6, with_density=True)
    ctx = ops.Context(sysm.cell, sysm.gset)
    with pytest.raises(_lib.McError):
        ctx.set_ions(sysm.tau, sysm.cell.avec)            # no local PP yet
    with pytest.raises(_lib.McError):
        ctx.nat_ions = 3
        ctx.vofrho(np.zeros(int(np.prod(sysm.cell.kr))))
