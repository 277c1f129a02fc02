"""GPU parity of rhoofr and vpsi (HIP, through the C ABI) against the oracle
on the same seeded inputs, plus size-independent properties at full size.

Tolerances (FP64 path): max |d rho| < 1e-11 (absolute, rho in e/bohr^3),
E_kin / charge sums to 1e-12 relative, vpsi force max |d c2| < 1e-11 * max|c2|.
"""
import numpy as np
import pytest

import oracle
from cpmd_amd import ops
from cpmd_amd import system as S
from util import small_system, config_system

pytestmark = pytest.mark.gpu

RHO_TOL = 1e-11
C2_TOL = 1e-11


def _check_rho(ctx, L, c, f):
    rho_g, ek_g, cg_g, cr_g = ctx.rhoofr(c, f)
    rho_o, ek_o, cg_o, cr_o = oracle.rhoofr(L, c, f)
    err = np.abs(rho_g - rho_o).max()
    assert err < RHO_TOL, err
    assert ek_g == pytest.approx(ek_o, rel=1e-12, abs=1e-300)
    assert cg_g == pytest.approx(cg_o, rel=1e-12, abs=1e-300)
    assert cr_g == pytest.approx(cr_o, rel=1e-12, abs=1e-300)
    return rho_g


def _check_vpsi(ctx, L, c, f, v, seed=3):
    rng = np.random.default_rng(seed)
    c2 = np.asfortranarray(rng.standard_normal(c.shape) + 1j * rng.standard_normal(c.shape))
    out_g = ctx.vpsi(c, c2.copy(order="F"), f, v)
    out_o = oracle.vpsi(L, c, c2, f, v)
    scale = np.abs(out_o).max()
    err = np.abs(out_g - out_o).max()
    assert err < C2_TOL * scale, (err, scale)


@pytest.mark.parametrize("kind,nstate", [("cubic", 6), ("cubic", 5), ("ortho", 4), ("ortho2", 7), ("ortho2", 1),
                                         ("ortho3", 5)])
def test_rhoofr_small(gpu, kind, nstate):
    cell, gs, c = small_system(kind, nstate)
    ctx = ops.Context(cell, gs)
    L = oracle.Layout(cell, gs)
    f = np.linspace(2.0, 0.5, nstate)
    _check_rho(ctx, L, c, f)


def test_rhoofr_zero_occupations(gpu):
    cell, gs, c = small_system("cubic", 6)
    ctx = ops.Context(cell, gs)
    L = oracle.Layout(cell, gs)
    _check_rho(ctx, L, c, np.array([2.0, 2.0, 0.0, 0.0, 1.0, 0.0]))
    rho, ek, cg, cr = ctx.rhoofr(c, np.zeros(6))
    assert not rho.any() and ek == 0.0 and cg == 0.0 and cr == 0.0


@pytest.mark.parametrize("kind,nstate", [("cubic", 6), ("cubic", 3), ("ortho", 4), ("ortho2", 5), ("ortho2", 1),
                                         ("ortho3", 4)])
def test_vpsi_small(gpu, kind, nstate):
    cell, gs, c = small_system(kind, nstate)
    ctx = ops.Context(cell, gs)
    L = oracle.Layout(cell, gs)
    sysm = S.SyntheticSystem(cell=cell, gset=gs, tau=np.zeros((1, 3)), nstate=nstate, f=np.full(nstate, 2.0))
    v = S.real_space_potential(sysm, seed=11)
    f = np.array([2.0, 0.0, 1.0, 2.0, 0.5, 2.0, 2.0][:nstate])
    _check_vpsi(ctx, L, c, f, v)


@pytest.mark.parametrize("batch", [1, 3, 64])
def test_cfg1_parity(gpu, batch):
    sysm, c = config_system("cfg1")
    ctx = ops.Context(sysm.cell, sysm.gset)
    ctx.set_batch(batch, batch)
    L = oracle.Layout(sysm.cell, sysm.gset)
    rho = _check_rho(ctx, L, c, sysm.f)
    assert rho.min() > -1e-14
    v = S.real_space_potential(sysm)
    _check_vpsi(ctx, L, c, sysm.f, v)


def test_cfg5_pair(gpu):
    """One state pair on the 240^3 mesh of config 5 (the FP64 tile of the fused
    x / y engine is above 64 KB of LDS there): rhoofr and vpsi in FP64 to the
    parity tolerances and with FP32 transforms to 2e-6, against the oracle."""
    sysm, c = config_system("cfg5", nstate=2)
    assert tuple(sysm.cell.nr) == (240, 240, 240)
    ctx = ops.Context(sysm.cell, sysm.gset)
    L = oracle.Layout(sysm.cell, sysm.gset)
    f = sysm.f[:2]
    v = S.real_space_potential(sysm)
    rho_o, ek_o, cg_o, cr_o = oracle.rhoofr(L, c, f)
    c2_o = oracle.vpsi(L, c, np.zeros_like(c), f, v)
    for mode, rtol in (("fp64", None), ("fp32", 2e-6)):
        ctx.set_precision(mode)
        rho_g, ek_g, cg_g, cr_g = ctx.rhoofr(c, f)
        c2_g = ctx.vpsi(c, np.zeros_like(c), f, v)
        if rtol is None:
            assert np.abs(rho_g - rho_o).max() < RHO_TOL
            assert ek_g == pytest.approx(ek_o, rel=1e-12) and cg_g == pytest.approx(cg_o, rel=1e-12)
            assert np.abs(c2_g - c2_o).max() < C2_TOL * np.abs(c2_o).max()
        else:
            assert np.abs(rho_g - rho_o).max() < rtol * np.abs(rho_o).max(), mode
            assert np.abs(c2_g - c2_o).max() < rtol * np.abs(c2_o).max(), mode


@pytest.mark.parametrize("nstate", [3, 4])
def test_sphere_beyond_quarter_band(gpu, nstate):
    """dual 3 on a 128^3 mesh: the wavefunction sphere reaches |m| = 36 > N/4,
    so the x and y passes keep every frequency (the unpruned 16 x 8 engine);
    rhoofr and vpsi against the oracle in FP64 and with FP32 transforms."""
    cell = S.cubic_cell(19.5, 135.0, dual=3.0)
    assert tuple(cell.nr) == (128, 128, 128)
    gs = S.gvectors(cell, with_density=False)
    assert np.abs(gs.miller[:gs.ngw, :2]).max() >= 128 // 4
    sysm = S.SyntheticSystem(cell=cell, gset=gs, tau=np.zeros((1, 3)), nstate=nstate, f=np.full(nstate, 2.0))
    c = S.initial_coefficients(sysm, seed=31, nstate=nstate)
    ctx = ops.Context(cell, gs)
    L = oracle.Layout(cell, gs)
    f = np.linspace(2.0, 1.0, nstate)
    v = S.real_space_potential(sysm, seed=7)
    _check_rho(ctx, L, c, f)
    _check_vpsi(ctx, L, c, f, v)
    ctx.set_precision("fp32")
    rho_o = oracle.rhoofr(L, c, f)[0]
    assert np.abs(ctx.rhoofr(c, f)[0] - rho_o).max() < 2e-6 * np.abs(rho_o).max()
    c2_o = oracle.vpsi(L, c, np.zeros_like(c), f, v)
    assert np.abs(ctx.vpsi(c, np.zeros_like(c), f, v) - c2_o).max() < 2e-6 * np.abs(c2_o).max()


def test_cfg2_parity_subset(gpu):
    sysm, c = config_system("cfg2", nstate=6)
    ctx = ops.Context(sysm.cell, sysm.gset)
    L = oracle.Layout(sysm.cell, sysm.gset)
    f = sysm.f[:6]
    _check_rho(ctx, L, c, f)
    _check_vpsi(ctx, L, c, f, S.real_space_potential(sysm))


def test_cfg2_all_states(gpu):
    """rhoofr and vpsi on all 128 states of config 2 (128^3, 68,338 plane
    waves) against the oracle's threaded CPU path (oracle/cpu_step: the same
    arithmetic as oracle.c, checked against it to 1e-14 in
    tests/test_oracle_cpu_step.py), at the FP64 parity tolerances."""
    from oracle.cpu_step import CpuStep
    sysm, c = config_system("cfg2")
    assert c.shape[1] == 128
    ctx = ops.Context(sysm.cell, sysm.gset)
    cs = CpuStep(sysm.cell, sysm.gset)
    f = sysm.f.copy()
    f[-3:] = [1.0, 0.0, 0.0]                       # a partly filled and two empty states
    rho_g, ek_g, cg_g, cr_g = ctx.rhoofr(c, f)
    rho_o = cs.rhoofr(c, f)
    assert np.abs(rho_g - rho_o).max() < RHO_TOL
    assert cr_g == pytest.approx(cg_g, rel=1e-12)  # Parseval: the FP64 charge sums agree
    v = S.real_space_potential(sysm)
    rng = np.random.default_rng(8)
    c2 = np.asfortranarray(rng.standard_normal(c.shape) + 1j * rng.standard_normal(c.shape))
    out_g = ctx.vpsi(c, c2.copy(order="F"), f, v)
    out_o = cs.vpsi(c, c2, f, v)
    assert np.abs(out_g - out_o).max() < C2_TOL * np.abs(out_o).max()


def test_device_resident_roundtrip(gpu):
    sysm, c = config_system("cfg1")
    ctx = ops.Context(sysm.cell, sysm.gset)
    n = c.shape[1]
    d = ctx.upload(c)
    assert np.array_equal(ctx.download(d, n), c)            # permutation round trip is exact
    rho_d = ops.DeviceArray(8 * ctx.nnr1)
    ek, cg, cr = ctx.rhoofr_dev(d, sysm.f, n, rho_d)
    rho_h, ek_h, cg_h, cr_h = ctx.rhoofr(c, sysm.f)
    assert np.array_equal(rho_d.to_host(ctx.nnr1), rho_h)
    assert (ek, cg, cr) == (ek_h, cg_h, cr_h)
    v = S.real_space_potential(sysm)
    vd = ops.DeviceArray(8 * ctx.nnr1).from_host(v)
    c2d = ctx.alloc_wfn(n)
    ctx.vpsi_dev(d, c2d, sysm.f, vd, n)
    out_h = ctx.vpsi(c, np.zeros_like(c), sysm.f, v)
    assert np.array_equal(ctx.download(c2d, n), out_h)


def test_cfg4_full_size_properties(gpu):
    """512 states on the 192^3 mesh: Parseval (csumr = csumg = sum f), rho >= 0,
    and the vpsi force of a constant potential is -f (G^2/2 + V0) c exactly."""
    sysm, c = config_system("cfg4")
    ctx = ops.Context(sysm.cell, sysm.gset)
    n = c.shape[1]
    d = ctx.upload(c)
    rho_d = ops.DeviceArray(8 * ctx.nnr1)
    ek, cg, cr = ctx.rhoofr_dev(d, sysm.f, n, rho_d)
    assert cg == pytest.approx(sysm.f.sum(), rel=1e-12)
    assert cr == pytest.approx(cg, rel=1e-12)
    rho = rho_d.to_host(ctx.nnr1)
    assert rho.min() > -1e-14
    g2 = sysm.cell.tpiba2 * sysm.gset.hg[:sysm.gset.ngw]
    ek_ref = sysm.cell.tpiba2 * float(np.sum(sysm.f * (sysm.gset.hg[:sysm.gset.ngw, None] * np.abs(c) ** 2).sum(0)))
    assert ek == pytest.approx(ek_ref, rel=1e-12)
    v0 = 0.25
    vd = ops.DeviceArray(8 * ctx.nnr1).from_host(np.full(ctx.nnr1, v0))
    c2d = ctx.alloc_wfn(n)
    ctx.vpsi_dev(d, c2d, sysm.f, vd, n)
    out = ctx.download(c2d, n)
    ref = -(sysm.f[None, :]) * (0.5 * g2[:, None] + v0) * c
    assert np.abs(out - ref).max() < 1e-11 * np.abs(ref).max()


# ---------------------------------------------------------------------------
# FP32 transform mode (mixed precision): same kernels in float, FP64 C0/C2/rho.
# Tolerances are FP32 round-off over three passes: 2e-6 relative.
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("name,nstate", [("cfg1", None), ("cfg2", 6)])
def test_fp32_transforms(gpu, name, nstate):
    sysm, c = config_system(name, nstate=nstate)
    f = sysm.f[:c.shape[1]]
    ctx = ops.Context(sysm.cell, sysm.gset)
    ctx.set_precision("fp32")
    L = oracle.Layout(sysm.cell, sysm.gset)
    rho_g, ek_g, cg_g, cr_g = ctx.rhoofr(c, f)
    rho_o, ek_o, cg_o, cr_o = oracle.rhoofr(L, c, f)
    assert np.abs(rho_g - rho_o).max() < 2e-6 * np.abs(rho_o).max()
    assert ek_g == pytest.approx(ek_o, rel=1e-12)        # E_kin is summed from C0 in FP64
    assert cg_g == pytest.approx(cg_o, rel=1e-12)
    assert cr_g == pytest.approx(cr_o, rel=2e-6)
    assert cr_g == pytest.approx(cg_g, rel=1e-12)        # FP32 density rescaled to the FP64 charge
    v = S.real_space_potential(sysm)
    c2 = ctx.vpsi(c, np.zeros_like(c), f, v)
    c2_o = oracle.vpsi(L, c, np.zeros_like(c), f, v)
    assert np.abs(c2 - c2_o).max() < 2e-6 * np.abs(c2_o).max()
    # and back to FP64 on the same context
    ctx.set_precision("fp64")
    assert np.abs(ctx.rhoofr(c, f)[0] - rho_o).max() < 1e-11


@pytest.mark.parametrize("kind,nstate", [("ortho", 5), ("ortho3", 4)])
def test_fp32_transforms_small(gpu, kind, nstate):
    """FP32 transforms on a ragged y tile (n2 = 18 < lines per workgroup) and on
    the runtime-plan lengths (n1 = 30, n2 = n3 = 20)."""
    cell, gs, c = small_system(kind, nstate)
    f = np.linspace(2.0, 0.5, nstate)
    ctx = ops.Context(cell, gs)
    ctx.set_precision("fp32")
    L = oracle.Layout(cell, gs)
    rho_g, ek_g, cg_g, cr_g = ctx.rhoofr(c, f)
    rho_o, ek_o, cg_o, cr_o = oracle.rhoofr(L, c, f)
    assert np.abs(rho_g - rho_o).max() < 2e-6 * np.abs(rho_o).max()
    assert cr_g == pytest.approx(cg_g, rel=1e-12)
    sysm = S.SyntheticSystem(cell=cell, gset=gs, tau=np.zeros((1, 3)), nstate=nstate, f=np.full(nstate, 2.0))
    v = S.real_space_potential(sysm, seed=11)
    c2 = ctx.vpsi(c, np.zeros_like(c), f, v)
    c2_o = oracle.vpsi(L, c, np.zeros_like(c), f, v)
    assert np.abs(c2 - c2_o).max() < 2e-6 * np.abs(c2_o).max()


def test_forces_c2_local_only(gpu):
    """Without projectors the electronic gradient is vpsi alone: equal to vpsi
    into a cleared c2, whatever c2 held before, padding rows included."""
    cell, gs, c = small_system("cubic", 5)
    n = c.shape[1]
    ctx = ops.Context(cell, gs)
    sysm = S.SyntheticSystem(cell=cell, gset=gs, tau=np.zeros((1, 3)), nstate=n, f=np.full(n, 2.0))
    v = S.real_space_potential(sysm, seed=13)
    f = np.array([2.0, 0.0, 1.0, 2.0, 0.5])
    d = ctx.upload(c)
    vd = ops.DeviceArray(8 * ctx.nnr1).from_host(v)
    ref = ctx.alloc_wfn(n)
    ctx.vpsi_dev(d, ref, f, vd, n)
    nd = 2 * ctx.ld * n
    out = ctx.alloc_wfn(n).from_host(np.random.default_rng(2).standard_normal(nd))
    ctx.forces_c2_dev(d, out, f, vd, n)
    ctx.synchronize()
    assert np.array_equal(out.to_host(nd), ref.to_host(nd))
