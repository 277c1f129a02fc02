"""Transform lengths with a factor 7 (the meshes a CPMD built with FFTW
chooses, src/fftchk_utils.mod.F90:73-97): rhoofr and vpsi against the oracle.

* a 28 x 42 x 14 mesh (4*7, 2*3*7, 2*7: the runtime-plan engine, radix-7
  stages on every axis), FP64 to the parity tolerances;
* E_KS, c2 and the ionic forces on a 28^3 mesh (the density transforms of
  vofrho too) against the oracle, FP64 to 1e-11;
* one state pair on the 224^3 mesh of the 512-atom 36 bohr system with the
  FFTW table (the compile-time 16 x 14 engine), FP64 to the parity
  tolerances and FP32 transforms to 2e-6 relative."""
import numpy as np
import pytest

import oracle
from cpmd_amd import ops
from cpmd_amd import system as S
from cpmd_amd.ks import KSDevice
from test_gpu_wfft import RHO_TOL, C2_TOL, _check_rho, _check_vpsi
from test_oracle_energy import ks_system

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("nstate", [5, 6])
def test_radix7_runtime_engine(gpu, nstate):
    cell = S.orthorhombic_cell(9.5, 14.5, 4.8, 20.0, table="fftw")
    assert tuple(cell.nr) == (28, 42, 14)
    gs = S.gvectors(cell, with_density=False)
    sysm = S.SyntheticSystem(cell=cell, gset=gs, tau=np.zeros((1, 3)), nstate=nstate, f=np.full(nstate, 2.0))
    c = S.initial_coefficients(sysm, seed=17, nstate=nstate)
    ctx = ops.Context(cell, gs)
    L = oracle.Layout(cell, gs)
    f = np.linspace(2.0, 0.5, nstate)
    _check_rho(ctx, L, c, f)
    v = S.real_space_potential(sysm, seed=5)
    _check_vpsi(ctx, L, c, f, v)


def test_radix7_ks_energy(gpu):
    sysm, K, c = ks_system(8.5, 3, 25.0, 6, seed=9, fft_table="fftw")
    assert tuple(sysm.cell.nr) == (28, 28, 28)
    ks = KSDevice(sysm)
    ks.set_ions(sysm.tau)
    d_c = ks.ctx.upload(c)
    n = c.shape[1]
    d_c2 = ks.ctx.alloc_wfn(n)
    r = ks.energy(d_c, sysm.f, gradient=d_c2, forces=True)
    o = oracle.ks_energy(K, c, sysm.f, sysm.tau, forces=True, gradient=True)
    for k in ("ekin", "ehep", "epseu", "exc", "esr", "eself", "enl", "etot"):
        assert r[k] == pytest.approx(o[k], rel=1e-11, abs=1e-11), k
    c2 = ks.ctx.download(d_c2, n)
    assert np.abs(c2 - o["c2"]).max() < 1e-11 * np.abs(o["c2"]).max()
    assert np.abs(r["fion"] - o["fion"]).max() < 1e-9 * max(1.0, np.abs(o["fion"]).max())
    ks.ctx.close()


def test_fftw_mesh_224_pair(gpu):
    cfg = S.CONFIGS["cfg5"]
    sysm = S.make_system(cfg["a"], cfg["nat"], cfg["ecut"], nstate=2, fft_table="fftw")
    assert tuple(sysm.cell.nr) == (224, 224, 224)
    c = S.initial_coefficients(sysm, seed=4321)
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
