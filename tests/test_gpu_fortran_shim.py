"""The Fortran ISO_C_BINDING shim (cpmd_amd/fortran/micpmd.F90) called from
Fortran programs:

* CPMD's routine signatures (tests/fortran/shim_driver.F90) compared with the
  oracle on the cfg1 system (54^3 mesh, 3625 plane waves);
* the device-resident velocity Verlet CP step (tests/fortran/cp_driver.F90),
  in mixed precision and FP64, compared with the same step driven from
  Python."""
import subprocess
from pathlib import Path

import numpy as np
import pytest

import oracle
from cpmd_amd import build as B
from cpmd_amd import system as S
from util import config_system

pytestmark = pytest.mark.gpu


def test_fortran_shim_cfg1(gpu, tmp_path):
    assert B.DRIVER.exists(), f"{B.DRIVER} not built (run __graft_entry__.build())"
    sysm, c = config_system("cfg1")
    n = c.shape[1]
    gs, cell = sysm.gset, sysm.cell
    ngw = gs.ngw
    nh = np.array([x // 2 + 1 for x in cell.nr], np.int32)
    inyh = (gs.miller[:ngw] + nh).astype(np.int32)
    v = S.real_space_potential(sysm)
    rng = np.random.default_rng(5)
    cp = np.asfortranarray(c + 0.02 * (rng.standard_normal(c.shape) + 1j * rng.standard_normal(c.shape))
                           * np.exp(-gs.hg[:ngw, None] * cell.tpiba2 / 8))
    cp[0] = cp[0].real
    na, nhxs, ngh, lang, hmat, twnl, eigr = S.projector_set(sysm, gs.miller[:ngw], gs.hg[:ngw])
    nat = int(na[0])
    fin, fout = tmp_path / "in.bin", tmp_path / "out.bin"
    with open(fin, "wb") as fh:
        np.array(list(cell.nr) + list(cell.kr) + [ngw, n, nat], np.int32).tofile(fh)
        np.array([cell.tpiba2, cell.omega]).tofile(fh)
        inyh.tofile(fh)                                   # (ngw,3) C order == inyh(3,ngw) Fortran order
        gs.hg[:ngw].astype(np.float64).tofile(fh)
        np.asfortranarray(c).T.tofile(fh)                 # column-major complex
        sysm.f.astype(np.float64).tofile(fh)
        v.tofile(fh)
        cp.T.tofile(fh)
        twnl[:, 0, 0].astype(np.float64).tofile(fh)
        np.array([hmat[0, 0, 0]]).tofile(fh)
        np.asfortranarray(eigr).T.tofile(fh)
    r = subprocess.run([str(B.DRIVER), str(fin), str(fout)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    nnr1 = int(np.prod(cell.kr))
    out = np.fromfile(fout, dtype=np.uint8)
    off = 0

    def take(count, dtype, shape=None):
        nonlocal off
        a = np.frombuffer(out, dtype=dtype, count=count, offset=off)
        off += a.nbytes
        return a if shape is None else a.reshape(shape, order="F")

    rho = take(nnr1, np.float64)
    ekin, csumg, csumr = take(3, np.float64)
    c2 = take(ngw * n, np.complex128, (ngw, n))
    crgs = take(ngw * n, np.complex128, (ngw, n))
    smat = take(n * n, np.float64, (n, n))
    clow = take(ngw * n, np.complex128, (ngw, n))
    fnl = take(nat * n, np.float64, (nat, 1, n))
    c2nl = take(ngw * n, np.complex128, (ngw, n))
    assert off == out.size

    L = oracle.Layout(cell, gs)
    rho_o, ek_o, cg_o, cr_o = oracle.rhoofr(L, c, sysm.f)
    assert np.abs(rho - rho_o).max() < 1e-11
    assert ekin == pytest.approx(ek_o, rel=1e-12) and csumr == pytest.approx(cr_o, rel=1e-12)
    c2_o = oracle.vpsi(L, c, np.zeros_like(c), sysm.f, v)
    assert np.abs(c2 - c2_o).max() < 1e-11 * np.abs(c2_o).max()
    crgs_o, smat_o = oracle.rgs(cp)
    assert np.abs(crgs - crgs_o).max() < 1e-12 and np.abs(smat - smat_o).max() < 1e-11
    clow_o, _ = oracle.lowdin(cp)
    assert np.abs(clow - clow_o).max() < 1e-12
    P = oracle.Projectors(na=na, nhxs=1, ngh=ngh, lang=lang, hmat=hmat, twnl=twnl, eigr=eigr)
    fnl_o = oracle.rnlsm1(P, c)
    assert np.abs(fnl - fnl_o).max() < 1e-12 * np.abs(fnl_o).max()
    c2nl_o = oracle.fnonloc(P, np.zeros_like(c), sysm.f, fnl_o)
    assert np.abs(c2nl - c2nl_o).max() < 1e-12 * np.abs(c2nl_o).max()


ION_MASS = 12.0 * 1822.888486


@pytest.mark.parametrize("mode", ["MIXED", "FP64"])
def test_fortran_cp_step_matches_python(gpu, tmp_path, mode):
    """CPMD's velocity Verlet CP step driven from Fortran through the module's
    device-resident entry points (tests/fortran/cp_driver.F90) and from Python
    (KSDevice.cp_init / cp_step) on the same inputs: the coefficients,
    velocities, positions, forces and ekinc agree bit for bit, E_KS to 1e-12
    (its E_NL sum is ordered differently)."""
    from cpmd_amd.ks import KSDevice
    assert B.CP_DRIVER.exists(), f"{B.CP_DRIVER} not built (run __graft_entry__.build())"
    sysm = S.make_system(12.0, 8, 40.0, nstate=16, with_density=True)
    gs, cell = sysm.gset, sysm.cell
    ngw, nhg, n, nat = gs.ngw, gs.nhg, sysm.nstate, sysm.nat
    c0 = S.initial_coefficients(sysm)
    dt, emass, nsteps, raggio, iesr = 1.0, 400.0, 2, 1.2, 2
    zv, rg, vps, rhops = S.local_pp_set(sysm, gs.hg, raggio=raggio)
    na, nhxs, ngh, lang, hmat, twnl, _ = S.projector_set(sysm, gs.miller[:ngw], gs.hg[:ngw], with_eigr=False)
    nh = np.array([x // 2 + 1 for x in cell.nr], np.int32)
    fin, fout = tmp_path / "in.bin", tmp_path / "out.bin"
    with open(fin, "wb") as fh:
        np.array(list(cell.nr) + list(cell.kr) + [ngw, nhg, n, nat, iesr, 0], np.int32).tofile(fh)
        np.array([cell.tpiba2, cell.omega, dt, emass, ION_MASS, hmat[0, 0, 0], zv[0], rg[0]]).tofile(fh)
        (gs.miller + nh).astype(np.int32).tofile(fh)
        gs.hg.astype(np.float64).tofile(fh)
        (gs.miller @ cell.bvec).astype(np.float64).tofile(fh)
        vps[0].astype(np.float64).tofile(fh)
        rhops[0].astype(np.float64).tofile(fh)
        twnl[:, 0, 0].astype(np.float64).tofile(fh)
        sysm.tau.astype(np.float64).tofile(fh)
        np.asfortranarray(cell.avec).T.astype(np.float64).tofile(fh)
        np.asfortranarray(c0).T.tofile(fh)
        sysm.f.astype(np.float64).tofile(fh)
    r = subprocess.run([str(B.CP_DRIVER), str(fin), str(fout), mode, str(nsteps)], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    out = np.fromfile(fout, dtype=np.uint8)
    off = 0

    def take(count, dtype, shape=None):
        nonlocal off
        a = np.frombuffer(out, dtype=dtype, count=count, offset=off)
        off += a.nbytes
        return a if shape is None else a.reshape(shape, order="F")

    recs = [(take(1, np.float64)[0], take(3 * nat, np.float64, (3, nat)), take(1, np.float64)[0], None)]
    for _ in range(nsteps):
        recs.append((take(1, np.float64)[0], take(3 * nat, np.float64, (3, nat)), take(1, np.float64)[0],
                     int(take(1, np.int32)[0])))
    tau_f = take(3 * nat, np.float64, (3, nat)).T
    vion_f = take(3 * nat, np.float64, (3, nat)).T
    c0_f = take(ngw * n, np.complex128, (ngw, n))
    cm_f = take(ngw * n, np.complex128, (ngw, n))
    assert off == out.size

    ks = KSDevice(sysm, raggio=raggio, iesr=iesr)
    ctx = ks.ctx
    ctx.set_precision({"MIXED": "mixed", "FP64": "fp64"}[mode])
    bufs = {k: ctx.alloc_wfn(n) for k in ("c0", "cm", "cp", "c2")}
    ctx.upload(c0, bufs["c0"])
    st = ks.cp_init(bufs["c0"], bufs["cm"], bufs["c2"], sysm.f, sysm.tau, emass=emass, ion_mass=ION_MASS)
    ref = [(st["etot"], st["fion"].copy(), st["ekinc"], None)]
    for _ in range(nsteps):
        ks.cp_step(st, bufs, sysm.f, dt, epsog=1e-8, maxit=60)
        ref.append((st["etot"], st["fion"].copy(), st["ekinc"], st["iters"]))
    for (ef, ff, kf, itf), (er, fr, kr, itr) in zip(recs, ref):
        assert ef == pytest.approx(er, rel=1e-12)
        assert np.array_equal(ff, fr) and kf == kr and itf == itr
    assert np.array_equal(tau_f, st["tau"]) and np.array_equal(vion_f, st["vion"])
    assert np.array_equal(c0_f, ctx.download(bufs["c0"], n)) and np.array_equal(cm_f, ctx.download(bufs["cm"], n))
