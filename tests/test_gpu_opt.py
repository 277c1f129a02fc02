"""Wavefunction optimisation on the device: hesele, csize, the
preconditioned steepest-descent update and ODIIS against the oracle, and the precision
policy (FP32 transforms / split-bf16 contractions while the gradient size
gemax >= 1e-4, FP64 after) against a plain FP64 optimisation from the same
start: the same converged E_KS (|dE| < 1e-6 Ha/atom) at CPMD's default
convergence gemax < 1e-5."""
import numpy as np
import pytest

import oracle
from cpmd_amd import system as S
from cpmd_amd.ks import KSDevice

from test_oracle_energy import ks_system

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def small(gpu):
    sysm, K, c = ks_system(a=9.0, nat=3, ecut=25.0, nstate=6)
    ks = KSDevice(sysm)
    return sysm, K, c, ks


def _projectors(sysm, K):
    return K.projectors(sysm.tau)


@pytest.mark.parametrize("prec", [1, 0])
def test_sd_update_parity(small, prec):
    sysm, K, c, ks = small
    ctx = ks.ctx
    gs = sysm.gset
    ngw = gs.ngw
    rng = np.random.default_rng(11)
    n = c.shape[1]
    c2 = np.asfortranarray(rng.standard_normal((ngw, n)) + 1j * rng.standard_normal((ngw, n)))
    c2[0] = c2[0].real
    svar2, hthrs, pf1 = 25.0 / 400.0, 0.5, 0.3
    ctx.hesele(bool(prec), svar2, ks.vploc, hthrs)
    d0, d2 = ctx.upload(c), ctx.upload(c2)
    ctx.sd_update_dev(d0, d2, n, pf1)
    got = ctx.download(d0, n)
    vpp = oracle.hesele(prec, svar2, ks.vploc, hthrs, sysm.cell.tpiba2, gs.hg[:ngw], _projectors(sysm, K))
    want = oracle.sd_update(c, c2, vpp, pf1)
    assert np.abs(got - want).max() <= 1e-13 * np.abs(want).max()
    assert np.abs(got - c).max() > 1e-3          # the update did something


def test_csize_parity(small):
    sysm, K, c, ks = small
    ctx = ks.ctx
    ngw = sysm.gset.ngw
    rng = np.random.default_rng(12)
    n = 6
    c2 = np.asfortranarray(rng.standard_normal((ngw, n)) + 1j * rng.standard_normal((ngw, n)))
    c2[0] = c2[0].real
    # izamax picks the largest |Re| + |Im|, not the largest modulus
    c2[7, 2] = 3.0 + 3.0j          # |.|_1 = 6, modulus 4.24
    c2[9, 1] = 5.0 + 0.0j          # |.|_1 = 5, modulus 5
    c2[4, 5] = 50.0                # state 5 is empty: ignored
    f = np.array([2.0, 2.0, 2.0, 2.0, 2.0, 0.0])
    d2 = ctx.upload(c2)
    ge, cn = ctx.csize_dev(d2, f, n)
    ge_o, cn_o = oracle.csize(c2, f)
    assert ge == pytest.approx(3.0 * np.sqrt(2.0), rel=1e-15)
    assert ge == ge_o
    assert cn == pytest.approx(cn_o, rel=1e-13)


def _optimize(ks, c_start, d, f, mixed, tol):
    ctx = ks.ctx
    n = len(f)
    ctx.upload(c_start, d["c0"])
    r = ks.optimize(d["c0"], f, d["c2"], tol=tol, switch=1e-4, mixed=mixed, dt=20.0, hthrs=0.5,
                    maxiter=2500, method="sd")
    ctx.set_precision("fp64")
    c = ctx.download(d["c0"], n)
    return r, c


def test_optimize_mixed_then_fp64(gpu):
    # 16 atoms in a 16 bohr cell at 70 Ry: 2 ngw >= 16384, so the split-bf16
    # contractions and FP32 transforms are active in the mixed phase
    sysm = S.make_system(16.0, 16, 70.0, nstate=32, with_density=True)
    assert 2 * sysm.gset.ngw >= 16384
    ks = KSDevice(sysm)
    ctx = ks.ctx
    f = sysm.f
    n = len(f)
    d = {k: ctx.alloc_wfn(n) for k in ("c0", "c2")}
    ks.set_ions(sysm.tau)
    c0 = S.initial_coefficients(sysm)
    tol = 1e-5                 # CPMD's default orbital convergence (src/control_test_utils.mod.F90:595-601)
    r64, c64 = _optimize(ks, c0, d, f, False, tol)
    rmx, cmx = _optimize(ks, c0, d, f, True, tol)
    print(f"fp64: {r64['iters_fp64']} iterations, E {r64['etot']:.12f}; mixed->fp64: {rmx['iters_mixed']} + "
          f"{rmx['iters_fp64']} iterations, E {rmx['etot']:.12f}", flush=True)
    assert r64["converged"] and rmx["converged"]
    assert r64["iters_mixed"] == 0
    assert rmx["iters_mixed"] > 0 and rmx["iters_fp64"] > 0
    modes = [h[0] for h in rmx["history"]]
    k = modes.index("fp64")
    assert all(m == "mixed" for m in modes[:k]) and all(m == "fp64" for m in modes[k:])
    assert rmx["history"][k - 1][2] < 1e-4 <= min(h[2] for h in rmx["history"][:k - 1])
    nat = sysm.nat
    assert abs(rmx["etot"] - r64["etot"]) / nat < 1e-6
    # both end states are orthonormal in FP64
    for c in (c64, cmx):
        s = S.dotp_gamma(c, c)
        assert np.abs(s - np.eye(n)).max() < 1e-10
    # the energy went down from the start
    assert rmx["history"][0][1] > rmx["etot"]


def test_odiis_parity(small):
    """Seven ODIIS steps with a 3-deep history (the ring wraps, and the
    progress reset fires: the gradients do not shrink), an empty state and an
    energy rise (damping), device against the oracle on the same inputs."""
    sysm, K, c, ks = small
    ctx = ks.ctx
    gs = sysm.gset
    ngw = gs.ngw
    n = c.shape[1]
    f = np.array([2.0, 2.0, 2.0, 2.0, 2.0, 0.0])
    hthrs, svar2 = 0.5, 25.0 / 400.0
    ctx.hesele(True, 0.0, ks.vploc, hthrs)
    vpp = oracle.hesele(1, 0.0, ks.vploc, hthrs, sysm.cell.tpiba2, gs.hg[:ngw], _projectors(sysm, K))
    D = oracle.ODIIS(mdiis=3)
    rng = np.random.default_rng(21)
    cur = c.copy()
    d0, d2 = ctx.alloc_wfn(n), ctx.alloc_wfn(n)
    resets = []
    for k, e in enumerate([10.0, 9.0, 9.5, 8.0, 7.0, 6.5, 6.0]):
        c2 = np.asfortranarray(0.05 * (rng.standard_normal((ngw, n)) + 1j * rng.standard_normal((ngw, n))))
        c2[0] = c2[0].real
        ctx.upload(cur, d0)
        ctx.upload(c2, d2)
        reset = ctx.odiis_dev(d0, d2, f, n, e, svar2, mdiis=3, reinit=(k == 0))
        got = ctx.download(d0, n)
        want = D.step(cur, c2, vpp, f, e, svar2)
        err = np.abs(got - want).max() / np.abs(want).max()
        assert err < 1e-11, (k, err)
        assert reset == D.reset, k       # the reset is reported where the oracle's fires
        resets.append(reset)
        cur = want
    assert D.ndiis < 7 and any(resets)   # the history was restarted on the way


def test_optimize_diis_mixed_then_fp64(gpu):
    # 8 atoms in a 12 bohr cell at 70 Ry (2 ngw >= 16384: split-bf16 products in use)
    sysm = S.make_system(12.0, 8, 70.0, nstate=16, seed=5, with_density=True)
    assert 2 * sysm.gset.ngw >= 16384
    ks = KSDevice(sysm)
    ctx = ks.ctx
    f = sysm.f
    n = len(f)
    d = {k: ctx.alloc_wfn(n) for k in ("c0", "c2")}
    ks.set_ions(sysm.tau)
    c0 = S.initial_coefficients(sysm, seed=5)
    out = {}
    for mixed in (False, True):
        ctx.upload(c0, d["c0"])
        out[mixed] = ks.optimize(d["c0"], f, d["c2"], tol=1e-6, mixed=mixed, method="diis", maxiter=400)
    r64, rmx = out[False], out[True]
    print(f"diis fp64: {r64['iters_fp64']} iterations, E {r64['etot']:.12f}; mixed->fp64: "
          f"{rmx['iters_mixed']} + {rmx['iters_fp64']}, E {rmx['etot']:.12f}", flush=True)
    assert r64["converged"] and rmx["converged"]
    assert rmx["iters_mixed"] > 0 and rmx["iters_fp64"] > 0
    assert abs(rmx["etot"] - r64["etot"]) / sysm.nat < 1e-6
    # ODIIS with its progress reset converges in a few hundred iterations at most
    assert r64["iters_fp64"] < 300 and rmx["iters_mixed"] + rmx["iters_fp64"] < 300
