"""GPU checks of the mixed-precision contractions (MC_PREC_MIXED): 3-term
bf16 split MFMA products with FP32 accumulation per K chunk.

Error model.  hi = bf16(x) and lo = bf16(x - hi) leave |x - hi| <= 2^-8 |x| and
|d| = |x - hi - lo| <= 2^-16 |x|; with the dropped lo*lo term one product
carries a relative error of at most 3 * 2^-16; the FP32 accumulation of a K
chunk of length L adds at most L * 2^-24 of the sum of |terms|.  So every entry
obeys
    |out_mixed - out_fp64| <= (3 * 2^-16 + L * 2^-24) * sum_k |a_k| |b_k|
(L <= 2048 for the split-K products, L = n for the n-column rotations).  The
tests check that bound entrywise against the oracle, and tighter typical-error
bounds for near-orthonormal sets.  The diagonal of a square overlap is
recomputed in FP64 (its terms are correlated: the dropped lo^2 and the FP32
accumulation would bias it by ~1e-6), so it matches FP64 to rounding."""
import math

import numpy as np
import pytest

import oracle
from cpmd_amd import ops
from cpmd_amd import system as S

pytestmark = pytest.mark.gpu

U_SPLIT = 3 * 2.0 ** -16
U32 = 2.0 ** -24


def real_rows(c):
    """(2*ngw x n) real view of the half-sphere coefficients."""
    c = np.asarray(c)
    out = np.empty((2 * c.shape[0], c.shape[1]))
    out[0::2] = c.real
    out[1::2] = c.imag
    return out


@pytest.fixture(scope="module")
def cfg2_130():
    """cfg2 plane-wave set (68,338 PW, 136,676 real rows) with 130 states."""
    sysm = S.make_system(24.0, 64, 70.0, nstate=130)
    c = S.initial_coefficients(sysm, seed=31)
    rng = np.random.default_rng(32)
    d = (rng.standard_normal(c.shape) + 1j * rng.standard_normal(c.shape)) * np.exp(
        -sysm.gset.hg[:sysm.gset.ngw, None] * sysm.cell.tpiba2 / 8)
    d[0] = d[0].real
    cp = np.asfortranarray(c + 0.02 * d / np.sqrt(np.sum(np.abs(d) ** 2, axis=0)))
    return sysm, c, cp


def test_mixed_ovlap(gpu, cfg2_130):
    sysm, c, cp = cfg2_130
    ctx = ops.Context(sysm.cell, sysm.gset)
    ctx.set_precision("mixed")
    for a, b in ((cp, cp), (cp, c)):
        got = ctx.ovlap(a, b)
        ref = oracle.ovlap(a, b)
        bound = 2.0 * np.abs(real_rows(a)).T @ np.abs(real_rows(b))
        err = np.abs(got - ref)
        assert np.all(err <= (U_SPLIT + 2048 * U32) * bound + 1e-300)
        dg = np.abs(np.diag(got - ref))
        off = err - np.diag(dg)
        msg = f"diag err max {dg.max():.3e} off-diagonal err max {off.max():.3e}"
        # six-term split, FP32 chunk sums: off-diagonal errors at the 1e-8
        # level; on the diagonal (sums of squares, entries ~1) the MFMA's
        # alignment of each product to the running sum costs a few 1e-8
        assert off.max() < 3e-8, msg
        assert dg.max() < 1e-7, msg
        assert off.max() > 0.0                         # the split kernels ran
    s = ctx.ovlap(cp, cp)
    assert np.array_equal(s, s.T)
    ctx.set_precision("fp64")
    assert np.abs(ctx.ovlap(cp, c) - oracle.ovlap(cp, c)).max() < 1e-13


def test_mixed_rotate(gpu, cfg2_130):
    sysm, c, cp = cfg2_130
    ctx = ops.Context(sysm.cell, sysm.gset)
    ctx.set_precision("mixed")
    rng = np.random.default_rng(5)
    n = c.shape[1]
    gam = rng.standard_normal((n, n))
    got = ctx.rotate(0.5, c, -2.0, cp, gam)
    ref = oracle.rotate(0.5, c, -2.0, cp, gam)
    bound = 0.5 * np.abs(real_rows(c)) @ np.abs(gam)
    err = np.abs(real_rows(got) - real_rows(ref))
    assert np.all(err <= (U_SPLIT + n * U32) * bound + 1e-15 * np.abs(real_rows(ref)))
    assert err.max() > 0.0


def test_mixed_lowdin_rgs(gpu, cfg2_130):
    """Orthonormality after the mixed-precision factorisations, measured in FP64."""
    sysm, c, cp = cfg2_130
    ctx = ops.Context(sysm.cell, sysm.gset)
    n = c.shape[1]
    ref = oracle.lowdin(cp)[0]
    ctx.set_precision("mixed")
    g = ctx.lowdin(cp)
    r = ctx.rgs(cp)[0]
    for out in (g, r):
        assert np.abs(S.dotp_gamma(out, out) - np.eye(n)).max() < 1e-6
    # C (I + E) with E = S^-1/2 - I: the product error scales with |E| ~ 0.02
    assert np.abs(g - ref).max() < 1e-6 * np.abs(ref).max()


def test_mixed_projectors(gpu):
    sysm = S.make_system(24.0, 64, 70.0, nstate=64)
    c = S.initial_coefficients(sysm, seed=41)
    ngw = sysm.gset.ngw
    na, nhxs, ngh, lang, hmat, twnl, eigr = S.projector_set(sysm, sysm.gset.miller[:ngw], sysm.gset.hg[:ngw])
    P = dict(na=na, nhxs=nhxs, ngh=ngh, lang=lang, hmat=hmat, twnl=twnl, eigr=eigr)
    OP = oracle.Projectors(**P)
    ctx = ops.Context(sysm.cell, sysm.gset)
    ctx.set_projectors(na, nhxs, ngh, lang, hmat, twnl)
    ctx.set_eigr(eigr)
    ctx.set_precision("mixed")
    fnl = ctx.rnlsm1(c)
    ofnl = oracle.rnlsm1(OP, c)
    # |Re/Im of (-i)^l twnl eigr| <= |twnl|: sum_k |a||b| <= 2 sum_G |twnl| (|Re c| + |Im c|)
    tw = np.abs(twnl[:, 0, 0])
    b = 2.0 * (tw[:, None] * (np.abs(c.real) + np.abs(c.imag))).sum(0)        # (nstate,)
    err = np.abs(fnl - ofnl)
    assert np.all(err <= (U_SPLIT + 2048 * U32) * b[None, None, :] * math.sqrt(2) + 1e-300)
    assert err.max() > 0.0
    # projector operands are split in three bf16 parts: FP32-level products
    # (a 3-term split leaves 4e-6 rms relative error on these random sums)
    assert np.sqrt(np.mean(err ** 2) / np.mean(ofnl ** 2)) < 5e-7
    f = np.full(c.shape[1], 2.0)
    f[3] = 0.0
    rng = np.random.default_rng(7)
    c2 = np.asfortranarray(rng.standard_normal(c.shape) + 1j * rng.standard_normal(c.shape))
    out = ctx.fnonloc(c2, f, ofnl)
    oout = oracle.fnonloc(OP, c2, f, ofnl)
    # term bound: f * |twnl(G)| * sum_ia |dd(ia, i)|, dd = h fnl; the (re, im)
    # epilogue mixes the two components of the product (factor 2)
    h = float(hmat[0, 0, 0])
    fe = np.where(f < 1e-5, 1.0, f)
    bnd = (fe[None, :] * tw[:, None] * (abs(h) * np.abs(ofnl[:, 0, :]).sum(0))[None, :]) * 2.0
    nat = int(np.sum(na))
    for part in (np.real, np.imag):
        err2 = np.abs(part(out) - part(oout))
        assert np.all(err2 <= (U_SPLIT + nat * U32) * bnd + 1e-15 * np.abs(oout).max())


def test_mixed_rnlsm_combined(gpu):
    """rnlsm (fnl + dfnl in one call: one stacked contraction of the four
    projector variants over the bf16 planes of c0) equals the separate rnlsm1 /
    rnlsm2 calls exactly in FP64 and to rounding in mixed precision (the
    stacking changes the K chunking of the FP32 sums), and the mixed fnl and
    dfnl stay at FP32-level error against the oracle."""
    sysm = S.make_system(24.0, 64, 70.0, nstate=64)
    c = S.initial_coefficients(sysm, seed=43)
    ngw = sysm.gset.ngw
    na, nhxs, ngh, lang, hmat, twnl, eigr = S.projector_set(sysm, sysm.gset.miller[:ngw], sysm.gset.hg[:ngw])
    OP = oracle.Projectors(na=na, nhxs=nhxs, ngh=ngh, lang=lang, hmat=hmat, twnl=twnl, eigr=eigr)
    gk = (sysm.gset.miller[:ngw] @ sysm.cell.bvec).T
    ctx = ops.Context(sysm.cell, sysm.gset)
    ctx.set_projectors(na, nhxs, ngh, lang, hmat, twnl)
    ctx.set_eigr(eigr)
    ctx.set_gk(gk)
    n = c.shape[1]
    d = ctx.upload(c)
    for mode in ("fp64", "mixed"):
        ctx.set_precision(mode)
        fnl, dfnl = ctx.rnlsm_dev(d, n, tfor=True, want_fnl=True, want_dfnl=True)
        fnl1 = ctx.rnlsm1_dev(d, n, want_fnl=True)
        dfnl1 = ctx.rnlsm2_dev(d, n, want_dfnl=True)
        if mode == "fp64":
            assert np.array_equal(fnl, fnl1) and np.array_equal(dfnl, dfnl1)
        else:
            for a, b in ((fnl, fnl1), (dfnl, dfnl1)):
                assert np.sqrt(np.mean((a - b) ** 2) / np.mean(b ** 2)) < 3e-7
        fnl0, _ = ctx.rnlsm_dev(d, n, tfor=False, want_fnl=True)
        assert np.array_equal(fnl0, fnl1)
    ofnl = oracle.rnlsm1(OP, c)
    assert np.sqrt(np.mean((fnl - ofnl) ** 2) / np.mean(ofnl ** 2)) < 5e-7
    odfnl = oracle.rnlsm2(OP, c, gk, sysm.cell.tpiba)
    rel = np.sqrt(np.mean((dfnl - odfnl) ** 2) / np.mean(odfnl ** 2))
    assert rel < 5e-7, rel


def test_posupa_combined(gpu, cfg2_130):
    """mc_posupa_dev (velupa kick + posupa + rortog + velocities in one call,
    the bf16 planes of c0 and cp built by its first pass in mixed precision)
    equals the separate velupa/posupa, rortog and velocity calls exactly in
    FP64 and in mixed precision."""
    sysm, c, cp = cfg2_130
    ctx = ops.Context(sysm.cell, sysm.gset)
    n = c.shape[1]
    rng = np.random.default_rng(61)
    c2 = np.asfortranarray(rng.standard_normal(c.shape) + 1j * rng.standard_normal(c.shape)) * 1e-3
    c2[0] = c2[0].real
    cm = np.asfortranarray((cp - c) / 0.5)
    for mode in ("fp64", "mixed"):
        ctx.set_precision(mode)
        d0, d2 = ctx.upload(c), ctx.upload(c2)
        dm_a, dm_b = ctx.upload(cm), ctx.upload(cm)
        dp_a, dp_b = ctx.alloc_wfn(n), ctx.alloc_wfn(n)
        gam_a, it_a = ctx.posupa_dev(d0, d2, dm_a, dp_a, n, 0.01, 0.5, want_gam=True)
        ctx.cp_posupa(d0, d2, dm_b, dp_b, n, 0.01, 0.5)
        gam_b, it_b = ctx.rortog_dev(d0, dp_b, n, epsog=1e-8, maxit=60)
        ctx.cp_velocity(d0, dp_b, dm_b, n, 0.5)
        assert it_a == it_b and np.array_equal(gam_a, gam_b), mode
        pa, pb = ctx.download(dp_a, n), ctx.download(dp_b, n)
        assert np.array_equal(pa, pb) and np.array_equal(ctx.download(dm_a, n), ctx.download(dm_b, n)), mode


def test_posupa_combined_odd_no_gradient(gpu, cfg2_130):
    """An odd number of states (the last column block part empty) and no
    gradient (c2 = NULL: no velupa kick) against the separate calls with a
    zero gradient, in both precisions."""
    sysm, c, cp = cfg2_130
    c, cp = np.asfortranarray(c[:, :129]), np.asfortranarray(cp[:, :129])
    ctx = ops.Context(sysm.cell, sysm.gset)
    n = c.shape[1]
    cm = np.asfortranarray((cp - c) / 0.5)
    for mode in ("fp64", "mixed"):
        ctx.set_precision(mode)
        d0, dz = ctx.upload(c), ctx.alloc_wfn(n)
        dm_a, dm_b = ctx.upload(cm), ctx.upload(cm)
        dp_a, dp_b = ctx.alloc_wfn(n), ctx.alloc_wfn(n)
        gam_a, it_a = ctx.posupa_dev(d0, None, dm_a, dp_a, n, 0.01, 0.5, want_gam=True)
        ctx.cp_posupa(d0, dz, dm_b, dp_b, n, 0.01, 0.5)
        gam_b, it_b = ctx.rortog_dev(d0, dp_b, n, epsog=1e-8, maxit=60)
        ctx.cp_velocity(d0, dp_b, dm_b, n, 0.5)
        assert it_a == it_b and np.array_equal(gam_a, gam_b), mode
        assert np.array_equal(ctx.download(dp_a, n), ctx.download(dp_b, n)), mode
        assert np.array_equal(ctx.download(dm_a, n), ctx.download(dm_b, n)), mode


@pytest.mark.parametrize("nstate", [64, 63])
def test_forces_c2_combined(gpu, nstate):
    """forces_c2 (vpsi writing c2 instead of adding to it, then fnonloc with the
    device fnl) equals c2 <- 0, vpsi, fnonloc bit for bit in both precisions,
    whatever c2 held before, padding rows included."""
    sysm = S.make_system(24.0, 64, 70.0, nstate=nstate)
    c = S.initial_coefficients(sysm, seed=47)
    ngw = sysm.gset.ngw
    na, nhxs, ngh, lang, hmat, twnl, eigr = S.projector_set(sysm, sysm.gset.miller[:ngw], sysm.gset.hg[:ngw])
    ctx = ops.Context(sysm.cell, sysm.gset)
    ctx.set_projectors(na, nhxs, ngh, lang, hmat, twnl)
    ctx.set_eigr(eigr)
    n = c.shape[1]
    f = np.full(n, 2.0)
    f[5] = 0.0
    d = ctx.upload(c)
    v = S.real_space_potential(sysm)
    vd = ops.DeviceArray(8 * ctx.nnr1).from_host(v)
    nd = 2 * ctx.ld * n
    rng = np.random.default_rng(11)
    for mode in ("fp64", "mixed"):
        ctx.set_precision(mode)
        ctx.rnlsm_dev(d, n, tfor=False)
        ref = ctx.alloc_wfn(n)
        ctx.zero_wfn(ref, n)
        ctx.vpsi_dev(d, ref, f, vd, n)
        ctx.fnonloc_dev(ref, f, n)
        out = ctx.alloc_wfn(n).from_host(rng.standard_normal(nd))
        ctx.forces_c2_dev(d, out, f, vd, n)
        ctx.synchronize()          # to_host copies on the null stream
        a, b = out.to_host(nd), ref.to_host(nd)
        assert np.array_equal(a, b)
        assert np.abs(b).max() > 0.0


def test_cfg4_full_size_mixed_against_fp64(gpu):
    """Config 4 at full size (512 states, 192^3, 256 atoms): from the seeded
    coefficients after 5 FP64 steepest-descent steps, E_KS and F_ion of the
    mixed path against FP64 at the same C and R (|dE_KS| < 1e-6 Ha/atom,
    max |dF_ion| < 1e-5 Ha/bohr), and ||C^T C - I|| < 1e-6 after one mixed
    mc_posupa_dev (velupa + posupa + rortog, dt 1 a.u.)."""
    from cpmd_amd.ks import KSDevice
    cfg = S.CONFIGS["cfg4"]
    sysm = S.make_system(cfg["a"], cfg["nat"], cfg["ecut"], nstate=cfg["nstate"], with_density=True)
    n, nat, f = sysm.nstate, sysm.nat, sysm.f
    ks = KSDevice(sysm)
    ctx = ks.ctx
    buf = {k: ctx.alloc_wfn(n) for k in ("c0", "cm", "cp", "c2")}
    ctx.set_precision("fp64")
    ctx.upload(S.initial_coefficients(sysm), buf["c0"])
    ks.set_ions(sysm.tau)
    for _ in range(5):
        ks.energy(buf["c0"], f, gradient=buf["c2"])
        ctx.lagrange_dev(buf["c0"], buf["c2"], f, n)
        ctx.axpy_wfn(0.02, buf["c2"], buf["c0"], n)
        ctx.lowdin_dev(buf["c0"], n)
    print("cfg4: FP64 and mixed E_KS / F_ion at the same state", flush=True)
    e64 = ks.energy(buf["c0"], f, tau=sysm.tau, forces=True)
    ctx.set_precision("mixed")
    emx = ks.energy(buf["c0"], f, tau=sysm.tau, gradient=buf["c2"], forces=True)
    assert emx["etot"] != e64["etot"]
    assert abs(emx["etot"] - e64["etot"]) / nat < 1e-6
    assert np.abs(emx["fion"] - e64["fion"]).max() < 1e-5
    ctx.axpby_wfn(0.0, buf["c0"], 0.0, buf["cm"], n)
    _, it = ctx.posupa_dev(buf["c0"], buf["c2"], buf["cm"], buf["cp"], n, 0.5 / 400.0, 1.0, epsog=1e-8, maxit=60)
    ctx.set_precision("fp64")
    ov = ctx.ovlap_dev(buf["cp"], buf["cp"], n)
    assert np.abs(ov - np.eye(n)).max() < 1e-6



def test_fp32_charge_loss_is_the_engine_contraction(gpu):
    """The FP32 rhoofr's real-space charge before its rescale, against the
    first-order prediction from the FP32 engine's per-frequency norm
    contraction (oracle/fp32_engine.py) for config 2's 128 states on 128^3:
    the measured loss is that deterministic contraction (within 5%), and
    after the rescale the returned csumr equals csumg."""
    from oracle import fp32_engine as E
    from util import config_system
    sysm, c = config_system("cfg2")
    ctx = ops.Context(sysm.cell, sysm.gset)
    ctx.set_precision("fp32")
    rho, ek, cg, cr = ctx.rhoofr(c, sysm.f)
    g, cr_raw = ctx.rhoofr_raw_charge()
    assert g == cg and cr == pytest.approx(cg, rel=1e-14)
    measured = (cr_raw - cg) / cg
    predicted = E.predicted_charge_loss(sysm.cell, sysm.gset.miller, c, sysm.f)
    print(f"FP32 charge loss: measured {measured:.4e}, predicted {predicted:.4e}", flush=True)
    assert predicted < -5e-8
    assert measured == pytest.approx(predicted, rel=0.05)
    ctx.set_precision("fp64")
    ctx.rhoofr(c, sysm.f)
    g64, cr64 = ctx.rhoofr_raw_charge()
    assert cr64 == pytest.approx(g64, rel=1e-13)          # FP64: nothing to rescale
