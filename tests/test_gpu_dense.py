"""GPU parity of the FP64-MFMA contractions, orthonormalisation and projector
routines against the oracle (same seeded inputs).  Tolerance 1e-12 relative
to the largest entry unless stated."""
import math

import numpy as np
import pytest

import oracle
from cpmd_amd import ops, _lib
from cpmd_amd import system as S
from util import small_system, config_system

pytestmark = pytest.mark.gpu


def rel(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


@pytest.fixture(scope="module")
def cfg1_70():
    """cfg1 plane-wave set with 70 states (ragged 64-tiles)."""
    sysm = S.make_system(12.0, 8, 40.0, nstate=70)
    c = S.initial_coefficients(sysm, seed=21)
    rng = np.random.default_rng(22)
    d = (rng.standard_normal(c.shape) + 1j * rng.standard_normal(c.shape)) * np.exp(
        -sysm.gset.hg[:sysm.gset.ngw, None] * sysm.cell.tpiba2 / 8)
    d[0] = d[0].real
    cp = np.asfortranarray(c + 0.03 * d)                 # not orthonormal
    return sysm, c, cp


def test_ovlap(gpu, cfg1_70):
    sysm, c, cp = cfg1_70
    ctx = ops.Context(sysm.cell, sysm.gset)
    assert rel(ctx.ovlap(cp, cp), oracle.ovlap(cp, cp)) < 1e-13
    a = ctx.ovlap(cp, cp)
    assert np.array_equal(a, a.T)
    assert rel(ctx.ovlap(cp, c), oracle.ovlap(cp, c)) < 1e-13
    # 130 states: three row/column tiles, split-K over 7250 real rows
    sysm2 = S.make_system(12.0, 8, 40.0, nstate=130)
    c2 = S.initial_coefficients(sysm2, seed=3)
    assert np.abs(ctx.ovlap(c2, c2) - np.eye(130)).max() < 1e-13


def test_rgs(gpu, cfg1_70):
    sysm, c, cp = cfg1_70
    ctx = ops.Context(sysm.cell, sysm.gset)
    g, smat = ctx.rgs(cp)
    o, osmat = oracle.rgs(cp)
    assert rel(smat, osmat) < 1e-12
    assert rel(g, o) < 1e-12
    assert np.all(g[0].imag == 0)
    assert np.abs(S.dotp_gamma(g, g) - np.eye(70)).max() < 1e-13


def test_rgs_not_positive_definite(gpu, cfg1_70):
    sysm, c, cp = cfg1_70
    ctx = ops.Context(sysm.cell, sysm.gset)
    bad = np.asfortranarray(cp.copy())
    bad[:, 5] = 0.0                                       # a null state: S(6,6) = 0
    with pytest.raises(_lib.McError) as e:
        ctx.rgs(bad)
    assert e.value.code == 3 and "6" in str(e.value)


@pytest.mark.parametrize("n", [200, 300])
def test_rgs_blocked(gpu, n):
    """The blocked Cholesky beyond two blocks: 64-column diagonal blocks with
    a ragged last one (200 = 3 x 64 + 8, 300 = 4 x 64 + 44), several trsm
    workgroups and update tiles above and below the diagonal."""
    sysm = S.make_system(12.0, 8, 40.0, nstate=n)
    rng = np.random.default_rng(n)
    ngw = sysm.gset.ngw
    cp = (rng.standard_normal((ngw, n)) + 1j * rng.standard_normal((ngw, n))) * np.exp(
        -sysm.gset.hg[:ngw, None] * sysm.cell.tpiba2 / 8)
    cp[0] = cp[0].real
    cp = np.asfortranarray(cp)
    ctx = ops.Context(sysm.cell, sysm.gset)
    g, smat = ctx.rgs(cp)
    o, osmat = oracle.rgs(cp)
    assert rel(smat, osmat) < 1e-12
    assert rel(g, o) < 1e-12
    assert np.abs(S.dotp_gamma(g, g) - np.eye(n)).max() < 1e-12
    # a null state in the third block: the leading minor of order 151 fails,
    # reported by the diagonal step of block 2 (info path of a later block)
    bad = np.asfortranarray(cp.copy())
    bad[:, 150] = 0.0
    with pytest.raises(_lib.McError) as e:
        ctx.rgs(bad)
    assert e.value.code == 3 and "order 151 " in str(e.value)
    with pytest.raises(np.linalg.LinAlgError, match="leading minor 151 "):
        oracle.rgs(bad)


def test_lowdin(gpu, cfg1_70):
    sysm, c, cp = cfg1_70
    ctx = ops.Context(sysm.cell, sysm.gset)
    g = ctx.lowdin(cp)
    o, _ = oracle.lowdin(cp)
    assert rel(g, o) < 1e-12
    assert np.abs(S.dotp_gamma(g, g) - np.eye(70)).max() < 1e-13


def test_rotate_lagrange(gpu, cfg1_70):
    sysm, c, cp = cfg1_70
    ctx = ops.Context(sysm.cell, sysm.gset)
    rng = np.random.default_rng(4)
    gam = rng.standard_normal((70, 70))
    assert rel(ctx.rotate(0.5, c, -2.0, cp, gam), oracle.rotate(0.5, c, -2.0, cp, gam)) < 1e-13
    f = np.full(70, 2.0)
    f[[3, 40, 69]] = 0.0
    g2 = np.asfortranarray(cp - c)
    assert rel(ctx.lagrange(c, g2, f), oracle.lagrange(c, g2, f)) < 1e-12


def test_rortog(gpu, cfg1_70):
    sysm, c, cp = cfg1_70
    ctx = ops.Context(sysm.cell, sysm.gset)
    rng = np.random.default_rng(6)
    d = np.asfortranarray(cp - c)
    d = d - c @ S.dotp_gamma(c, d)
    m = rng.standard_normal((70, 70))
    cm = np.asfortranarray(c + 0.5 * d + 0.01 * c @ (m + m.T))
    for tnon in (False, True):
        out, gam, it = ctx.rortog(c, cm, tnon=tnon, epsog=1e-12, maxit=60)
        oout, ogam, oit = oracle.rortho(c, cm, tnon=tnon, epsog=1e-12, maxit=60)
        assert abs(it - oit) <= 1
        assert rel(gam, ogam) < 1e-10
        assert rel(out, oout) < 1e-11
    with pytest.raises(_lib.McError) as e:
        ctx.rortog(c, cm, epsog=1e-30, maxit=2)
    assert e.value.code == 4


def _two_species_projectors(sysm, ngw):
    """Two species (3 + 5 atoms), projectors s, s', p_x: exercises (-i)^l and h(iv,jv)."""
    rng = np.random.default_rng(8)
    g = sysm.gset.miller[:ngw] @ sysm.cell.bvec * sysm.cell.tpiba
    g2 = (g ** 2).sum(1)
    tw = np.zeros((ngw, 3, 2))
    tw[:, 0, 0] = np.exp(-0.2 * g2)
    tw[:, 1, 0] = g2 * np.exp(-0.25 * g2)
    tw[:, 2, 0] = g[:, 0] * np.exp(-0.3 * g2)
    tw[:, 0, 1] = 2.0 * np.exp(-0.15 * g2)
    tw[:, 1, 1] = g[:, 1] * np.exp(-0.2 * g2)
    lang = np.array([[0, 0, 1], [0, 1, 0]]).T                       # (nhxs, nsp)
    h = np.zeros((3, 3, 2))
    h[:2, :2, 0] = [[1.5, -0.3], [-0.3, 0.8]]
    h[2, 2, 0] = 0.6
    h[0, 0, 1] = 2.0
    h[1, 1, 1] = -0.4
    tau = rng.uniform(0, 12.0, size=(8, 3))
    frac = tau @ np.linalg.inv(sysm.cell.avec)
    eigr = np.exp(-2j * math.pi * (sysm.gset.miller[:ngw] @ frac.T))
    return dict(na=[3, 5], nhxs=3, ngh=[3, 2], lang=lang, hmat=h, twnl=tw, eigr=eigr)


@pytest.mark.parametrize("kind", ["cfg1", "two_species"])
def test_rnlsm1_fnonloc(gpu, kind):
    sysm, c = config_system("cfg1")
    ngw = sysm.gset.ngw
    if kind == "cfg1":
        na, nhxs, ngh, lang, hmat, twnl, eigr = S.projector_set(sysm, sysm.gset.miller[:ngw], sysm.gset.hg[:ngw])
        P = dict(na=na, nhxs=nhxs, ngh=ngh, lang=lang, hmat=hmat, twnl=twnl, eigr=eigr)
    else:
        P = _two_species_projectors(sysm, ngw)
    ctx = ops.Context(sysm.cell, sysm.gset)
    ctx.set_projectors(P["na"], P["nhxs"], P["ngh"], P["lang"], P["hmat"], P["twnl"])
    ctx.set_eigr(P["eigr"])
    OP = oracle.Projectors(**P)
    fnl = ctx.rnlsm1(c)
    ofnl = oracle.rnlsm1(OP, c)
    assert rel(fnl, ofnl) < 1e-12
    f = np.full(c.shape[1], 2.0)
    f[1] = 0.0
    rng = np.random.default_rng(1)
    c2 = np.asfortranarray(rng.standard_normal(c.shape) + 1j * rng.standard_normal(c.shape))
    out = ctx.fnonloc(c2, f)                              # device copy of fnl from rnlsm1
    oout = oracle.fnonloc(OP, c2, f, ofnl)
    assert rel(out - c2, oout - c2) < 1e-12
    out2 = ctx.fnonloc(c2, f, fnl)                        # host fnl
    assert np.array_equal(out, out2)


def test_device_resident_ortho_cfg4(gpu):
    """512 states over 198,204 plane waves: rgs and lowdin on the device make
    the perturbed set orthonormal; Lagrange projection removes the span."""
    sysm, c = config_system("cfg4")
    ctx = ops.Context(sysm.cell, sysm.gset)
    n = c.shape[1]
    rng = np.random.default_rng(2)
    d = ctx.upload(c)
    s = ctx.ovlap_dev(d, d, n)
    assert np.abs(s - np.eye(n)).max() < 1e-12
    # perturb on the device: cp = c + 0.01 c M  (rotate), M random
    m = 0.01 * rng.standard_normal((n, n))
    cp = ctx.alloc_wfn(n)
    ctx.rotate_dev(1.0, d, 0.0, cp, np.eye(n) + m, n)
    ctx.rgs_dev(cp, n)
    assert np.abs(ctx.ovlap_dev(cp, cp, n) - np.eye(n)).max() < 1e-12
    ctx.rotate_dev(1.0, d, 0.0, cp, np.eye(n) + m, n)
    it = ctx.lowdin_dev(cp, n)
    assert 1 <= it <= 100
    assert np.abs(ctx.ovlap_dev(cp, cp, n) - np.eye(n)).max() < 1e-12
    # Loewdin of C (I + M) with C orthonormal is C (I+M) ((I+M)^T (I+M))^-1/2
    a = np.eye(n) + m
    w, z = np.linalg.eigh(a.T @ a)
    ref = a @ z @ np.diag(w ** -0.5) @ z.T
    sub = ctx.download(cp, n)[:2000]
    assert np.abs(sub - c[:2000] @ ref).max() < 1e-12


@pytest.mark.parametrize("kind", ["cfg1", "two_species"])
def test_rnlsm2_rnlfor(gpu, kind):
    """dfnl and the nonlocal forces against the oracle (itself pinned by finite
    differences of fnl and E_NL), host and device-resident entry points."""
    sysm, c = config_system("cfg1")
    ngw = sysm.gset.ngw
    if kind == "cfg1":
        na, nhxs, ngh, lang, hmat, twnl, eigr = S.projector_set(sysm, sysm.gset.miller[:ngw], sysm.gset.hg[:ngw])
        P = dict(na=na, nhxs=nhxs, ngh=ngh, lang=lang, hmat=hmat, twnl=twnl, eigr=eigr)
    else:
        P = _two_species_projectors(sysm, ngw)
    gk = (sysm.gset.miller[:ngw] @ sysm.cell.bvec).T
    ctx = ops.Context(sysm.cell, sysm.gset)
    ctx.set_projectors(P["na"], P["nhxs"], P["ngh"], P["lang"], P["hmat"], P["twnl"])
    ctx.set_eigr(P["eigr"])
    ctx.set_gk(gk)
    OP = oracle.Projectors(**P)
    dfnl = ctx.rnlsm2(c)
    odfnl = oracle.rnlsm2(OP, c, gk, sysm.cell.tpiba)
    assert rel(dfnl, odfnl) < 1e-12
    f = np.full(c.shape[1], 2.0)
    f[2] = 0.0
    ofnl = oracle.rnlsm1(OP, c)
    ofion = oracle.rnlfor(OP, ofnl, odfnl, f)
    assert rel(ctx.rnlfor(f, ofnl, odfnl), ofion) < 1e-12          # host fnl / dfnl
    d = ctx.upload(c)
    n = c.shape[1]
    ctx.rnlsm1_dev(d, n)
    ctx.rnlsm2_dev(d, n)
    base = np.ones((3, OP.nat))
    got = ctx.rnlfor(f, fion=base)                                   # device copies, accumulates
    assert rel(got - base, ofion) < 1e-12
    oenl = oracle.enl(OP, ofnl, f)                                   # E_NL from the device fnl
    assert abs(ctx.rnlrh(f) - oenl) < 1e-12 * abs(oenl)
    with pytest.raises(_lib.McError):
        ctx.rnlrh(f[:-1])                                            # no fnl for that many states


def test_cp_posupa_velocity(gpu):
    """velupa + posupa (cm += dtm c2, cp = c0 + dt cm) and cm = (cp - c0)/dt
    against numpy on the same coefficients (two roundings per element)."""
    cell, gs, c = small_system("ortho", 5)
    ctx = ops.Context(cell, gs)
    rng = np.random.default_rng(17)
    shape = c.shape
    c2 = np.asfortranarray(rng.standard_normal(shape) + 1j * rng.standard_normal(shape))
    cm = np.asfortranarray(rng.standard_normal(shape) + 1j * rng.standard_normal(shape))
    n = shape[1]
    d0, d2, dm = ctx.upload(c), ctx.upload(c2), ctx.upload(cm)
    dp = ctx.alloc_wfn(n)
    dtm, dt = 0.0125, 4.0
    ctx.cp_posupa(d0, d2, dm, dp, n, dtm, dt)
    cm_ref = cm + dtm * c2
    cp_ref = c + dt * cm_ref
    assert np.abs(ctx.download(dm, n) - cm_ref).max() < 1e-14 * np.abs(cm_ref).max()
    assert np.abs(ctx.download(dp, n) - cp_ref).max() < 1e-14 * np.abs(cp_ref).max()
    ctx.cp_velocity(d0, dp, dm, n, dt)
    v_ref = (cp_ref - c) / dt
    assert np.abs(ctx.download(dm, n) - v_ref).max() < 1e-13 * np.abs(v_ref).max()
    with pytest.raises(_lib.McError):
        ctx.cp_posupa(d0, d2, dm, d0, n, dtm, dt)      # cp aliasing c0 is refused
