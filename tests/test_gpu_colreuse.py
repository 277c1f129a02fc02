"""Column reuse (mc_ctx_set_column_reuse): vpsi_dev / forces_c2_dev right after
rhoofr_dev on the same device wavefunctions start from rhoofr's y-transformed
columns.  The results must be bit-identical to the unshared transforms, the
inverse z and y passes must actually be skipped (kernel launch counts), and
every change that makes the kept columns stale must fall back to a full
transform: a library write to the wavefunctions, another wavefunction
buffer, another precision, a device memory limit.
"""
import numpy as np
import pytest

from cpmd_amd import ops
from cpmd_amd import system as S
from util import small_system

pytestmark = pytest.mark.gpu


def _setup(kind, nstate):
    if kind == "mesh128":   # fused 16 x 8 engine, pruned x / y passes
        cell = S.cubic_cell(19.5, 100.0)
        assert tuple(cell.nr) == (128, 128, 128)
        gs = S.gvectors(cell, with_density=False)
        sysm = S.SyntheticSystem(cell=cell, gset=gs, tau=np.zeros((1, 3)), nstate=nstate,
                                 f=np.full(nstate, 2.0))
        c = S.initial_coefficients(sysm, seed=17, nstate=nstate)
    else:
        cell, gs, c = small_system(kind, nstate)
        sysm = S.SyntheticSystem(cell=cell, gset=gs, tau=np.zeros((1, 3)), nstate=nstate,
                                 f=np.full(nstate, 2.0))
    ctx = ops.Context(cell, gs)
    v = S.real_space_potential(sysm, seed=5)
    f = np.linspace(2.0, 0.25, nstate)
    f[nstate // 2] = 0.0   # an empty state: rhoofr's x pass skips it, vpsi still needs its columns
    return ctx, c, f, v


def _run(ctx, d_c0, f, n, d_v, c2_init):
    """rhoofr_dev, then vpsi_dev into a copy of c2_init: (rho, ekin, c2).
    Buffers are written before rhoofr: a library write in between (an
    upload) makes the kept columns stale."""
    d_rho = ops.DeviceArray(8 * ctx.nnr1)
    d_c2 = ctx.upload(c2_init)
    ek = ctx.rhoofr_dev(d_c0, f, n, d_rho)[0]
    ctx.vpsi_dev(d_c0, d_c2, f, d_v, n)
    return d_rho.to_host(ctx.nnr1), ek, ctx.download(d_c2, n)


def _launches(ctx, name):
    ctx.synchronize()
    return ctx.timer_get(name)[1]


@pytest.mark.parametrize("mode", ["fp64", "fp32"])
@pytest.mark.parametrize("kind,nstate,batch", [("cubic", 6, (3, 2)), ("ortho2", 5, (64, 64)),
                                               ("ortho3", 7, (2, 3)), ("mesh128", 9, (4, 3))])
def test_bit_identical(gpu, mode, kind, nstate, batch):
    ctx, c, f, v = _setup(kind, nstate)
    ctx.set_precision(mode)
    ctx.set_batch(*batch)   # rhoofr and vpsi batches differ: the kept slabs are per pair
    rng = np.random.default_rng(2)
    c2_init = np.asfortranarray(rng.standard_normal(c.shape) + 1j * rng.standard_normal(c.shape))
    d_c0 = ctx.upload(c)
    d_v = ops.DeviceArray(8 * ctx.nnr1).from_host(v)
    rho0, ek0, c20 = _run(ctx, d_c0, f, nstate, d_v, c2_init)
    ctx.set_column_reuse(True)
    ctx.timer_enable(True)
    rho1, ek1, c21 = _run(ctx, d_c0, f, nstate, d_v, c2_init)
    zinv_rho = _launches(ctx, "zinv")
    assert zinv_rho > 0
    assert np.array_equal(rho1, rho0) and ek1 == ek0
    assert np.array_equal(c21, c20)
    # one more step: vpsi with the kept columns launched no inverse z pass
    d_rho = ops.DeviceArray(8 * ctx.nnr1)
    d_c2 = ctx.upload(c2_init)
    ctx.timer_reset()
    ctx.rhoofr_dev(d_c0, f, nstate, d_rho)
    z_after_rho = _launches(ctx, "zinv")
    y_after_rho = _launches(ctx, "yinv")
    ctx.timer_reset()
    ctx.vpsi_dev(d_c0, d_c2, f, d_v, nstate)
    assert _launches(ctx, "zinv") == 0 and _launches(ctx, "yinv") == 0
    assert z_after_rho > 0 and y_after_rho > 0
    assert np.array_equal(ctx.download(d_c2, nstate), c20)
    # the columns were consumed: a second vpsi transforms again
    ctx.upload(c2_init, d_c2)
    ctx.timer_reset()
    ctx.vpsi_dev(d_c0, d_c2, f, d_v, nstate)
    assert _launches(ctx, "zinv") > 0
    assert np.array_equal(ctx.download(d_c2, nstate), c20)


def test_stale_columns_fall_back(gpu):
    ctx, c, f, v = _setup("mesh128", 8)
    n = c.shape[1]
    rng = np.random.default_rng(4)
    c2_init = np.asfortranarray(rng.standard_normal(c.shape) + 1j * rng.standard_normal(c.shape))
    d_v = ops.DeviceArray(8 * ctx.nnr1).from_host(v)
    d_c0 = ctx.upload(c)
    d_x = ctx.upload(np.asfortranarray(rng.standard_normal(c.shape) + 1j * rng.standard_normal(c.shape)))
    d_rho = ops.DeviceArray(8 * ctx.nnr1)

    def vpsi_fresh(d_c):
        ctx.set_column_reuse(False)
        d_c2 = ctx.upload(c2_init)
        ctx.vpsi_dev(d_c, d_c2, f, d_v, n)
        ctx.set_column_reuse(True)
        return ctx.download(d_c2, n)

    ctx.set_column_reuse(True)
    ctx.timer_enable(True)
    # 1. a library write to c0 between rhoofr and vpsi (c0 += 0.5 x)
    d_c2 = ctx.upload(c2_init)
    ctx.rhoofr_dev(d_c0, f, n, d_rho)
    ctx.axpy_wfn(0.5, d_x, d_c0, n)
    ctx.timer_reset()
    ctx.vpsi_dev(d_c0, d_c2, f, d_v, n)
    assert _launches(ctx, "zinv") > 0
    got = ctx.download(d_c2, n)
    assert np.array_equal(got, vpsi_fresh(d_c0))
    # 2. vpsi on another wavefunction buffer
    d_c2 = ctx.upload(c2_init)
    ctx.rhoofr_dev(d_c0, f, n, d_rho)
    ctx.timer_reset()
    ctx.vpsi_dev(d_x, d_c2, f, d_v, n)
    assert _launches(ctx, "zinv") > 0
    assert np.array_equal(ctx.download(d_c2, n), vpsi_fresh(d_x))
    # 3. another precision between the two calls
    ctx.set_precision("fp32")
    d_c2 = ctx.upload(c2_init)
    ctx.rhoofr_dev(d_c0, f, n, d_rho)
    ctx.set_precision("fp64")
    ctx.vpsi_dev(d_c0, d_c2, f, d_v, n)
    assert np.array_equal(ctx.download(d_c2, n), vpsi_fresh(d_c0))
    # 4. fewer states than rhoofr kept
    d_c2 = ctx.upload(c2_init)
    ctx.rhoofr_dev(d_c0, f, n, d_rho)
    ctx.timer_reset()
    ctx.vpsi_dev(d_c0, d_c2, f[:n - 2], d_v, n - 2)
    assert _launches(ctx, "zinv") > 0
    ctx.set_column_reuse(False)
    d_r = ctx.upload(c2_init)
    ctx.vpsi_dev(d_c0, d_r, f[:n - 2], d_v, n - 2)
    assert np.array_equal(ctx.download(d_c2, n), ctx.download(d_r, n))


def test_memory_limit_disables_keep(gpu):
    ctx, c, f, v = _setup("mesh128", 6)
    n = c.shape[1]
    d_c0 = ctx.upload(c)
    d_v = ops.DeviceArray(8 * ctx.nnr1).from_host(v)
    c2_init = np.zeros(c.shape, dtype=np.complex128, order="F")
    ref = _run(ctx, d_c0, f, n, d_v, c2_init)
    ctx.set_column_reuse(True)
    cur0 = ops.mem_stats()[0]
    try:
        ops.set_mem_limit(cur0 + (4 << 30))
        ctx.timer_enable(True)
        d_rho = ops.DeviceArray(8 * ctx.nnr1)
        d_c2 = ctx.upload(c2_init)
        ctx.rhoofr_dev(d_c0, f, n, d_rho)
        ctx.timer_reset()
        ctx.vpsi_dev(d_c0, d_c2, f, d_v, n)
        assert _launches(ctx, "zinv") > 0   # no kept columns under a limit
        assert np.array_equal(ctx.download(d_c2, n), ref[2])
    finally:
        ops.set_mem_limit(0)
    # without the limit the columns are kept: the library holds them
    d_rho = ops.DeviceArray(8 * ctx.nnr1)
    before = ops.mem_stats()[0]
    ctx.rhoofr_dev(d_c0, f, n, d_rho)
    held = ops.mem_stats()[0] - before
    assert held > 0, held
    ctx.set_column_reuse(False)
    assert ops.mem_stats()[0] < before + held   # off releases them


@pytest.mark.parametrize("mode,nstate,ecut", [("mixed", 128, 70.0), ("mixed_bf16x3", 128, 70.0), ("fp64", 128, 70.0),
                                              ("mixed", 127, 70.0), ("mixed", 128, 66.0)])
def test_cp_steps_bit_identical(gpu, mode, nstate, ecut):
    """Three velocity Verlet CP steps of the 128-state system on the 128^3 mesh
    with column reuse on (rhoofr's columns for vpsi; in mixed precision also
    the kept bf16 planes of the coefficients: written by posupa's final
    rotation for even state counts, split by rnlsm otherwise, and used by the
    next rortv and posupa) and off: every energy, force and coefficient bit
    for bit.  At 66 Ry the padded leading dimension exceeds the plane rows by
    a whole 32-row block (ldg - ngw = 52 >= 16): the planes written by the
    rotation must follow the K = 2 ngw layout of split_planes, not M = 2 ldg."""
    from cpmd_amd.ks import KSDevice
    from test_oracle_energy import ks_system
    sysm, _, c = ks_system(24.0, 64, ecut, nstate, seed=3)
    if ecut == 66.0:
        ngw = sysm.gset.ngw
        assert (2 * ngw + 31) // 32 < 2 * ((ngw + 63) // 64 * 64) // 32
    assert 2 * sysm.gset.ngw >= 16384   # the split-bf16 contraction path in mixed precision
    out = {}
    for reuse in (False, True):
        ks = KSDevice(sysm, column_reuse=reuse)
        ctx = ks.ctx
        ctx.set_precision(mode)
        n = c.shape[1]
        bufs = {k: ctx.alloc_wfn(n) for k in ("c0", "cm", "cp", "c2")}
        ctx.upload(c, bufs["c0"])
        print(f"CP steps, {mode}, {n} states, column reuse {reuse}", flush=True)
        st = ks.cp_init(bufs["c0"], bufs["cm"], bufs["c2"], sysm.f, sysm.tau, emass=400.0, ion_mass=21874.7)
        hist = []
        for _ in range(3):
            ks.cp_step(st, bufs, sysm.f, 4.0)
            hist.append((st["etot"], st["ekinc"], st["iters"], np.array(st["fion"])))
        # rekine (from the rotation's epilogue in mixed precision) against the
        # per-state Gamma norms of the final cm
        assert st["ekinc"] == pytest.approx(400.0 * ctx.dotp_wfn(bufs["cm"], bufs["cm"], n).sum(), rel=1e-12)
        out[reuse] = (hist, ctx.download(bufs["c0"], n), ctx.download(bufs["cm"], n))
        ctx.close()
    (h0, c00, cm0), (h1, c01, cm1) = out[False], out[True]
    for a, b in zip(h0, h1):
        assert a[0] == b[0] and a[1] == b[1] and a[2] == b[2]
        assert np.array_equal(a[3], b[3])
    assert np.array_equal(c00, c01) and np.array_equal(cm0, cm1)


def test_kept_planes_stale_cases(gpu):
    """The kept bf16 planes of c0 (mixed precision, column reuse on) are used
    only when nothing could have changed c0 since they were made; each case
    below runs with reuse on and off and must agree bit for bit:
    rnlsm, a library write to c0, then velupa + rortv; rnlsm on one buffer,
    velupa + rortv on another; rnlsm, velupa + rortv, then posupa after a
    library write to c0."""
    from cpmd_amd.ks import KSDevice
    from test_oracle_energy import ks_system
    sysm, _, c = ks_system(24.0, 64, 70.0, 64, seed=5)
    n = c.shape[1]
    rng = np.random.default_rng(9)
    # small perturbations: c0 stays orthonormal to 1e-5, so that rortog converges
    x = np.asfortranarray(1e-8 * (rng.standard_normal(c.shape) + 1j * rng.standard_normal(c.shape)))
    x[0] = x[0].real
    g = np.asfortranarray(1e-5 * (rng.standard_normal(c.shape) + 1j * rng.standard_normal(c.shape)))
    g[0] = g[0].real

    def run(reuse):
        ks = KSDevice(sysm, column_reuse=reuse)
        ctx = ks.ctx
        ctx.set_precision("mixed")
        ks.set_ions(sysm.tau)
        d = {k: ctx.upload(v) for k, v in (("c0", c), ("c1", c), ("x", x), ("c2", g), ("cm", g))}
        d["cp"] = ctx.alloc_wfn(n)
        out = []
        # 1. rnlsm(c0); c0 += x (library write); velupa + rortv on c0
        ctx.rnlsm_dev(d["c0"], n, tfor=True)
        ctx.axpy_wfn(1.0, d["x"], d["c0"], n)
        out.append(ctx.cp_velupa_rortv_dev(d["c0"], d["c2"], d["cm"], n, 0.01, 400.0)[0])
        out.append(ctx.download(d["cm"], n))
        # 2. rnlsm on c1, velupa + rortv on c0
        ctx.rnlsm_dev(d["c1"], n, tfor=True)
        out.append(ctx.cp_velupa_rortv_dev(d["c0"], d["c2"], d["cm"], n, 0.01, 400.0)[0])
        out.append(ctx.download(d["cm"], n))
        # 3. rnlsm(c1), velupa + rortv on c1 (planes kept), c1 += x, posupa on c1
        ctx.rnlsm_dev(d["c1"], n, tfor=True)
        ctx.cp_velupa_rortv_dev(d["c1"], d["c2"], d["cm"], n, 0.01, 400.0)
        ctx.axpy_wfn(1.0, d["x"], d["c1"], n)
        it = ctx.posupa_dev(d["c1"], d["c2"], d["cm"], d["cp"], n, 0.01, 4.0)[1]
        out.append(it)
        out.append(ctx.download(d["cp"], n))
        out.append(ctx.download(d["cm"], n))
        ctx.close()
        return out

    print("kept planes: stale cases, reuse on and off", flush=True)
    a, b = run(False), run(True)
    for u, v in zip(a, b):
        if isinstance(u, np.ndarray):
            assert np.array_equal(u, v)
        else:
            assert u == v
