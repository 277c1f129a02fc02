"""The allocation-free vpsi (mc_vpsi_dev_noalloc) against mc_vpsi_dev, through
the C ABI: the same kernels on the same inputs, so c2 must agree bit for bit.
Two calls in a row on each context, the second with fewer states, so the
second call runs on scratch contents the first one left behind; the 128 MiB
scratch holds fewer state pairs than the batch size at 128^3 (5 in FP64, 11 in
FP32 against 64), so the pairs-per-launch path differs from mc_vpsi_dev's.
A state pair whose workspaces exceed the scratch (FP64 transforms at 240^3)
is refused with MC_ERR_ARG before any work is queued."""
import ctypes as C

import numpy as np
import pytest

from cpmd_amd import ops
from cpmd_amd import system as S
from cpmd_amd._lib import lib
from util import config_system

pytestmark = pytest.mark.gpu

MC_ERR_ARG = 1


def _p(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _compare(ctx, sysm, c, nstate, seed):
    """c2 += vpsi via mc_vpsi_dev and via mc_vpsi_dev_noalloc from the same c2."""
    rng = np.random.default_rng(seed)
    cc = np.asfortranarray(c[:, :nstate])
    f = np.ascontiguousarray(np.linspace(2.0, 0.0, nstate))   # last state empty: the f = 0 -> 1 rule
    v = S.real_space_potential(sysm, seed=seed)
    vd = ops.DeviceArray(8 * ctx.nnr1).from_host(v)
    c0d = ctx.upload(cc)
    c2h = np.asfortranarray(rng.standard_normal(cc.shape) + 1j * rng.standard_normal(cc.shape))
    ref = ctx.upload(c2h)
    out = ctx.upload(c2h)
    assert lib().mc_vpsi_dev(ctx.h, c0d.ptr, ref.ptr, _p(f), vd.ptr, nstate) == 0, lib().mc_last_error()
    st = lib().mc_vpsi_dev_noalloc(ctx.h, c0d.ptr, out.ptr, _p(f), vd.ptr, nstate)
    assert st == 0, lib().mc_last_error()
    ctx.synchronize()
    a, b = ctx.download(ref, nstate), ctx.download(out, nstate)
    assert np.array_equal(a, b), np.abs(a - b).max()
    assert not np.array_equal(a, c2h)   # vpsi did add a force


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_noalloc_vpsi_matches_vpsi_dev(gpu, prec):
    sysm, c = config_system("cfg2")            # 128^3, 68,338 plane waves, 128 states
    ctx = ops.Context(sysm.cell, sysm.gset)
    ctx.set_precision(prec)
    print(f"noalloc vpsi cfg2 {prec}: 128 states, then 37", flush=True)
    _compare(ctx, sysm, c, 128, seed=1)
    _compare(ctx, sysm, c, 37, seed=2)          # smaller, odd: a lone last state
    ctx.close()


def test_noalloc_vpsi_small_mesh_twice(gpu):
    sysm, c = config_system("cfg1")            # 54^3: every pair of the batch fits
    ctx = ops.Context(sysm.cell, sysm.gset)
    print("noalloc vpsi cfg1 fp64: 16 states, then 3", flush=True)
    _compare(ctx, sysm, c, 16, seed=3)
    _compare(ctx, sysm, c, 3, seed=4)
    ctx.close()


def test_noalloc_vpsi_refuses_beyond_scratch(gpu):
    sysm, c = config_system("cfg5", nstate=2)  # 240^3: one FP64 pair needs 136 MB
    ctx = ops.Context(sysm.cell, sysm.gset)
    print("noalloc vpsi cfg5: fp64 refused, then fp32 (one pair per launch)", flush=True)
    f = np.array([2.0, 2.0])
    vd = ops.DeviceArray(8 * ctx.nnr1).zero()
    c0d = ctx.upload(c)
    c2d = ctx.alloc_wfn(2)
    st = lib().mc_vpsi_dev_noalloc(ctx.h, c0d.ptr, c2d.ptr, _p(f), vd.ptr, 2)
    assert st == MC_ERR_ARG
    assert "scratch" in lib().mc_last_error().decode()
    ctx.set_precision("fp32")
    _compare(ctx, sysm, c, 2, seed=5)
    ctx.close()
