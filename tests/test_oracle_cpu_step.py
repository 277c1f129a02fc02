"""The threaded CPU path of the bench's CPU baseline (oracle/cpu_step.py and
oracle/cpu_step.c) against the plain-C oracle on the same seeded inputs:
rhoofr, vpsi (with an empty state: the f/2 = 0 -> 1 rule, and an odd state
count: a lone last state), Loewdin, rnlsm1 and fnonloc, all to rounding."""
import numpy as np
import pytest

import oracle
from cpmd_amd import system as S
from oracle.cpu_step import CpuStep
from util import small_system


@pytest.mark.parametrize("kind,nstate", [("cubic", 5), ("ortho", 4), ("ortho3", 3), ("ortho2", 1)])
def test_transforms_match_oracle(kind, nstate):
    cell, gs, c = small_system(kind, nstate)
    L = oracle.Layout(cell, gs)
    cs = CpuStep(cell, gs)
    f = np.linspace(2.0, 0.0, nstate) if nstate > 1 else np.array([2.0])
    rho = cs.rhoofr(c, f)
    assert np.abs(rho - oracle.rhoofr(L, c, f)[0]).max() < 1e-14
    sysm = S.SyntheticSystem(cell=cell, gset=gs, tau=np.zeros((1, 3)), nstate=nstate, f=np.full(nstate, 2.0))
    v = S.real_space_potential(sysm, seed=11)
    rng = np.random.default_rng(1)
    c2 = np.asfortranarray(rng.standard_normal(c.shape) + 1j * rng.standard_normal(c.shape))
    out, ref = cs.vpsi(c, c2, f, v), oracle.vpsi(L, c, c2, f, v)
    assert np.abs(out - ref).max() < 1e-13 * np.abs(ref).max()


def test_contractions_match_oracle():
    cell, gs, c = small_system("ortho2", 6)
    cs = CpuStep(cell, gs)
    cc = np.asfortranarray(c * 1.1 + 0.01)
    cc[0] = cc[0].real
    assert np.abs(cs.ovlap(c, cc) - oracle.ovlap(c, cc)).max() < 1e-13
    assert np.abs(cs.lowdin(cc) - oracle.lowdin(cc)[0]).max() < 1e-13
    sysm = S.SyntheticSystem(cell=cell, gset=gs, tau=np.array([[1.0, 2.0, 0.5], [3.0, 0.2, 4.0]]), nstate=6,
                             f=np.full(6, 2.0))
    na, nhxs, ngh, lang, hmat, twnl, eigr = S.projector_set(sysm, gs.miller[:gs.ngw], gs.hg[:gs.ngw])
    P = oracle.Projectors(na=na, nhxs=1, ngh=ngh, lang=lang, hmat=hmat, twnl=twnl, eigr=eigr)
    beta = twnl[:, 0, 0][:, None] * eigr
    fnl, fnl_o = cs.rnlsm1(c, beta), oracle.rnlsm1(P, c)
    assert np.abs(fnl - fnl_o[:, 0, :]).max() < 1e-14 * np.abs(fnl_o).max()
    f = np.linspace(2.0, 0.5, 6)
    c2 = cs.fnonloc(np.zeros_like(c), beta, -(f[None, :] * hmat[0, 0, 0]) * fnl)
    c2_o = oracle.fnonloc(P, np.zeros_like(c), f, fnl_o)
    assert np.abs(c2 - c2_o).max() < 1e-13 * np.abs(c2_o).max()
