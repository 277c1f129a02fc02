"""The whole force evaluation of the 128-state system on the 128^3 mesh
(a = 24 bohr, 64 atoms, 70 Ry, 68,338 plane waves): E_KS and its terms, the
ionic forces and the gradient c2 = -f H c from the device against the oracle's
values for the same seeded inputs (tests/golden/ks_128states_128cubed.npz,
written by tests/golden/make_ks_128states.py; the oracle takes ~90 s at this
size).  c2 is compared through per-state fingerprints: <c2_i|c2_i> and
<r_i|c2_i> with seeded random r_i (Gamma metric).

FP64: energies to 1e-11 relative, forces to 1e-9 of the largest, fingerprints
to 1e-10 of the largest.  Mixed (FP32 transforms, split-bf16 contractions):
|dE_KS| < 1e-6 Ha per atom, max |dF| < 1e-5 Ha/bohr, fingerprints to 1e-5."""
from pathlib import Path

import numpy as np
import pytest

from cpmd_amd.ks import KSDevice
from test_oracle_energy import ks_system

pytestmark = pytest.mark.gpu

GOLD = Path(__file__).resolve().parent / "golden" / "ks_128states_128cubed.npz"
CASE = dict(a=24.0, nat=64, ecut=70.0, nstate=128, seed=3)


def _probes(c):
    rng = np.random.default_rng(2024)
    r = rng.standard_normal(c.shape) + 1j * rng.standard_normal(c.shape)
    r[0] = r[0].real
    return r


def _gamma_dots(a, b):
    return 2.0 * np.real(np.sum(np.conj(a) * b, axis=0)) - a[0].real * b[0].real


@pytest.fixture(scope="module")
def step128():
    g = np.load(GOLD)
    sysm, K, c = ks_system(CASE["a"], CASE["nat"], CASE["ecut"], CASE["nstate"], seed=CASE["seed"])
    assert tuple(sysm.cell.nr) == tuple(g["mesh"])
    ks = KSDevice(sysm)
    ks.set_ions(sysm.tau)
    d_c = ks.ctx.upload(c)
    d_c2 = ks.ctx.alloc_wfn(c.shape[1])
    yield g, sysm, c, ks, d_c, d_c2
    ks.ctx.set_precision("fp64")
    ks.ctx.close()


@pytest.mark.parametrize("mode", ["fp64", "mixed"])
def test_step128_against_oracle(gpu, step128, mode):
    g, sysm, c, ks, d_c, d_c2 = step128
    ks.ctx.set_precision(mode)
    print(f"128 states / 128^3, {mode}: E_KS, c2, F_ion", flush=True)
    r = ks.energy(d_c, sysm.f, gradient=d_c2, forces=True)
    c2 = ks.ctx.download(d_c2, c.shape[1])
    keys = [str(k) for k in g["keys"]]
    e_o = dict(zip(keys, g["energies"]))
    norms, over = _gamma_dots(c2, c2), _gamma_dots(_probes(c), c2)
    nat = sysm.nat
    if mode == "fp64":
        for k in keys:
            assert r[k] == pytest.approx(e_o[k], rel=1e-11, abs=1e-11), k
        assert np.abs(r["fion"] - g["fion"]).max() < 1e-9 * max(1.0, np.abs(g["fion"]).max())
        assert np.abs(norms - g["c2_norms"]).max() < 1e-10 * np.abs(g["c2_norms"]).max()
        assert np.abs(over - g["c2_overlaps"]).max() < 1e-10 * np.abs(g["c2_overlaps"]).max()
    else:
        assert abs(r["etot"] - e_o["etot"]) < 1e-6 * nat
        assert np.abs(r["fion"] - g["fion"]).max() < 1e-5
        assert np.abs(norms - g["c2_norms"]).max() < 1e-5 * np.abs(g["c2_norms"]).max()
        assert np.abs(over - g["c2_overlaps"]).max() < 1e-5 * np.abs(g["c2_overlaps"]).max()
