"""Writes ks_128states_128cubed.npz: the oracle's E_KS terms, ionic forces and
fingerprints of the gradient c2 = -f H c for the seeded 128-state system on the
128^3 mesh (a = 24 bohr, 64 atoms, 70 Ry; the same inputs the GPU test builds),
so that the GPU test checks the whole step at that size without spending the
oracle's ~90 s of host time.  Fingerprints of c2: per state the Gamma-metric norm
<c2_i|c2_i> and the overlap <r_i|c2_i> with a seeded random vector r_i.

Run (about 90 s): python tests/golden/make_ks_128states.py"""
import sys
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))
sys.path.insert(0, str(ROOT / "tests"))

import oracle                                   # noqa: E402
from test_oracle_energy import ks_system        # noqa: E402

CASE = dict(a=24.0, nat=64, ecut=70.0, nstate=128, seed=3)


def probes(c):
    """Seeded random coefficient vectors, one per state (G = 0 real)."""
    rng = np.random.default_rng(2024)
    r = rng.standard_normal(c.shape) + 1j * rng.standard_normal(c.shape)
    r[0] = r[0].real
    return r


def gamma_dots(a, b):
    """Per column the Gamma-metric product 2 Re sum_G conj(a) b - a(0) b(0)
    (src/dotp_utils.mod.F90:58-101, G = 0 first and real)."""
    return 2.0 * np.real(np.sum(np.conj(a) * b, axis=0)) - a[0].real * b[0].real


def fingerprints(c2, r):
    return gamma_dots(c2, c2), gamma_dots(r, c2)


def main():
    sysm, K, c = ks_system(CASE["a"], CASE["nat"], CASE["ecut"], CASE["nstate"], seed=CASE["seed"])
    o = oracle.ks_energy(K, c, sysm.f, sysm.tau, forces=True, gradient=True)
    norms, over = fingerprints(o["c2"], probes(c))
    keys = ("ekin", "ehep", "epseu", "exc", "esr", "eself", "enl", "etot")
    np.savez(Path(__file__).with_name("ks_128states_128cubed.npz"), energies=np.array([o[k] for k in keys]),
             keys=np.array(keys), fion=o["fion"], c2_norms=norms, c2_overlaps=over, mesh=np.array(sysm.cell.nr))
    print("etot", o["etot"])


if __name__ == "__main__":
    main()
