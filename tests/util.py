"""Shared helpers of the test suite: small seeded systems and closed-form
references that do not go through any FFT."""
from __future__ import annotations

import math

import numpy as np

from cpmd_amd import system as S


def small_system(kind: str = "cubic", nstate: int = 6, seed: int = 7):
    if kind == "cubic":
        cell = S.cubic_cell(6.0, 20.0)
    elif kind == "ortho":
        cell = S.orthorhombic_cell(6.0, 6.0, 8.0, 20.0)
    elif kind == "ortho2":
        cell = S.orthorhombic_cell(5.0, 6.0, 7.0, 15.0)
    elif kind == "ortho3":   # n1 = 30: x pass through the runtime-plan transform
        cell = S.orthorhombic_cell(9.0, 6.5, 7.0, 20.0)
    else:
        raise ValueError(kind)
    gs = S.gvectors(cell, with_density=False)
    sysm = S.SyntheticSystem(cell=cell, gset=gs, tau=np.zeros((1, 3)), nstate=nstate,
                             f=np.full(nstate, 2.0))
    c = S.initial_coefficients(sysm, seed=seed, nstate=nstate)
    return cell, gs, c


def config_system(name: str, nstate: int | None = None, seed: int = 4321):
    c = S.CONFIGS[name]
    sysm = S.make_system(c["a"], c["nat"], c["ecut"], nstate=c["nstate"] if nstate is None else nstate)
    coef = S.initial_coefficients(sysm, seed=seed)
    return sysm, coef


def frac_grid(cell):
    n1, n2, n3 = cell.nr
    s1 = np.arange(n1) / n1
    s2 = np.arange(n2) / n2
    s3 = np.arange(n3) / n3
    return np.meshgrid(s1, s2, s3, indexing="ij")


def psi_direct(cell, miller: np.ndarray, c: np.ndarray) -> np.ndarray:
    """psi(r) = c(0) + sum_{G != 0 on the half sphere} (c e^{iG.r} + conj(c) e^{-iG.r}),
    by direct summation on the (n1,n2,n3) mesh (G=0 first row)."""
    X, Y, Z = frac_grid(cell)
    psi = np.full(X.shape, c[0], dtype=np.complex128)
    for g in range(1, len(c)):
        m = miller[g]
        ph = np.exp(2j * math.pi * (m[0] * X + m[1] * Y + m[2] * Z))
        psi += c[g] * ph + np.conj(c[g]) * np.conj(ph)
    return psi


def to_padded(cell, a: np.ndarray) -> np.ndarray:
    """(n1,n2,n3) array -> CPMD real-space vector v[ix + kr1*(iy + kr2*iz)] with zero padding."""
    n1, n2, n3 = cell.nr
    k1, k2, k3 = cell.kr
    out = np.zeros((k3, k2, k1), dtype=a.dtype)
    out[:n3, :n2, :n1] = a.transpose(2, 1, 0)
    return out.ravel()


def from_padded(cell, v: np.ndarray) -> np.ndarray:
    n1, n2, n3 = cell.nr
    k1, k2, k3 = cell.kr
    return v.reshape(k3, k2, k1)[:n3, :n2, :n1].transpose(2, 1, 0)


def miller_index(miller: np.ndarray):
    """dict (m1,m2,m3) -> (row, conj flag) over the half sphere and its mirror."""
    d = {}
    for g, m in enumerate(miller):
        d[(int(m[0]), int(m[1]), int(m[2]))] = (g, False)
        d[(-int(m[0]), -int(m[1]), -int(m[2]))] = (g, True)
    return d


def coef_at(d, c_col: np.ndarray, key) -> complex:
    if key not in d:
        return 0.0
    g, cj = d[key]
    return np.conj(c_col[g]) if cj else c_col[g]
