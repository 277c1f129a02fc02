"""Plane-wave setup on the host: FFT mesh choice, Gamma-point G-vector sets in
CPMD's order, and the seeded synthetic systems the kernels are measured on.

This is the host-side mirror of the setup CPMD performs before its hot loop:

* mesh choice      -- ``src/numpw_utils.mod.F90:188-219`` (``change_gcut``) and the
                      default transform-length table of ``src/fftchk_utils.mod.F90:149-156``
                      or that of an FFTW build (``:73-97``, radices 2, 3, 5, 7);
* leading dims     -- ``leadim`` ``src/loadpa_utils.mod.F90:508-525``;
* G-vector set     -- generation loop ``src/loadpa_utils.mod.F90:283-331`` (half sphere
                      i>0 | i=0,j>0 | i=j=0,k>=0; shell tie-break key
                      g2*(1 +- sqrt(ig-1)*epsgx)), sorted by ``gsort``
                      (``src/loadpa_utils.mod.F90:748-792``), exact |G|^2 recomputed in
                      ``rggen`` (``src/rggen_utils.mod.F90:120-130``).

Everything here is numpy on the host; the GPU library receives the results
(Miller indices, |G|^2) through its C ABI exactly as CPMD would pass ``inyh``/``hg``.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

# Acceptable transform lengths of CPMD's default FFT (fftchk_def,
# src/fftchk_utils.mod.F90:149-156).  The table has to be identical for the
# mesh choice to agree with CPMD's; it is not the set of all 5-smooth numbers
# (50 is missing, for example).
FFT_LENGTHS = (
    3, 4, 5, 6, 8, 9, 12, 15, 16, 18, 20, 24, 25, 27, 30, 32, 36, 40, 45, 48, 54,
    60, 64, 72, 75, 80, 81, 90, 96, 100, 108, 120, 125, 128, 135, 144, 150, 160,
    162, 180, 192, 200, 216, 225, 240, 243, 256, 270, 288, 300, 320, 324, 360,
    375, 384, 400, 405, 432, 450, 480, 486, 500, 512, 540, 576, 600, 625, 640,
    648, 675, 720, 729, 750, 768, 800, 810, 864, 900, 960, 972, 1000, 1024,
    1080, 1152, 1200, 1280, 1296, 1350, 1440, 1458, 1500, 1536, 1600, 1620,
    1728, 1800, 1920, 1944, 2000, 2048)


# The table of a CPMD built with FFTW (fftchk_fftw, roots 2, 3, 5, 7,
# src/fftchk_utils.mod.F90:73-97): the 7-smooth numbers from 2 to 8192, none
# left out, so it is generated rather than listed.
def _smooth7(limit: int) -> tuple:
    out = []
    for v in range(2, limit + 1):
        m = v
        for r in (2, 3, 5, 7):
            while m % r == 0:
                m //= r
        if m == 1:
            out.append(v)
    return tuple(out)


FFT_LENGTHS_FFTW = _smooth7(8192)
FFT_TABLES = {"default": FFT_LENGTHS, "fftw": FFT_LENGTHS_FFTW}

EPSGX = 1.0e-12  # src/gvec.mod.F90:12


def fftchk(m: int, n: int = 2, table: str = "default") -> int:
    """Next acceptable transform length >= m (n=2: next even one, n=1: next one,
    n<0: next smaller one) from CPMD's default table (fftchk_def,
    src/fftchk_utils.mod.F90:135-170) or, with table="fftw", the table of an
    FFTW build (fftchk_fftw, :59-133)."""
    lft = FFT_TABLES[table]
    for i, v in enumerate(lft):
        if v >= m:
            if n < 0:
                return lft[i - 1]
            if n == 1:
                return v
            if n == 2 and v % 2 == 0:
                return v
    raise ValueError(f"FFTCHK: minimal mesh size is larger than {lft[-1]} (m={m})")


def fortran_nint(x: float) -> int:
    """Fortran NINT: round half away from zero."""
    return int(math.floor(abs(x) + 0.5)) * (1 if x >= 0 else -1)


def leadim(nr: int) -> int:
    """Padded leading dimension kr = nr + mod(nr+1, 2) (src/loadpa_utils.mod.F90:519-521)."""
    return nr + (nr + 1) % 2


def mesh_cubic(alat: float, ecut_ry: float, dual: float = 4.0, mesh_min: int = 0, table: str = "default") -> int:
    """Mesh for a simple cubic cell, src/numpw_utils.mod.F90:188-193,213-219."""
    mnr = fortran_nint(alat / math.pi * math.sqrt(dual * ecut_ry) + 0.5)
    return fftchk(max(mnr, mesh_min), 2, table)


@dataclass
class GSet:
    """Gamma-point G vectors in CPMD order.

    ``miller[ig]`` = (i, j, k) with G = i*b1 + j*b2 + k*b3 (units 2*pi/alat),
    ``hg[ig]`` = |G|^2 in (2*pi/alat)^2, the first ``ngw`` are the wavefunction set.
    CPMD's ``inyh(:,ig)`` is ``miller[ig] + (nr//2 + 1)``.
    """
    miller: np.ndarray      # (nhg, 3) int32
    hg: np.ndarray          # (nhg,) float64
    ngw: int
    nhg: int


@dataclass
class Cell:
    alat: float                     # bohr
    avec: np.ndarray                # (3,3) rows a1,a2,a3 in bohr
    ecut: float                     # Ry, wavefunction cutoff
    dual: float = 4.0
    nr: tuple = (0, 0, 0)

    @property
    def omega(self) -> float:
        return float(abs(np.linalg.det(self.avec)))

    @property
    def tpiba(self) -> float:
        return 2.0 * math.pi / self.alat

    @property
    def tpiba2(self) -> float:
        return self.tpiba ** 2

    @property
    def bvec(self) -> np.ndarray:
        """Reciprocal vectors in units of 2*pi/alat (rows b1,b2,b3), as ``recips``."""
        return np.linalg.inv(self.avec / self.alat).T

    @property
    def kr(self) -> tuple:
        return tuple(leadim(n) for n in self.nr)

    @property
    def gcutw(self) -> float:
        return self.ecut / self.tpiba2

    @property
    def gcut(self) -> float:
        return self.gcutw * self.dual


def cubic_cell(a: float, ecut: float, dual: float = 4.0, mesh: int = 0, table: str = "default") -> Cell:
    n = mesh_cubic(a, ecut, dual, mesh, table)
    return Cell(alat=a, avec=np.eye(3) * a, ecut=ecut, dual=dual, nr=(n, n, n))


def orthorhombic_cell(a: float, b: float, c: float, ecut: float, dual: float = 4.0, table: str = "default") -> Cell:
    """Orthorhombic cell (alat = a); mesh per axis as the general branch of
    src/numpw_utils.mod.F90:206-212."""
    ns = tuple(fftchk(fortran_nint(x / math.pi * math.sqrt(dual * ecut) + 0.5), 2, table) for x in (a, b, c))
    return Cell(alat=a, avec=np.diag([a, b, c]).astype(float), ecut=ecut, dual=dual, nr=ns)


def gvectors(cell: Cell, with_density: bool = True) -> GSet:
    """G vectors with |G|^2 < gcut (density) / gcutw (wavefunction), CPMD order.

    Follows src/loadpa_utils.mod.F90:283-331 (generation order and tie-break key)
    and ``gsort``.  When ``with_density`` is False only the wavefunction set is
    returned (its order is the same: every wavefunction key sorts below every
    density-only key).
    """
    b = cell.bvec
    gcut = cell.gcut if with_density else cell.gcutw
    gcutw = cell.gcutw
    nr1, nr2, nr3 = cell.nr
    # |i| <= sqrt(gcut)*|a_i|/alat bounds the sphere; restricting the loop to it
    # keeps CPMD's loop order and generation index of the in-sphere vectors.
    lim = [int(math.floor(math.sqrt(gcut) * np.linalg.norm(cell.avec[d]) / cell.alat)) + 1
           for d in range(3)]
    imax = min(nr1 - 1, lim[0])
    jmax = min(nr2 - 1, lim[1])
    kmax = min(nr3 - 1, lim[2])
    ii = np.arange(0, imax + 1)
    jj = np.arange(-jmax, jmax + 1)
    kk = np.arange(-kmax, kmax + 1)
    I, J, K = np.meshgrid(ii, jj, kk, indexing="ij")
    I = I.ravel(); J = J.ravel(); K = K.ravel()
    half = (I > 0) | ((I == 0) & (J > 0)) | ((I == 0) & (J == 0) & (K >= 0))
    I = I[half]; J = J[half]; K = K[half]
    t = (I[:, None] * b[0][None, :] + J[:, None] * b[1][None, :] + K[:, None] * b[2][None, :])
    g2 = t[:, 0] * t[:, 0] + t[:, 1] * t[:, 1] + t[:, 2] * t[:, 2]
    inside = g2 < gcut
    I = I[inside]; J = J[inside]; K = K[inside]; g2 = g2[inside]
    ig = np.arange(len(g2), dtype=np.float64)          # ig-1 of the reference loop
    sign = np.where(g2 < gcutw, -1.0, 1.0)
    key = g2 * (1.0 + np.sqrt(ig) * EPSGX * sign)
    order = np.argsort(key, kind="stable")
    miller = np.stack([I, J, K], axis=1)[order].astype(np.int32)
    t = miller.astype(np.float64) @ b
    hg = t[:, 0] * t[:, 0] + t[:, 1] * t[:, 1] + t[:, 2] * t[:, 2]
    ngw = int(np.count_nonzero(g2 < gcutw))
    return GSet(miller=np.ascontiguousarray(miller), hg=np.ascontiguousarray(hg),
                ngw=ngw, nhg=len(hg))


# ----------------------------------------------------------------------------
# Seeded synthetic systems (one species, analytic Goedecker pseudopotential)
# ----------------------------------------------------------------------------

@dataclass
class GthPP:
    """Goedecker-Teter-Hutter pseudopotential with one s projector
    (the &POTENTIAL GOEDECKER block format of src/recpnew_utils.mod.F90:286-322)."""
    zion: float = 4.0
    rloc: float = 0.45
    c1: float = -6.0
    c2: float = 0.9
    r0: float = 0.40
    h11: float = 5.0


@dataclass
class SyntheticSystem:
    cell: Cell
    gset: GSet
    tau: np.ndarray               # (nat,3) bohr
    nstate: int
    f: np.ndarray                 # occupations
    pp: GthPP = field(default_factory=GthPP)

    @property
    def nat(self) -> int:
        return self.tau.shape[0]


def place_atoms(a: float, nat: int, seed: int, dmin: float = 3.0) -> np.ndarray:
    """Uniform positions in [0,a)^3, rejecting any closer than dmin (minimum image)
    to an accepted one."""
    rng = np.random.default_rng(seed)
    tau = np.zeros((0, 3))
    tries = 0
    while tau.shape[0] < nat:
        p = rng.uniform(0.0, a, size=3)
        if tau.shape[0]:
            d = tau - p
            d -= a * np.round(d / a)
            if np.min(np.einsum("ij,ij->i", d, d)) < dmin * dmin:
                tries += 1
                if tries > 10_000_000:
                    raise RuntimeError("atom placement failed")
                continue
        tau = np.vstack([tau, p])
    return tau


def make_system(a: float, nat: int, ecut: float, nstate: int | None = None,
                seed: int = 1234, mesh: int = 0, with_density: bool = False,
                fft_table: str = "default") -> SyntheticSystem:
    """fft_table: "default" (CPMD's own table) or "fftw" (the mesh a CPMD
    built with FFTW chooses)."""
    cell = cubic_cell(a, ecut, mesh=mesh, table=fft_table)
    gs = gvectors(cell, with_density=with_density)
    tau = place_atoms(a, nat, seed)
    if nstate is None:
        nstate = 2 * nat          # Zion=4 -> 4*nat electrons -> 2*nat doubly occupied states
    f = np.full(nstate, 2.0)
    return SyntheticSystem(cell=cell, gset=gs, tau=tau, nstate=nstate, f=f)


# The five systems the measurements are quoted on (cubic a, atoms, ecut Ry,
# states).  The mesh and plane-wave counts follow from CPMD's rules above.
CONFIGS = {
    "cfg1": dict(a=12.0, nat=8, ecut=40.0, nstate=16),
    "cfg2": dict(a=24.0, nat=64, ecut=70.0, nstate=128),
    "cfg4": dict(a=32.0, nat=256, ecut=80.0, nstate=512),
    "cfg5": dict(a=36.0, nat=512, ecut=90.0, nstate=1024),
}


def _real_rows(a: np.ndarray) -> np.ndarray:
    """(ngw, n) complex -> (n, 2*ngw) float64 rows (re, im interleaved), no copy
    when `a` is Fortran-ordered."""
    return np.ascontiguousarray(np.asfortranarray(a, dtype=np.complex128).T).view(np.float64)


def dotp_gamma(a: np.ndarray, b: np.ndarray) -> np.ndarray:
    """Gamma-point real inner product of half-sphere coefficient columns:
    Re(a0)*Re(b0) + 2*sum_{G!=0} Re(conj(a)*b)  (src/dotp_utils.mod.F90:58-101;
    G=0 is the first row)."""
    ra = _real_rows(a)
    rb = ra if b is a else _real_rows(b)
    r = 2.0 * (ra @ rb.T)
    r -= np.outer(ra[:, 0], rb[:, 0])
    return r


def initial_coefficients(sys: SyntheticSystem, seed: int = 4321, nstate: int | None = None,
                         orthonormalize: bool = True) -> np.ndarray:
    """Seeded complex Gaussian c_i(G) = (x+iy) exp(-|G|^2/8), |G| in bohr^-1,
    c_i(0) real, Loewdin-orthonormalised in the Gamma metric.
    Returns (ngw, nstate) complex128, Fortran (column-major) order."""
    n = sys.nstate if nstate is None else nstate
    ngw = sys.gset.ngw
    rng = np.random.default_rng(seed)
    g2 = sys.gset.hg[:ngw] * sys.cell.tpiba2
    amp = np.exp(-g2 / 8.0)
    c = (rng.standard_normal((ngw, n)) + 1j * rng.standard_normal((ngw, n))) * amp[:, None]
    c[0, :] = c[0, :].real
    c = np.asfortranarray(c)
    if orthonormalize:
        s = dotp_gamma(c, c)
        w, z = np.linalg.eigh(s)
        a = z @ np.diag(1.0 / np.sqrt(w)) @ z.T
        r = a.T @ _real_rows(c)                    # rows of the real view: C <- C A
        c = np.asfortranarray(r.view(np.complex128).T)
        c[0, :] = c[0, :].real
    return c


def gth_local_potential_g(sys: SyntheticSystem, hg: np.ndarray) -> np.ndarray:
    """Analytic GTH local form factor V_loc(G)/Omega (Hartree) for G != 0; G=0 gets
    the finite part.  Used only to build a physical-looking V(r) for measurements."""
    pp = sys.pp
    om = sys.cell.omega
    g2 = hg * sys.cell.tpiba2
    rl2 = pp.rloc ** 2
    x = g2 * rl2
    e = np.exp(-0.5 * x)
    with np.errstate(divide="ignore", invalid="ignore"):
        v = (-4.0 * math.pi * pp.zion / om * e / g2
             + math.sqrt(8.0 * math.pi ** 3) * pp.rloc ** 3 / om * e
             * (pp.c1 + pp.c2 * (3.0 - x)))
    v0 = (2.0 * math.pi * pp.zion * rl2 / om
          + math.sqrt(8.0 * math.pi ** 3) * pp.rloc ** 3 / om * (pp.c1 + 3.0 * pp.c2))
    v = np.where(g2 > 0, v, v0)
    return v


def gth_projector_g(sys: SyntheticSystem, hg: np.ndarray) -> np.ndarray:
    """Analytic s-projector twnl(G) of the GTH pseudopotential: the radial
    form factor times Y_00 = 1/sqrt(4 pi) times 1/sqrt(Omega), as CPMD builds
    twnl = ylmb tw vol (src/rinforce_utils.mod.F90:564-576), so that the
    projector has unit norm (sum over the G sphere of twnl^2 -> 1) and the
    nonlocal operator of one atom has eigenvalue h11."""
    pp = sys.pp
    g2 = hg * sys.cell.tpiba2
    r0 = pp.r0
    return (4.0 * math.sqrt(2.0) * math.pi ** 1.25 * r0 ** 1.5 / math.sqrt(4.0 * math.pi)
            * np.exp(-0.5 * g2 * r0 * r0) / math.sqrt(sys.cell.omega))


def structure_factors(sys: SyntheticSystem, miller: np.ndarray) -> np.ndarray:
    """eigr(G, ia) = exp(-i G.R_ia), (ng, nat) complex128 column-major
    (the role of ``eigr`` from ``phfac``)."""
    frac = sys.tau @ np.linalg.inv(sys.cell.avec)       # fractional coordinates
    phase = -2.0 * math.pi * (miller.astype(np.float64) @ frac.T)
    return np.asfortranarray(np.exp(1j * phase))


def projector_set(sys: SyntheticSystem, miller: np.ndarray, hg: np.ndarray, with_eigr: bool = True):
    """The one-species, one-s-projector Goedecker set of the synthetic systems as
    the projector arguments of the library: (na, nhxs, ngh, lang, hmat, twnl, eigr);
    eigr is None when with_eigr is False (the device builds it, mc_phfac)."""
    twnl = gth_projector_g(sys, hg)[:, None, None]            # (ngw, nhxs=1, nsp=1)
    eigr = structure_factors(sys, miller) if with_eigr else None
    return (np.array([sys.nat], np.int32), 1, np.array([1], np.int32), np.zeros((1, 1), np.int32),
            np.full((1, 1, 1), sys.pp.h11), np.asfortranarray(twnl), eigr)


def local_pp_set(sys: SyntheticSystem, hg: np.ndarray, raggio: float = 1.2):
    """Local pseudopotential and Gaussian ion charge of the synthetic species on
    the density G set: (zv, raggio, vps, rhops), vps / rhops (nsp=1, nhg).

    vps = voo / Omega with CPMD's Goedecker local form factor
    (src/formf_utils.mod.F90:131-187: the difference of the erf-screened
    Coulomb tail and the Gaussian ion charge of radius raggio plus the C1, C2
    terms), evaluated directly instead of through CPMD's radial spline;
    rhops = -Zv exp(-G^2 raggio^2 / 4) / Omega (putps,
    src/rinforce_utils.mod.F90:430-457).  raggio 1.2 is CPMD's default radius."""
    pp = sys.pp
    om = sys.cell.omega
    gg = hg * sys.cell.tpiba2
    rc2 = pp.rloc ** 2
    r2 = raggio ** 2
    rs1 = math.sqrt((2.0 * math.pi) ** 3) * pp.rloc ** 3 * np.exp(-0.5 * gg * rc2)
    with np.errstate(divide="ignore", invalid="ignore"):
        voo = np.where(gg < 1e-8, math.pi * pp.zion * (2.0 * rc2 - r2),
                       -4.0 * math.pi * pp.zion / gg * (np.exp(-0.5 * gg * rc2) - np.exp(-0.25 * gg * r2)))
    voo = voo + pp.c1 * rs1 + pp.c2 * (3.0 - rc2 * gg) * rs1
    vps = (voo / om)[None, :]
    rhops = (-pp.zion * np.exp(-0.25 * r2 * gg) / om)[None, :]
    return np.array([pp.zion]), np.array([raggio]), np.asfortranarray(vps), np.asfortranarray(rhops)


def real_space_potential(sys: SyntheticSystem, seed: int = 99, amplitude: float = 1.0) -> np.ndarray:
    """A smooth, seeded real potential V(r) on the padded mesh, layout
    v[ix + kr1*(iy + kr2*iz)] (CPMD's vpot(nnr1)).  Padding entries are zero."""
    nr1, nr2, nr3 = sys.cell.nr
    kr1, kr2, kr3 = sys.cell.kr
    rng = np.random.default_rng(seed)
    x = np.arange(nr1) / nr1
    y = np.arange(nr2) / nr2
    z = np.arange(nr3) / nr3
    X, Y, Z = np.meshgrid(x, y, z, indexing="ij")
    v = np.zeros((nr1, nr2, nr3))
    for _ in range(6):
        h = rng.integers(-3, 4, size=3)
        ph = rng.uniform(0, 2 * math.pi)
        amp = rng.uniform(-0.5, 0.5) * amplitude
        v += amp * np.cos(2 * math.pi * (h[0] * X + h[1] * Y + h[2] * Z) + ph)
    v -= 0.3 * amplitude
    out = np.zeros((kr3, kr2, kr1))
    out[:nr3, :nr2, :nr1] = v.transpose(2, 1, 0)
    return out.ravel()
