"""ORACLE -- test infrastructure only.

ctypes wrapper of the plain-C FP64 restatement in oracle/oracle.c.  Only the
tests, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py``
use it, always as the checker; the product library never calls it.
"""
from __future__ import annotations

import ctypes as C
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "liboracle.so"
_lib = None

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)


def build() -> Path:
    """Compiles the oracle's C restatement and the threaded CPU path of the bench's baseline."""
    subprocess.run(["make", "-s", "-C", str(HERE), "-B", "liboracle.so", "libcpustep.so"], check=True)
    return LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            build()
        L = C.CDLL(str(LIB_PATH))
        L.or_fft3d.argtypes = [C.c_int] * 5 + [_dp, C.c_int]
        L.or_fft1d.argtypes = [C.c_int, _dp, C.c_int]
        L.or_rhoofr.argtypes = [_ip, _ip, _dp, C.c_double, C.c_double, _dp, C.c_int, _dp, C.c_int,
                                _dp, _dp, _dp, _dp]
        L.or_vpsi.argtypes = [_ip, _ip, _dp, C.c_double, C.c_double, _dp, C.c_int, _dp, _dp, _dp, C.c_int]
        _lib = L
    return _lib


def _p(a, t=_dp):
    return a.ctypes.data_as(t)


def fft1d(x: np.ndarray, sign: int) -> np.ndarray:
    y = np.ascontiguousarray(x, dtype=np.complex128).copy()
    lib().or_fft1d(len(y), _p(y.view(np.float64)), sign)
    return y


def fft3d(x: np.ndarray, sign: int) -> np.ndarray:
    """x indexed [z, y, x] (x fastest in memory), unscaled, sign +1 inverse / -1 forward."""
    y = np.ascontiguousarray(x, dtype=np.complex128).copy()
    n3, n2, n1 = y.shape
    lib().or_fft3d(n1, n2, n3, n1, n2, _p(y.view(np.float64)), sign)
    return y


class Layout:
    """The layout arguments of the oracle functions, from a GSet and Cell."""

    def __init__(self, cell, gset, ngw=None):
        ngw = gset.ngw if ngw is None else ngw
        nr = cell.nr
        kr = cell.kr
        self.ngw = ngw
        self.dims = np.array([nr[0], nr[1], nr[2], kr[0], kr[1], kr[2], ngw, 1], dtype=np.int32)
        nh = np.array([n // 2 + 1 for n in nr], dtype=np.int32)
        self.inyh = np.ascontiguousarray((gset.miller[:ngw] + nh).astype(np.int32))
        self.hg = np.ascontiguousarray(gset.hg[:ngw])
        self.tpiba2 = cell.tpiba2
        self.omega = cell.omega
        self.nnr1 = kr[0] * kr[1] * kr[2]


def rhoofr(L: Layout, c0: np.ndarray, f: np.ndarray):
    c0 = np.asfortranarray(c0, dtype=np.complex128)
    f = np.ascontiguousarray(f, dtype=np.float64)
    rho = np.zeros(L.nnr1)
    ek = C.c_double(); cg = C.c_double(); cr = C.c_double()
    lib().or_rhoofr(_p(L.dims, _ip), _p(L.inyh, _ip), _p(L.hg), L.tpiba2, L.omega,
                    _p(c0 if c0.size else np.zeros(2)), c0.shape[0], _p(f), c0.shape[1],
                    _p(rho), C.byref(ek), C.byref(cg), C.byref(cr))
    return rho, ek.value, cg.value, cr.value


def vpsi(L: Layout, c0: np.ndarray, c2: np.ndarray, f: np.ndarray, vpot: np.ndarray) -> np.ndarray:
    c0 = np.asfortranarray(c0, dtype=np.complex128)
    out = np.array(c2, dtype=np.complex128, order='F', copy=True)
    f = np.ascontiguousarray(f, dtype=np.float64)
    v = np.ascontiguousarray(vpot, dtype=np.float64)
    lib().or_vpsi(_p(L.dims, _ip), _p(L.inyh, _ip), _p(L.hg), L.tpiba2, L.omega,
                  _p(c0), c0.shape[0], _p(out), _p(f), _p(v), c0.shape[1])
    return out


# --------------------------------------------------------------------------
# dense contractions
# --------------------------------------------------------------------------
def _setup_dense(L):
    if getattr(L, "_dense", False):
        return
    L._dense = True
    _ci = C.c_int
    L.or_ovlap.argtypes = [_ci, _ci, _dp, _dp, _ci, _ci, _dp]
    L.or_potrf_upper.argtypes = [_dp, _ci]
    L.or_potrf_upper.restype = _ci
    L.or_trtri_upper.argtypes = [_dp, _ci, _dp]
    L.or_rgs.argtypes = [_ci, _ci, _dp, _ci, _ci, _dp]
    L.or_rgs.restype = _ci
    L.or_jacobi_eigen.argtypes = [_dp, _ci, _dp, _dp]
    L.or_lowdin.argtypes = [_ci, _ci, _dp, _ci, _ci, _dp]
    L.or_rotate.argtypes = [_ci, C.c_double, _dp, C.c_double, _dp, _ci, _dp, _ci]
    L.or_lagrange.argtypes = [_ci, _ci, _dp, _dp, _ci, _dp, _ci]
    L.or_rortho.argtypes = [_ci, _ci, _dp, _dp, _ci, _dp, _ci, _ci, C.c_double, _ci]
    L.or_hesele.argtypes = [_ci, _ci, C.c_double, C.c_double, C.c_double, C.c_double, _dp, _ci, _ip, _ci, _ip, _dp,
                            _dp, _ci, _dp]
    L.or_csize.argtypes = [_ci, _ci, _dp, _ci, _dp, _ci, _dp, _dp]
    L.or_sd_update.argtypes = [_ci, _dp, _dp, _ci, _ci, _dp, C.c_double]
    L.or_rortho.restype = _ci
    L.or_rnlsm1.argtypes = [_ci, _ci, _ci, _ip, _ci, _ip, _ip, _dp, _ci, _dp, _ci, _dp, _ci, _ci, _dp]
    L.or_fnonloc.argtypes = [_ci, _ci, _ip, _ci, _ip, _ip, _dp, _dp, _ci, _dp, _ci, _dp, _dp, _dp, _ci, _ci]
    L.or_rnlsm2.argtypes = [_ci, _ci, C.c_double, _dp, _ci, _ip, _ci, _ip, _ip, _dp, _ci, _dp, _ci, _dp, _ci, _ci,
                            _dp]
    L.or_rnlfor.argtypes = [_ci, _ip, _ci, _ip, _dp, _dp, _dp, _dp, _ci, _dp]
    L.or_enl.argtypes = [_ci, _ip, _ci, _ip, _dp, _dp, _dp, _ci]
    L.or_enl.restype = C.c_double
    L.or_lda_pade.argtypes = [C.c_double, _dp, _dp]
    L.or_xcener.argtypes = [_ip, _ip, _dp, _dp, _dp, _dp]
    L.or_ewald.argtypes = [_ci, _ip, _dp, _dp, _dp, _dp, _ci, _dp, _dp, _dp]
    L.or_eicalc.argtypes = [_ci, _ip, _ip, _ci, _ip, _dp, _dp, _dp, _dp, _dp]
    L.or_vofrho.argtypes = [_ip, _ci, _ip, _dp, _dp, C.c_double, C.c_double, _ci, _ci, _ip, _dp, _dp, _dp, _dp,
                            _dp, _dp, _dp, _dp]


def _L():
    L = lib()
    _setup_dense(L)
    return L


def _F(a, dtype=np.complex128):
    return np.array(a, dtype=dtype, order="F", copy=True)


def ovlap(c1, c2, geq0=True):
    c1 = np.asfortranarray(c1, dtype=np.complex128)
    c2 = np.asfortranarray(c2, dtype=np.complex128)
    n = c1.shape[1]
    a = np.zeros((n, n), order="F")
    _L().or_ovlap(c1.shape[0], int(geq0), _p(c1), _p(c2), c1.shape[0], n, _p(a))
    return a


def potrf_upper(s):
    u = _F(s, np.float64)
    info = _L().or_potrf_upper(_p(u), u.shape[0])
    return u, info


def trtri_upper(u):
    u = _F(u, np.float64)
    x = np.zeros_like(u, order="F")
    _L().or_trtri_upper(_p(u), u.shape[0], _p(x))
    return x


def rgs(c, geq0=True):
    c = _F(c)
    n = c.shape[1]
    smat = np.zeros((n, n), order="F")
    info = _L().or_rgs(c.shape[0], int(geq0), _p(c), c.shape[0], n, _p(smat))
    if info:
        raise np.linalg.LinAlgError(f"oracle rgs: leading minor {info} not positive definite")
    return c, smat


def jacobi_eigen(a):
    a = _F(a, np.float64)
    n = a.shape[0]
    w = np.zeros(n)
    z = np.zeros((n, n), order="F")
    _L().or_jacobi_eigen(_p(a), n, _p(w), _p(z))
    return w, z


def lowdin(c, geq0=True):
    c = _F(c)
    n = c.shape[1]
    amat = np.zeros((n, n), order="F")
    _L().or_lowdin(c.shape[0], int(geq0), _p(c), c.shape[0], n, _p(amat))
    return c, amat


def rotate(a, c1, b, c2, gam):
    c1 = np.asfortranarray(c1, dtype=np.complex128)
    out = _F(c2)
    gam = np.asfortranarray(gam, dtype=np.float64)
    _L().or_rotate(c1.shape[0], a, _p(c1), b, _p(out), c1.shape[0], _p(gam), c1.shape[1])
    return out


def lagrange(c0, c2, f, geq0=True):
    c0 = np.asfortranarray(c0, dtype=np.complex128)
    out = _F(c2)
    f = np.ascontiguousarray(f, dtype=np.float64)
    _L().or_lagrange(c0.shape[0], int(geq0), _p(c0), _p(out), c0.shape[0], _p(f), c0.shape[1])
    return out


def hesele(prec, svar2, vploc, hthrs, tpiba2, hg, P=None):
    """vpp(G) in CPMD G order; P: the projector set (None: no projectors)."""
    hg = np.ascontiguousarray(hg, dtype=np.float64)
    ngw = len(hg)
    vpp = np.zeros(ngw)
    if P is None:
        na = np.zeros(1, np.int32); ngh = np.zeros(1, np.int32)
        hm = np.zeros(1); tw = np.zeros(1); nsp, nhxs, ldtw = 0, 1, 1
    else:
        na = np.ascontiguousarray(P.na, np.int32); ngh = np.ascontiguousarray(P.ngh, np.int32)
        hm = np.asfortranarray(P.hmat, np.float64); tw = np.asfortranarray(P.twnl, np.float64)
        nsp, nhxs, ldtw = len(na), P.nhxs, tw.shape[0]
    _L().or_hesele(ngw, int(prec), svar2, vploc, hthrs, tpiba2, _p(hg), nsp, _p(na, _ip), nhxs, _p(ngh, _ip),
                   _p(hm), _p(tw), ldtw, _p(vpp))
    return vpp


def csize(c2, f, geq0=True):
    """(gemax, cnorm) of the electronic gradient c2 (ngw, n)."""
    c2 = np.asfortranarray(c2, dtype=np.complex128)
    f = np.ascontiguousarray(f, dtype=np.float64)
    ge, cn = C.c_double(), C.c_double()
    _L().or_csize(c2.shape[0], int(geq0), _p(c2), c2.shape[0], _p(f), c2.shape[1], C.byref(ge), C.byref(cn))
    return ge.value, cn.value


def sd_update(c0, c2, vpp, pf1):
    out = _F(c0)
    c2 = np.asfortranarray(c2, dtype=np.complex128)
    vpp = np.ascontiguousarray(vpp, dtype=np.float64)
    _L().or_sd_update(out.shape[0], _p(out), _p(c2), out.shape[0], out.shape[1], _p(vpp), pf1)
    return out


def _dotp_cols(a, b, geq0=True):
    """dotp per column (src/dotp_utils.mod.F90:58-101)."""
    s = 2.0 * (a.real * b.real + a.imag * b.imag).sum(axis=0)
    if geq0:
        s -= a[0].real * b[0].real
    return s
class ODIIS:
    """odiis, src/odiis_utils.mod.F90:492-740 (closed shell, uncompressed
    history): DIIS over the occupied states with the
    preconditioner vpp.  step() returns the new coefficients before the
    orthonormalisation that updwf applies after it."""

    def __init__(self, mdiis=10, nreset=None, prec=True):
        # nreset: CPMD sets it to mdiis for wavefunction optimisation
        # (src/control_test_utils.mod.F90:452-465)
        self.mdiis, self.prec = mdiis, prec
        self.nreset = mdiis if nreset is None else nreset
        self.istate = 0
    def _updis(self, iact):
        # updis, src/odiis_utils.mod.F90:36-58 (1-based nowv)
        if iact == 1:
            self.nowv = self.ndiis % self.mdiis + 1
            self.ndiis += 1
            self.nsize = min(self.ndiis + 1, self.mdiis + 1)
        else:
            self.ndiis = self.nowv = self.nsize = 0
    def step(self, c0, c2, vpp, f, etot, svar2, reinit=False, geq0=True):
        n = c0.shape[1]
        self.reset = False
        if reinit or self.istate == 0:
            self.istate, self.eold, self.gamma = 0, 9999.0, 0.0
            self.nocc = int(np.count_nonzero(np.asarray(f) > 1e-5))
            self.ndiis = 0
            m = self.mdiis
            self.diism = np.zeros((m, m))
            self.grmax, self.gimax, self.gnorm = np.zeros(m), np.zeros(m), np.zeros(m)
            self.pme, self.gde = [None] * m, [None] * m
        nocc = self.nocc
        if self.istate != 0:
            de1 = etot - self.eold
            if geq0:      # :541-551
                self.gamma = self.gamma + 1.0 / vpp[0] if de1 > 0.0 else 0.75 * self.gamma
            if self.ndiis > self.nreset and self.nreset >= 3:      # :553-569
                nowm = 1 if self.nowv == self.mdiis else self.nowv + 1
                g1 = np.sqrt(self.grmax[self.nowv - 1] ** 2 + self.gimax[self.nowv - 1] ** 2)
                g2 = np.sqrt(self.grmax[nowm - 1] ** 2 + self.gimax[nowm - 1] ** 2)
                raim = 1.0 - self.mdiis / 100.0
                if g1 / g2 > raim and self.gnorm[self.nowv - 1] / self.gnorm[nowm - 1] > raim:
                    self.istate, self.gamma = 0, 0.0
                    self.reset = True          # odiis returns reinit = .TRUE. (:584-593)
        if self.istate == 0:
            self._updis(0)
        self.istate = 2
        self._updis(1)
        w = self.nowv - 1
        v = vpp / (1.0 + vpp * self.gamma) if self.gamma != 0.0 else vpp
        g = -c2[:, :nocc]
        self.pme[w] = c0[:, :nocc].copy()
        self.gde[w] = g.copy()
        # grimax, src/odiis_utils.mod.F90:408-450
        self.grmax[w] = np.abs(g.real).max()
        gi = np.abs(g.imag).max()
        self.gimax[w] = gi if gi != 0.0 else np.finfo(np.float64).eps
        self.gnorm[w] = np.sqrt(_dotp_cols(g, g, geq0).sum() / (nocc * c0.shape[0]))
        ff = np.where(self.prec & (np.asarray(f[:nocc]) > 0.1), 1.0 / np.asarray(f[:nocc]), 1.0)
        sg = g * (2.0 * ff * ff)[None, :] * (v * v)[:, None]
        ns = self.nsize
        for i in range(ns - 1):
            self.diism[i, w] = _dotp_cols(self.gde[i], sg, geq0).sum()
        for i in range(ns - 1):
            self.diism[w, i] = self.diism[i, w]
        bc = np.zeros((ns, ns))
        bc[:ns - 1, :ns - 1] = self.diism[:ns - 1, :ns - 1]
        bc[:ns - 1, ns - 1] = bc[ns - 1, :ns - 1] = -1.0
        vc = np.zeros(ns); vc[ns - 1] = -1.0
        # solve: dgelss, minimum-norm least squares, rcond 0.2e-15 (:60-89)
        vc = np.linalg.lstsq(bc, vc, rcond=0.2e-15)[0]
        out = c0.copy()
        cn = np.zeros_like(c0[:, :nocc])
        for i in range(ns - 1):
            cn += vc[i] * self.pme[i]
        for i in range(ns - 1):
            cn -= vc[i] * ff[None, :] * v[:, None] * self.gde[i]
        out[:, :nocc] = cn
        if nocc < n:
            out[:, nocc:] += svar2 * c2[:, nocc:]
        self.eold = etot
        return out


def rortho(c0, cm, tnon=False, epsog=1e-8, maxit=30, geq0=True):
    c0 = np.asfortranarray(c0, dtype=np.complex128)
    out = _F(cm)
    n = c0.shape[1]
    gam = np.zeros((n, n), order="F")
    it = _L().or_rortho(c0.shape[0], int(geq0), _p(c0), _p(out), c0.shape[0], _p(gam), int(tnon), n, epsog, maxit)
    return out, gam, it


def rortv(c0, cm, geq0=True):
    """rortv, src/rortv_utils.mod.F90:149-160 (overlap constraints, no tharm /
    tmass / tdmal): gamy = ovlap(cm, c0); gamy(i,j) = gamy(j,i) =
    -(gamy(i,j) + gamy(j,i))/2; cm += c0 gamy.  Returns (cm, gamy)."""
    a = ovlap(cm, c0, geq0)
    gamy = -0.5 * (a + a.T)
    return rotate(1.0, c0, 1.0, cm, gamy), np.asfortranarray(gamy)


def rekine(cm, emass, geq0=True):
    """rekine, src/rekine_utils.mod.F90:46-50 (no tmass): emass sum_i dotp(cm_i, cm_i)."""
    cm = np.asarray(cm, dtype=np.complex128)
    return emass * float(_dotp_cols(cm, cm, geq0).sum())


class Projectors:
    """Arguments of the projector routines (CPMD's nlps/sgpp/cppt data)."""

    def __init__(self, na, nhxs, ngh, lang, hmat, twnl, eigr):
        self.na = np.ascontiguousarray(na, dtype=np.int32)
        self.nsp = len(self.na)
        self.nhxs = int(nhxs)
        self.ngh = np.ascontiguousarray(ngh, dtype=np.int32)
        self.lang = np.asfortranarray(np.asarray(lang, dtype=np.int32).reshape(self.nhxs, self.nsp, order="F"))
        self.hmat = np.asfortranarray(np.asarray(hmat, dtype=np.float64).reshape(self.nhxs, self.nhxs, self.nsp, order="F"))
        self.twnl = np.asfortranarray(np.asarray(twnl, dtype=np.float64))      # (ngw, nhxs, nsp)
        self.eigr = np.asfortranarray(np.asarray(eigr, dtype=np.complex128))   # (ngw, nat)
        self.nat = int(self.na.sum())


def rnlsm1(P: Projectors, c0, geq0=True):
    c0 = np.asfortranarray(c0, dtype=np.complex128)
    n = c0.shape[1]
    fnl = np.zeros((P.nat, P.nhxs, n), order="F")
    _L().or_rnlsm1(c0.shape[0], int(geq0), P.nsp, _p(P.na, _ip), P.nhxs, _p(P.ngh, _ip), _p(P.lang, _ip),
                   _p(P.twnl), P.twnl.shape[0], _p(P.eigr), P.eigr.shape[0], _p(c0), c0.shape[0], n, _p(fnl))
    return fnl


def fnonloc(P: Projectors, c2, f, fnl):
    out = _F(c2)
    f = np.ascontiguousarray(f, dtype=np.float64)
    fnl = np.asfortranarray(fnl, dtype=np.float64)
    _L().or_fnonloc(out.shape[0], P.nsp, _p(P.na, _ip), P.nhxs, _p(P.ngh, _ip), _p(P.lang, _ip), _p(P.hmat),
                    _p(P.twnl), P.twnl.shape[0], _p(P.eigr), P.eigr.shape[0], _p(fnl), _p(f), _p(out), out.shape[0],
                    out.shape[1])
    return out


def rnlsm2(P: Projectors, c0, gk, tpiba, geq0=True):
    """dfnl (nat, nhxs, 3, nstate); gk (3, ngw) in units of 2 pi / alat."""
    c0 = np.asfortranarray(c0, dtype=np.complex128)
    gk = np.asfortranarray(gk, dtype=np.float64)
    n = c0.shape[1]
    dfnl = np.zeros((P.nat, P.nhxs, 3, n), order="F")
    _L().or_rnlsm2(c0.shape[0], int(geq0), float(tpiba), _p(gk), P.nsp, _p(P.na, _ip), P.nhxs, _p(P.ngh, _ip),
                   _p(P.lang, _ip), _p(P.twnl), P.twnl.shape[0], _p(P.eigr), P.eigr.shape[0], _p(c0), c0.shape[0], n,
                   _p(dfnl))
    return dfnl


def rnlfor(P: Projectors, fnl, dfnl, f, fion=None):
    """fion (3, nat) -= the nonlocal forces (rnlfor accumulates)."""
    fnl = np.asfortranarray(fnl, dtype=np.float64)
    dfnl = np.asfortranarray(dfnl, dtype=np.float64)
    f = np.ascontiguousarray(f, dtype=np.float64)
    out = np.zeros((3, P.nat), order="F") if fion is None else _F(fion, np.float64)
    _L().or_rnlfor(P.nsp, _p(P.na, _ip), P.nhxs, _p(P.ngh, _ip), _p(P.hmat), _p(fnl), _p(dfnl), _p(f), len(f),
                   _p(out))
    return out


def enl(P: Projectors, fnl, f):
    fnl = np.asfortranarray(fnl, dtype=np.float64)
    f = np.ascontiguousarray(f, dtype=np.float64)
    return _L().or_enl(P.nsp, _p(P.na, _ip), P.nhxs, _p(P.ngh, _ip), _p(P.hmat), _p(fnl), _p(f), len(f))


# --------------------------------------------------------------------------
# local potential / Hartree / LDA / Ewald
# --------------------------------------------------------------------------
VOFRHO_KEYS = ("ehep", "ehii", "ehee", "epseu", "vploc", "exc", "vxc")


def lda_pade(rho):
    """(eps_xc, V_xc) per point, Goedecker-Teter-Hutter Pade LDA."""
    rho = np.atleast_1d(np.asarray(rho, dtype=np.float64))
    e = np.empty_like(rho)
    v = np.empty_like(rho)
    L = _L()
    for i, r in enumerate(rho):
        a = C.c_double(); b = C.c_double()
        L.or_lda_pade(float(r), C.byref(a), C.byref(b))
        e[i] = a.value
        v[i] = b.value
    return e, v


def ewald(na, zv, raggio, tau, ht, iesr=2, forces=True):
    """(esr, eself, fion (3, nat)); tau (nat, 3) bohr, ht rows = lattice vectors."""
    na = np.ascontiguousarray(na, dtype=np.int32)
    zv = np.ascontiguousarray(zv, dtype=np.float64)
    rg = np.ascontiguousarray(raggio, dtype=np.float64)
    t = np.ascontiguousarray(np.asarray(tau, dtype=np.float64))          # (nat, 3) C = (3, nat) F
    h = np.asfortranarray(np.asarray(ht, dtype=np.float64))
    fion = np.zeros((3, int(na.sum())), order="F")
    esr = C.c_double(); es = C.c_double()
    _L().or_ewald(len(na), _p(na, _ip), _p(zv), _p(rg), _p(t), _p(h), int(iesr), C.byref(esr), C.byref(es),
                  _p(fion) if forces else None)
    return esr.value, es.value, fion


def vofrho(L: Layout, inyh, hg, gk, na, vps, rhops, frac, rho, geq0=True, forces=True):
    """vofrhoa + vofrhob (closed shell, LDA Pade, no core/gradient correction).

    inyh (nhg, 3) centred 1-based indices, hg (nhg,), gk (nhg, 3) units 2 pi/alat,
    vps / rhops (nsp, nhg), frac (nat, 3).  Returns (v(r), energies dict,
    fion (3, nat) from potfor, rho(G))."""
    inyh = np.ascontiguousarray(inyh, dtype=np.int32)
    hg = np.ascontiguousarray(hg, dtype=np.float64)
    gk = np.ascontiguousarray(gk, dtype=np.float64)
    na = np.ascontiguousarray(na, dtype=np.int32)
    vps = np.asfortranarray(vps, dtype=np.float64)
    rhops = np.asfortranarray(rhops, dtype=np.float64)
    frac = np.ascontiguousarray(frac, dtype=np.float64)
    rho = np.ascontiguousarray(rho, dtype=np.float64)
    nhg = hg.shape[0]
    v = np.zeros(L.nnr1)
    out = np.zeros(7)
    fion = np.zeros((3, int(na.sum())), order="F")
    rhog = np.zeros(nhg, dtype=np.complex128)
    _L().or_vofrho(_p(L.dims, _ip), nhg, _p(inyh, _ip), _p(hg), _p(gk), L.tpiba2, L.omega, int(geq0), len(na),
                   _p(na, _ip), _p(vps), _p(rhops), _p(frac), _p(rho), _p(v), _p(out),
                   _p(fion) if forces else None, _p(rhog))
    return v, dict(zip(VOFRHO_KEYS, out)), fion, rhog


# --------------------------------------------------------------------------
# Kohn-Sham energy, gradient and forces composed from the routines above
# (the Gamma, closed-shell, LDA path of CPMD's forces driver,
# src/forces_driver.mod.F90:175-384)
# --------------------------------------------------------------------------
class KSSystem:
    """Everything the KS energy needs besides C: layout, density set, local and
    nonlocal pseudopotential data and the ions."""

    def __init__(self, L, inyh, hg, gk, na, zv, raggio, vps, rhops, avec, bvec, tpiba, miller_w, nl_args,
                 iesr=2):
        self.L = L
        self.inyh, self.hg, self.gk = inyh, hg, gk
        self.na, self.zv, self.raggio, self.vps, self.rhops = na, zv, raggio, vps, rhops
        self.avec = np.asarray(avec, dtype=np.float64)
        self.miller_w = miller_w          # (ngw, 3) wavefunction G set, for eigr
        self.gkw = np.asfortranarray((miller_w.astype(np.float64) @ np.asarray(bvec)).T)   # (3, ngw), 2 pi / alat
        self.tpiba = float(tpiba)
        self.nl = nl_args                 # dict(nhxs, ngh, lang, hmat, twnl)
        self.iesr = iesr

    def projectors(self, tau):
        frac = tau @ np.linalg.inv(self.avec)
        eigr = np.exp(-2j * np.pi * (self.miller_w.astype(np.float64) @ frac.T))
        return Projectors(na=self.na, eigr=eigr, **self.nl)


def ks_energy(K: KSSystem, c, f, tau, forces=False, gradient=False):
    """E_KS = ekin + ehep + epseu + exc + esr - eself + enl (ener_com%etot of
    CPMD for this path).  gradient: c2 = -f H c (vpsi + fnonloc); forces: fion
    (3, nat) = potfor + Ewald + rnlfor."""
    frac = tau @ np.linalg.inv(K.avec)
    rho, ekin, _, _ = rhoofr(K.L, c, f)
    v, en, floc, _ = vofrho(K.L, K.inyh, K.hg, K.gk, K.na, K.vps, K.rhops, frac, rho, forces=forces)
    esr, eself, fesr = ewald(K.na, K.zv, K.raggio, tau, K.avec, iesr=K.iesr, forces=forces)
    P = K.projectors(tau)
    fnl = rnlsm1(P, c)
    e_nl = enl(P, fnl, f)
    out = {"ekin": ekin, "ehep": en["ehep"], "epseu": en["epseu"], "exc": en["exc"], "esr": esr, "eself": eself,
           "enl": e_nl, "v": v, "rho": rho}
    out["etot"] = ekin + en["ehep"] + en["epseu"] + en["exc"] + esr - eself + e_nl
    if gradient:
        c2 = vpsi(K.L, c, np.zeros_like(c), f, v)
        out["c2"] = fnonloc(P, c2, f, fnl)
    if forces:
        dfnl = rnlsm2(P, c, K.gkw, K.tpiba)
        out["fion"] = floc + fesr + rnlfor(P, fnl, dfnl, f)
    return out
