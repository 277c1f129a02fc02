"""Host-side mirror of CPMD's hot-path routines over the micpmd C ABI.

``Context`` holds one plane-wave layout on one GPU (what CPMD keeps in the
``fft``/``cppt``/``system`` modules) and exposes the routines with CPMD's names
and argument meaning:

* ``rhoofr(c0, f)``            -- src/rhoofr_utils.mod.F90:122 (returns rhoe, ekin, csumg, csumr)
* ``vpsi(c0, c2, f, vpot)``    -- src/vpsi_utils.mod.F90:120  (c2 += local H psi force)

Arrays are numpy, complex128 wavefunctions (ngw, nstate) in CPMD G order and
real-space float64 arrays of length kr1*kr2*kr3.  ``*_dev`` variants take
``DeviceArray`` wavefunctions already resident on the GPU in the library's
layout (see ``Context.upload``).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from ._lib import check, lib

_dp = C.POINTER(C.c_double)


def _p(a):
    return a.ctypes.data_as(_dp)


class DeviceArray:
    """A hipMalloc'ed buffer owned by Python."""

    def __init__(self, nbytes: int):
        self.nbytes = int(nbytes)
        p = C.c_void_p()
        check(lib().mc_dev_alloc(self.nbytes, C.byref(p)))
        self.ptr = p

    @property
    def value(self) -> int:
        return self.ptr.value or 0

    def zero(self):
        check(lib().mc_memset_d(self.ptr, 0, self.nbytes))
        return self

    def from_host(self, a: np.ndarray):
        a = np.ascontiguousarray(a)
        assert a.nbytes <= self.nbytes
        check(lib().mc_memcpy_h2d(self.ptr, a.ctypes.data_as(C.c_void_p), a.nbytes))
        return self

    def to_host(self, shape, dtype=np.float64) -> np.ndarray:
        out = np.empty(shape, dtype=dtype)
        assert out.nbytes <= self.nbytes
        check(lib().mc_memcpy_d2h(out.ctypes.data_as(C.c_void_p), self.ptr, out.nbytes))
        return out

    def free(self):
        if self.ptr is not None and self.ptr.value:
            lib().mc_dev_free(self.ptr)
        self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def set_mem_limit(nbytes: int):
    """Cap the device memory the library holds (0: no cap); see mc_set_mem_limit."""
    check(lib().mc_set_mem_limit(int(nbytes)))


def mem_stats() -> tuple[int, int, int]:
    """(bytes held now, most held since the last reset, the cap)."""
    cur, peak, lim = C.c_size_t(), C.c_size_t(), C.c_size_t()
    check(lib().mc_mem_stats(C.byref(cur), C.byref(peak), C.byref(lim)))
    return cur.value, peak.value, lim.value


def mem_reset_peak():
    check(lib().mc_mem_reset_peak())


MFMA_PROBES = {"bf16_16x16x32": 0, "bf16_32x32x16": 1, "f64_16x16x4": 2, "f32_16x16x4": 3}


def probe_mfma(kind: str) -> float:
    """TFLOP/s one MFMA instruction reaches on this GPU (see mc_probe_mfma)."""
    r = C.c_double()
    check(lib().mc_probe_mfma(MFMA_PROBES[kind], C.byref(r)))
    return r.value


def make_layout(cell, gset, ngw: int | None = None):
    """mc_layout plus the arrays it points to (kept alive by the returned tuple)."""
    ngw = gset.ngw if ngw is None else ngw
    nh = np.array([n // 2 + 1 for n in cell.nr], dtype=np.int32)
    inyh = np.ascontiguousarray((gset.miller[:ngw] + nh).astype(np.int32))
    hg = np.ascontiguousarray(gset.hg[:ngw], dtype=np.float64)
    lay = _lib.mc_layout(cell.nr[0], cell.nr[1], cell.nr[2], cell.kr[0], cell.kr[1], cell.kr[2], ngw,
                         inyh.ctypes.data_as(C.POINTER(C.c_int)), hg.ctypes.data_as(_dp),
                         cell.tpiba2, cell.omega, 1)
    return lay, inyh, hg


def layout_stats(cell, gset, ngw: int | None = None):
    """(sticks, x-frequency columns, owner sticks) of the library's stick layout (host only)."""
    lay, _keep1, _keep2 = make_layout(cell, gset, ngw)
    s, c, o = C.c_int(), C.c_int(), C.c_int()
    check(lib().mc_layout_stats(C.byref(lay), C.byref(s), C.byref(c), C.byref(o)))
    return s.value, c.value, o.value


class Context:
    def __init__(self, cell, gset, ngw: int | None = None):
        self.cell = cell
        self.ngw = gset.ngw if ngw is None else ngw
        lay, self._inyh, self._hg = make_layout(cell, gset, self.ngw)
        h = C.c_void_p()
        check(lib().mc_ctx_create(C.byref(lay), C.byref(h)))
        self.h = h
        self.ld = int(lib().mc_wfn_ld(self.h))
        kr = cell.kr
        self.nnr1 = kr[0] * kr[1] * kr[2]

    def close(self):
        if getattr(self, "h", None) is not None and self.h.value:
            lib().mc_ctx_destroy(self.h)
        self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ------------------------------------------------------------ settings
    def set_batch(self, pairs_rhoofr: int, pairs_vpsi: int):
        check(lib().mc_ctx_set_batch(self.h, pairs_rhoofr, pairs_vpsi))

    def set_column_reuse(self, on: bool = True):
        """rhoofr_dev keeps the y-transformed columns of every state pair; the
        next vpsi_dev / forces_c2_dev on the same wavefunction tensor starts
        from them (bit-identical).  The caller must not write that tensor
        outside the library between the two calls."""
        check(lib().mc_ctx_set_column_reuse(self.h, 1 if on else 0))

    def set_precision(self, mode: str):
        """'fp64' (default), 'fp32' (FP32 rhoofr/vpsi transforms), 'mixed' (FP32
        transforms and split-bf16 MFMA wavefunction contractions) or
        'mixed_bf16x3' (mixed with every split product 3-term)."""
        check(lib().mc_set_precision(self.h, {"fp64": 0, "fp32": 1, "mixed": 2, "mixed_bf16x3": 3}[mode]))

    def synchronize(self):
        check(lib().mc_ctx_synchronize(self.h))

    # ------------------------------------------------------------ transfers
    def alloc_wfn(self, nstate: int) -> DeviceArray:
        return DeviceArray(16 * self.ld * max(nstate, 1)).zero()

    def zero_wfn(self, dev: DeviceArray, nstate: int):
        check(lib().mc_wfn_zero(self.h, dev.ptr, nstate))

    def axpy_wfn(self, alpha: float, x: DeviceArray, y: DeviceArray, nstate: int):
        check(lib().mc_wfn_axpy(self.h, alpha, x.ptr, y.ptr, nstate))

    def axpby_wfn(self, alpha: float, x: DeviceArray, beta: float, y: DeviceArray, nstate: int):
        """y = alpha x + beta y."""
        check(lib().mc_wfn_axpby(self.h, alpha, x.ptr, beta, y.ptr, nstate))

    def cp_posupa(self, c0: DeviceArray, c2: DeviceArray, cm: DeviceArray, cp: DeviceArray, nstate: int,
                  dtm: float, dt: float):
        """cm += dtm c2; cp = c0 + dt cm (velupa + posupa before rortog)."""
        check(lib().mc_cp_posupa(self.h, c0.ptr, c2.ptr, cm.ptr, cp.ptr, nstate, dtm, dt))

    def posupa_dev(self, c0: DeviceArray, c2: DeviceArray | None, cm: DeviceArray, cp: DeviceArray, nstate: int,
                   dtm: float, dt: float, epsog: float = 1e-8, maxit: int = 60, want_gam: bool = False):
        """cm += dtm c2; cp = c0 + dt cm; rortog(c0, cp); cm = (cp - c0)/dt.
        Returns (gam or None, rortog iterations)."""
        gam = np.zeros((nstate, nstate), order="F") if want_gam else None
        it = C.c_int(0)
        check(lib().mc_posupa_dev(self.h, c0.ptr, c2.ptr if c2 is not None else None, cm.ptr, cp.ptr, nstate, dtm,
                                  dt, epsog, maxit, _p(gam) if want_gam else None, C.byref(it)))
        return gam, it.value

    def cp_velocity(self, c0: DeviceArray, cp: DeviceArray, cm: DeviceArray, nstate: int, dt: float):
        """cm = (cp - c0) / dt (velocities after rortog)."""
        check(lib().mc_cp_velocity(self.h, c0.ptr, cp.ptr, cm.ptr, nstate, dt))

    def rortv(self, c0: np.ndarray, cm: np.ndarray, want_gam: bool = False):
        """rortv on host arrays: returns (cm after the velocity constraint, gamy or None)."""
        c0 = np.asfortranarray(c0, dtype=np.complex128)
        out = np.array(cm, dtype=np.complex128, order="F", copy=True)
        n = c0.shape[1]
        gam = np.zeros((n, n), order="F") if want_gam else None
        check(lib().mc_rortv(self.h, _p(c0), _p(out), c0.shape[0], _p(gam) if want_gam else None, n))
        return out, gam

    def rortv_dev(self, c0: DeviceArray, cm: DeviceArray, nstate: int, want_gam: bool = False):
        """cm += c0 gamy with gamy = -sym(<cm|c0>) (the CP velocity constraint); returns gamy or None."""
        gam = np.zeros((nstate, nstate), order="F") if want_gam else None
        check(lib().mc_rortv_dev(self.h, c0.ptr, cm.ptr, _p(gam) if want_gam else None, nstate))
        return gam

    def rekine_dev(self, cm: DeviceArray, nstate: int, emass: float) -> float:
        """ekinc = emass sum_i dotp(cm_i, cm_i)."""
        e = C.c_double()
        check(lib().mc_rekine_dev(self.h, cm.ptr, nstate, emass, C.byref(e)))
        return e.value

    def cp_velupa_rortv_dev(self, c0: DeviceArray, c2: DeviceArray | None, cm: DeviceArray, nstate: int,
                            dtm: float, emass: float, want_gam: bool = False):
        """End of the velocity Verlet CP step: cm += dtm c2, rortv, rekine.  Returns (ekinc, gamy or None)."""
        gam = np.zeros((nstate, nstate), order="F") if want_gam else None
        e = C.c_double()
        check(lib().mc_cp_velupa_rortv_dev(self.h, c0.ptr, c2.ptr if c2 is not None else None, cm.ptr, nstate, dtm,
                                           emass, _p(gam) if want_gam else None, C.byref(e)))
        return e.value, gam

    def hesele(self, prec: bool, svar2: float = 0.0, vploc: float = 0.0, hthrs: float = 0.5):
        """Set the diagonal preconditioner vpp(G) of the electronic gradient."""
        check(lib().mc_hesele(self.h, int(prec), svar2, vploc, hthrs))

    def csize_dev(self, c2: DeviceArray, f: np.ndarray, nstate: int) -> tuple[float, float]:
        """(gemax, cnorm) of the gradient c2 over the occupied states."""
        f = np.ascontiguousarray(f, dtype=np.float64)
        ge, cn = C.c_double(), C.c_double()
        check(lib().mc_csize_dev(self.h, c2.ptr, _p(f), nstate, C.byref(ge), C.byref(cn)))
        return ge.value, cn.value

    def sd_update_dev(self, c0: DeviceArray, c2: DeviceArray, nstate: int, pf1: float):
        """c0 += vpp pf1 c2 (preconditioned steepest descent)."""
        check(lib().mc_sd_update_dev(self.h, c0.ptr, c2.ptr, nstate, pf1))

    def odiis_dev(self, c0: DeviceArray, c2: DeviceArray, f: np.ndarray, nstate: int, etot: float, svar2: float,
                  mdiis: int = 10, reinit: bool = False, nreset: int | None = None):
        """One DIIS step of the occupied states (CPMD's odiis); the history stays in the context.
        nreset defaults to mdiis, CPMD's setting for wavefunction optimisation.  Returns True
        when the step restarted the history for insufficient progress (odiis's reinit)."""
        f = np.ascontiguousarray(f, dtype=np.float64)
        r = C.c_int(int(reinit))
        nr = mdiis if nreset is None else nreset
        check(lib().mc_odiis_dev(self.h, c0.ptr, c2.ptr, _p(f), nstate, etot, svar2, mdiis, nr, C.byref(r)))
        return bool(r.value)

    def odiis_release(self):
        """Free the ODIIS history held by the context."""
        check(lib().mc_odiis_release(self.h))

    def dotp_wfn(self, a: DeviceArray, b: DeviceArray, nstate: int) -> np.ndarray:
        out = np.zeros(nstate)
        check(lib().mc_wfn_dotp(self.h, a.ptr, b.ptr, nstate, _p(out)))
        return out

    def upload(self, c: np.ndarray, dev: DeviceArray | None = None) -> DeviceArray:
        c = np.asfortranarray(c, dtype=np.complex128)
        n = c.shape[1]
        dev = dev or self.alloc_wfn(n)
        check(lib().mc_wfn_upload(self.h, _p(c), c.shape[0], n, dev.ptr))
        return dev

    def download(self, dev: DeviceArray, nstate: int) -> np.ndarray:
        out = np.zeros((self.ngw, nstate), dtype=np.complex128, order="F")
        check(lib().mc_wfn_download(self.h, dev.ptr, nstate, _p(out), self.ngw))
        return out

    # ------------------------------------------------------------ rhoofr
    def rhoofr(self, c0: np.ndarray, f: np.ndarray):
        c0 = np.asfortranarray(c0, dtype=np.complex128)
        f = np.ascontiguousarray(f, dtype=np.float64)
        rho = np.zeros(self.nnr1)
        ek, cg, cr = C.c_double(), C.c_double(), C.c_double()
        nstate = c0.shape[1]
        buf = c0 if c0.size else np.zeros(2)
        check(lib().mc_rhoofr(self.h, _p(buf), c0.shape[0], _p(f if f.size else np.zeros(1)), nstate,
                              _p(rho), C.byref(ek), C.byref(cg), C.byref(cr)))
        return rho, ek.value, cg.value, cr.value

    def rhoofr_dev(self, c0: DeviceArray, f: np.ndarray, nstate: int, rho: DeviceArray):
        f = np.ascontiguousarray(f, dtype=np.float64)
        ek, cg, cr = C.c_double(), C.c_double(), C.c_double()
        check(lib().mc_rhoofr_dev(self.h, c0.ptr, _p(f), nstate, rho.ptr, C.byref(ek), C.byref(cg),
                                  C.byref(cr)))
        return ek.value, cg.value, cr.value

    def rhoofr_raw_charge(self) -> tuple[float, float]:
        """(csumg, csumr before the FP32 rescale) of the last rhoofr."""
        g, r = C.c_double(), C.c_double()
        check(lib().mc_rhoofr_raw_charge(self.h, C.byref(g), C.byref(r)))
        return g.value, r.value

    # ------------------------------------------------------------ vpsi
    def vpsi(self, c0: np.ndarray, c2: np.ndarray, f: np.ndarray, vpot: np.ndarray) -> np.ndarray:
        """c2 += vpsi force (returned; c2 is updated in place when it is a Fortran-ordered complex128)."""
        c0 = np.asfortranarray(c0, dtype=np.complex128)
        out = np.asfortranarray(c2, dtype=np.complex128)
        f = np.ascontiguousarray(f, dtype=np.float64)
        v = np.ascontiguousarray(vpot, dtype=np.float64)
        assert v.size >= self.nnr1
        check(lib().mc_vpsi(self.h, _p(c0), c0.shape[0], _p(out),
                            _p(f), _p(v), c0.shape[1]))
        return out

    def vpsi_dev(self, c0: DeviceArray, c2: DeviceArray, f: np.ndarray, vpot: DeviceArray, nstate: int):
        f = np.ascontiguousarray(f, dtype=np.float64)
        check(lib().mc_vpsi_dev(self.h, c0.ptr, c2.ptr, _p(f), vpot.ptr, nstate))

    # ------------------------------------------------------------ contractions
    def ovlap(self, c1: np.ndarray, c2: np.ndarray) -> np.ndarray:
        c1 = np.asfortranarray(c1, dtype=np.complex128)
        c2 = c1 if c2 is c1 else np.asfortranarray(c2, dtype=np.complex128)
        n = c1.shape[1]
        a = np.zeros((n, n), order="F")
        check(lib().mc_ovlap(self.h, n, _p(a), _p(c1), _p(c2), c1.shape[0]))
        return a

    def ovlap_dev(self, c1: DeviceArray, c2: DeviceArray, nstate: int) -> np.ndarray:
        a = np.zeros((nstate, nstate), order="F")
        check(lib().mc_ovlap_dev(self.h, nstate, _p(a), c1.ptr, c2.ptr))
        return a

    def rgs(self, cp: np.ndarray):
        """Returns (orthonormalised copy of cp, U^-1)."""
        out = np.array(cp, dtype=np.complex128, order="F", copy=True)
        n = out.shape[1]
        smat = np.zeros((n, n), order="F")
        check(lib().mc_rgs(self.h, _p(out), out.shape[0], n, _p(smat)))
        return out, smat

    def rgs_dev(self, c: DeviceArray, nstate: int, want_smat: bool = False):
        smat = np.zeros((nstate, nstate), order="F") if want_smat else None
        check(lib().mc_rgs_dev(self.h, c.ptr, nstate, _p(smat) if want_smat else None))
        return smat

    def lowdin(self, cp: np.ndarray) -> np.ndarray:
        out = np.array(cp, dtype=np.complex128, order="F", copy=True)
        check(lib().mc_lowdin(self.h, _p(out), out.shape[0], out.shape[1]))
        return out

    def lowdin_dev(self, c: DeviceArray, nstate: int) -> int:
        it = C.c_int()
        check(lib().mc_lowdin_dev(self.h, c.ptr, nstate, C.byref(it)))
        return it.value

    def rotate(self, a: float, c1: np.ndarray, b: float, c2: np.ndarray, gam: np.ndarray) -> np.ndarray:
        c1 = np.asfortranarray(c1, dtype=np.complex128)
        out = np.array(c2, dtype=np.complex128, order="F", copy=True)
        gam = np.asfortranarray(gam, dtype=np.float64)
        check(lib().mc_rotate(self.h, a, _p(c1), b, _p(out), c1.shape[0], _p(gam), c1.shape[1]))
        return out

    def rotate_dev(self, a: float, c1: DeviceArray, b: float, c2: DeviceArray, gam: np.ndarray, nstate: int):
        gam = np.asfortranarray(gam, dtype=np.float64)
        check(lib().mc_rotate_dev(self.h, a, c1.ptr, b, c2.ptr, _p(gam), nstate))

    def lagrange(self, c0: np.ndarray, c2: np.ndarray, f: np.ndarray) -> np.ndarray:
        c0 = np.asfortranarray(c0, dtype=np.complex128)
        out = np.array(c2, dtype=np.complex128, order="F", copy=True)
        f = np.ascontiguousarray(f, dtype=np.float64)
        check(lib().mc_lagrange(self.h, _p(c0), _p(out), c0.shape[0], _p(f), c0.shape[1]))
        return out

    def lagrange_dev(self, c0: DeviceArray, c2: DeviceArray, f: np.ndarray, nstate: int):
        f = np.ascontiguousarray(f, dtype=np.float64)
        check(lib().mc_lagrange_dev(self.h, c0.ptr, c2.ptr, _p(f), nstate))

    def rortog(self, c0: np.ndarray, cm: np.ndarray, tnon: bool = False, epsog: float = 1e-8, maxit: int = 30):
        c0 = np.asfortranarray(c0, dtype=np.complex128)
        out = np.array(cm, dtype=np.complex128, order="F", copy=True)
        n = c0.shape[1]
        gam = np.zeros((n, n), order="F")
        it = C.c_int()
        check(lib().mc_rortog(self.h, _p(c0), _p(out), c0.shape[0], _p(gam), int(tnon), n, epsog, maxit,
                              C.byref(it)))
        return out, gam, it.value

    def rortog_dev(self, c0: DeviceArray, cm: DeviceArray, nstate: int, tnon: bool = False, epsog: float = 1e-8,
                   maxit: int = 30):
        gam = np.zeros((nstate, nstate), order="F")
        it = C.c_int()
        check(lib().mc_rortog_dev(self.h, c0.ptr, cm.ptr, _p(gam), int(tnon), nstate, epsog, maxit, C.byref(it)))
        return gam, it.value

    # ------------------------------------------------------------ projectors
    def set_projectors(self, na, nhxs, ngh, lang, hmat, twnl):
        self._na = np.ascontiguousarray(na, dtype=np.int32)
        self._ngh = np.ascontiguousarray(ngh, dtype=np.int32)
        self._lang = np.asfortranarray(np.asarray(lang, dtype=np.int32))
        self._hmat = np.asfortranarray(np.asarray(hmat, dtype=np.float64))
        tw = np.asfortranarray(np.asarray(twnl, dtype=np.float64))
        self.nat = int(self._na.sum())
        self.nhxs = int(nhxs)
        check(lib().mc_set_projectors(self.h, len(self._na), self._na.ctypes.data_as(C.POINTER(C.c_int)), nhxs,
                                      self._ngh.ctypes.data_as(C.POINTER(C.c_int)),
                                      self._lang.ctypes.data_as(C.POINTER(C.c_int)), _p(self._hmat), _p(tw),
                                      tw.shape[0]))

    def set_eigr(self, eigr: np.ndarray):
        e = np.asfortranarray(eigr, dtype=np.complex128)
        check(lib().mc_set_eigr(self.h, _p(e), e.shape[0], e.shape[1]))

    def rnlsm1(self, c0: np.ndarray) -> np.ndarray:
        c0 = np.asfortranarray(c0, dtype=np.complex128)
        n = c0.shape[1]
        fnl = np.zeros((self.nat, self.nhxs, n), order="F")
        check(lib().mc_rnlsm1(self.h, _p(c0), c0.shape[0], n, _p(fnl)))
        return fnl

    def rnlsm1_dev(self, c0: DeviceArray, nstate: int, want_fnl: bool = False):
        fnl = np.zeros((self.nat, self.nhxs, nstate), order="F") if want_fnl else None
        check(lib().mc_rnlsm1_dev(self.h, c0.ptr, nstate, _p(fnl) if want_fnl else None))
        return fnl

    def set_gk(self, gk: np.ndarray):
        """gk (3, ngw) in units of 2 pi / alat, CPMD G order."""
        self._gk = np.asfortranarray(gk, dtype=np.float64)
        check(lib().mc_set_gk(self.h, _p(self._gk)))

    def rnlsm2(self, c0: np.ndarray) -> np.ndarray:
        c0 = np.asfortranarray(c0, dtype=np.complex128)
        n = c0.shape[1]
        dfnl = np.zeros((self.nat, self.nhxs, 3, n), order="F")
        check(lib().mc_rnlsm2(self.h, _p(c0), c0.shape[0], n, _p(dfnl)))
        return dfnl

    def rnlsm2_dev(self, c0: DeviceArray, nstate: int, want_dfnl: bool = False):
        dfnl = np.zeros((self.nat, self.nhxs, 3, nstate), order="F") if want_dfnl else None
        check(lib().mc_rnlsm2_dev(self.h, c0.ptr, nstate, _p(dfnl) if want_dfnl else None))
        return dfnl

    def rnlsm_dev(self, c0: DeviceArray, nstate: int, tfor: bool, want_fnl: bool = False,
                  want_dfnl: bool = False):
        """rnlsm1 (+ rnlsm2 with tfor) of the same coefficients; returns (fnl, dfnl)
        host copies as requested (None otherwise)."""
        fnl = np.zeros((self.nat, self.nhxs, nstate), order="F") if want_fnl else None
        dfnl = np.zeros((self.nat, self.nhxs, 3, nstate), order="F") if (tfor and want_dfnl) else None
        check(lib().mc_rnlsm_dev(self.h, c0.ptr, nstate, 1 if tfor else 0, _p(fnl) if fnl is not None else None,
                                 _p(dfnl) if dfnl is not None else None))
        return fnl, dfnl

    def rnlfor(self, f: np.ndarray, fnl: np.ndarray | None = None, dfnl: np.ndarray | None = None,
               fion: np.ndarray | None = None) -> np.ndarray:
        """fion (3, nat) with the nonlocal forces subtracted (zeros if not given);
        fnl / dfnl None: the device copies of the last rnlsm1 / rnlsm2."""
        f = np.ascontiguousarray(f, dtype=np.float64)
        out = np.zeros((3, self.nat), order="F") if fion is None else np.array(fion, dtype=np.float64, order="F")
        a = np.asfortranarray(fnl, dtype=np.float64) if fnl is not None else None
        b = np.asfortranarray(dfnl, dtype=np.float64) if dfnl is not None else None
        check(lib().mc_rnlfor(self.h, _p(f), len(f), _p(a) if a is not None else None,
                              _p(b) if b is not None else None, _p(out)))
        return out

    def rnlrh(self, f: np.ndarray) -> float:
        """E_NL from the device fnl of the last rnlsm1 / rnlsm for len(f) states."""
        f = np.ascontiguousarray(f, dtype=np.float64)
        e = C.c_double(0.0)
        check(lib().mc_rnlrh(self.h, _p(f), len(f), C.byref(e)))
        return e.value

    # ------------------------------------------------------------ vofrho / ewald
    VOFRHO_KEYS = ("ehep", "ehii", "ehee", "epseu", "vploc", "exc", "vxc")

    def set_density_gset(self, inyh: np.ndarray, hg: np.ndarray, gk: np.ndarray):
        """inyh (nhg, 3) centred 1-based indices, hg (nhg,), gk (nhg, 3) in 2 pi / alat."""
        self._inyh = np.ascontiguousarray(inyh, dtype=np.int32)
        self._hgd = np.ascontiguousarray(hg, dtype=np.float64)
        self._gkd = np.ascontiguousarray(gk, dtype=np.float64)
        check(lib().mc_set_density_gset(self.h, len(self._hgd), self._inyh.ctypes.data_as(C.POINTER(C.c_int)),
                                        _p(self._hgd), _p(self._gkd)))

    def set_local_pp(self, na, zv, raggio, vps: np.ndarray, rhops: np.ndarray):
        """vps / rhops (nsp, nhg)."""
        na = np.ascontiguousarray(na, dtype=np.int32)
        zv = np.ascontiguousarray(zv, dtype=np.float64)
        rg = np.ascontiguousarray(raggio, dtype=np.float64)
        vp = np.asfortranarray(vps, dtype=np.float64)
        rp = np.asfortranarray(rhops, dtype=np.float64)
        self.nat_ions = int(na.sum())
        check(lib().mc_set_local_pp(self.h, len(na), na.ctypes.data_as(C.POINTER(C.c_int)), _p(zv), _p(rg), _p(vp),
                                    _p(rp)))

    def set_ions(self, tau: np.ndarray, ht: np.ndarray):
        """tau (nat, 3) bohr; ht (3, 3) with rows = lattice vectors."""
        t = np.ascontiguousarray(tau, dtype=np.float64)
        h = np.asfortranarray(ht, dtype=np.float64)
        check(lib().mc_set_ions(self.h, _p(t), _p(h)))

    def vofrho(self, rho: np.ndarray, forces: bool = True):
        """(v(r), energies dict, fion (3, nat) from potfor or None)."""
        r = np.ascontiguousarray(rho, dtype=np.float64)
        v = np.zeros_like(r)
        e = np.zeros(7)
        f = np.zeros((3, self.nat_ions), order="F") if forces else None
        check(lib().mc_vofrho(self.h, _p(r), _p(v), _p(e), _p(f) if forces else None))
        return v, dict(zip(self.VOFRHO_KEYS, e)), f

    def vofrho_dev(self, rho: DeviceArray, v: DeviceArray, forces: bool = False):
        e = np.zeros(7)
        f = np.zeros((3, self.nat_ions), order="F") if forces else None
        check(lib().mc_vofrho_dev(self.h, rho.ptr, v.ptr, _p(e), _p(f) if forces else None))
        return dict(zip(self.VOFRHO_KEYS, e)), f

    def phfac(self):
        """Projector structure factors of the ions of set_ions, built on the device."""
        check(lib().mc_phfac(self.h))

    def ewald(self, iesr: int = 2, forces: bool = True):
        """(esr, eself, fion (3, nat) or None)."""
        esr = C.c_double()
        es = C.c_double()
        f = np.zeros((3, self.nat_ions), order="F") if forces else None
        check(lib().mc_ewald(self.h, int(iesr), C.byref(esr), C.byref(es), _p(f) if forces else None))
        return esr.value, es.value, f

    def fnonloc(self, c2: np.ndarray, f: np.ndarray, fnl: np.ndarray | None = None) -> np.ndarray:
        out = np.array(c2, dtype=np.complex128, order="F", copy=True)
        f = np.ascontiguousarray(f, dtype=np.float64)
        a = np.asfortranarray(fnl, dtype=np.float64) if fnl is not None else None
        check(lib().mc_fnonloc(self.h, _p(out), out.shape[0], _p(f), out.shape[1], _p(a) if a is not None else None))
        return out

    def fnonloc_dev(self, c2: DeviceArray, f: np.ndarray, nstate: int):
        f = np.ascontiguousarray(f, dtype=np.float64)
        check(lib().mc_fnonloc_dev(self.h, c2.ptr, _p(f), nstate, None))

    def forces_c2_dev(self, c0: DeviceArray, c2: DeviceArray, f: np.ndarray, vpot: DeviceArray, nstate: int):
        """c2 <- vpsi + fnonloc gradient of c0 (c2 overwritten, not accumulated);
        fnonloc uses the device fnl of the last rnlsm1."""
        f = np.ascontiguousarray(f, dtype=np.float64)
        check(lib().mc_forces_c2_dev(self.h, c0.ptr, c2.ptr, _p(f), vpot.ptr, nstate))

    # ------------------------------------------------------------ timing
    def timer_enable(self, on: bool = True):
        check(lib().mc_timer_enable(self.h, 1 if on else 0))

    def timer_reset(self):
        check(lib().mc_timer_reset(self.h))

    def timer_select(self, names=None):
        """Time only the kernels named (None: all) while the timer is enabled."""
        mask = (1 << 64) - 1 if names is None else sum(1 << _lib.KID[n] for n in names)
        check(lib().mc_timer_select(self.h, mask))

    def timer_get(self, name: str):
        ms, cnt = C.c_double(), C.c_longlong()
        check(lib().mc_timer_get(self.h, _lib.KID[name], C.byref(ms), C.byref(cnt)))
        return ms.value, cnt.value
