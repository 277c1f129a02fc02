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

    def synchronize(self):
        check(lib().mc_ctx_synchronize(self.h))

    # ------------------------------------------------------------ transfers
    def alloc_wfn(self, nstate: int) -> DeviceArray:
        return DeviceArray(16 * self.ld * max(nstate, 1)).zero()

    def zero_wfn(self, dev: DeviceArray, nstate: int):
        check(lib().mc_wfn_zero(self.h, dev.ptr, nstate))

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

    # ------------------------------------------------------------ timing
    def timer_enable(self, on: bool = True):
        check(lib().mc_timer_enable(self.h, 1 if on else 0))

    def timer_reset(self):
        check(lib().mc_timer_reset(self.h))

    def timer_get(self, name: str):
        ms, cnt = C.c_double(), C.c_longlong()
        check(lib().mc_timer_get(self.h, _lib.KID[name], C.byref(ms), C.byref(cnt)))
        return ms.value, cnt.value
