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
    subprocess.run(["make", "-s", "-C", str(HERE), "-B", "liboracle.so"], check=True)
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
