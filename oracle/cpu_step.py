"""ORACLE -- test infrastructure and the CPU baseline of bench.py only.

CPMD's CPU algorithm for the SCF-step content that the GPU path runs, on the
host's cores, used to time the CPU path (the product library never calls it):

* rhoofr and vpsi over Gamma state pairs with CPMD's stick-sparse 3D
  transforms, one state pair per OpenMP thread, the 1D transforms through the
  FFTW3 interface (libcpustep.so, oracle/cpu_step.c);
* the projector contractions rnlsm1 (src/rnlsm1_utils.mod.F90:84-163) and
  fnonloc (src/fnonloc_utils.mod.F90:209-300) and the Loewdin
  orthonormalisation (src/lowdin_utils.mod.F90:27-192: S = ovlap(C, C),
  S = Z W Z^T, C <- C Z W^-1/2 Z^T) as threaded BLAS / LAPACK calls through
  numpy / scipy.

Its results are checked against the plain-C oracle in tests/test_oracle.py.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
from pathlib import Path

import numpy as np
import scipy.linalg as sla

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "libcpustep.so"
_lib = None
_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)


def lib():
    global _lib
    if _lib is None:
        # each OpenMP thread runs its own transforms: the FFTW3 interface single-threaded
        os.environ.setdefault("MKL_THREADING_LAYER", "SEQUENTIAL")
        if not LIB_PATH.exists():
            subprocess.run(["make", "-s", "-C", str(HERE), "libcpustep.so"], check=True)
        L = C.CDLL(str(LIB_PATH))
        L.cs_rhoofr.argtypes = [_ip, _ip, _dp, C.c_double, C.c_double, _dp, C.c_int, _dp, C.c_int, C.c_int, _dp]
        L.cs_vpsi.argtypes = [_ip, _ip, _dp, C.c_double, C.c_double, _dp, C.c_int, _dp, _dp, _dp, C.c_int, C.c_int]
        L.cs_threads.restype = C.c_int
        _lib = L
    return _lib


def _p(a, t=_dp):
    return a.ctypes.data_as(t)


class CpuStep:
    def __init__(self, cell, gset, ngw: int | None = None):
        self.cell = cell
        self.ngw = ngw = gset.ngw if ngw is None else ngw
        nr, kr = cell.nr, cell.kr
        self.dims = np.array([nr[0], nr[1], nr[2], kr[0], kr[1], kr[2], ngw, 1], dtype=np.int32)
        nh = np.array([n // 2 + 1 for n in nr], dtype=np.int32)
        self.inyh = np.ascontiguousarray((gset.miller[:ngw] + nh).astype(np.int32))
        self.hg = np.ascontiguousarray(gset.hg[:ngw], dtype=np.float64)
        self.tpiba2, self.omega = cell.tpiba2, cell.omega
        self.nnr1 = kr[0] * kr[1] * kr[2]

    @staticmethod
    def threads() -> int:
        return int(lib().cs_threads())

    def rhoofr(self, c0: np.ndarray, f: np.ndarray, pairs: int | None = None) -> np.ndarray:
        """rho (CPMD layout) of the first `pairs` state pairs (all by default)."""
        c0 = np.asfortranarray(c0, dtype=np.complex128)
        f = np.ascontiguousarray(f, dtype=np.float64)
        n = c0.shape[1]
        npairs = (n + 1) // 2 if pairs is None else pairs
        rho = np.zeros(self.nnr1)
        lib().cs_rhoofr(_p(self.dims, _ip), _p(self.inyh, _ip), _p(self.hg), self.tpiba2, self.omega,
                        _p(c0 if c0.size else np.zeros(2)), c0.shape[0], _p(f), n, npairs, _p(rho))
        return rho

    def vpsi(self, c0: np.ndarray, c2: np.ndarray, f: np.ndarray, vpot: np.ndarray,
             pairs: int | None = None) -> np.ndarray:
        """c2 += -(f_i/2)(tpiba2 hg c0_i + 2 P_G[V psi_i]) (f_i/2 = 0 -> 1) over the first `pairs` pairs."""
        c0 = np.asfortranarray(c0, dtype=np.complex128)
        out = np.array(c2, dtype=np.complex128, order="F", copy=True)
        f = np.ascontiguousarray(f, dtype=np.float64)
        v = np.ascontiguousarray(vpot, dtype=np.float64)
        n = c0.shape[1]
        npairs = (n + 1) // 2 if pairs is None else pairs
        lib().cs_vpsi(_p(self.dims, _ip), _p(self.inyh, _ip), _p(self.hg), self.tpiba2, self.omega,
                      _p(c0), c0.shape[0], _p(out), _p(f), _p(v), n, npairs)
        return out

    # ---------------------------------------------------------------- contractions
    # Wavefunctions (ngw, n) complex column-major are handled through their
    # transposed real view: row i of (n, 2 ngw) holds state i as (re, im) pairs.
    @staticmethod
    def _rows(c: np.ndarray) -> np.ndarray:
        return np.ascontiguousarray(np.asarray(c, dtype=np.complex128).T).view(np.float64)

    @staticmethod
    def _from_rows(r: np.ndarray) -> np.ndarray:
        return np.ascontiguousarray(r).view(np.complex128).T

    def ovlap(self, c1: np.ndarray, c2: np.ndarray) -> np.ndarray:
        x1, x2 = self._rows(c1), self._rows(c2)
        return 2.0 * (x1 @ x2.T) - np.outer(x1[:, 0], x2[:, 0])

    def lowdin(self, c: np.ndarray) -> np.ndarray:
        """C S^-1/2 through the eigendecomposition of S (CPMD's lowdin), Im c(G=0) = 0."""
        x = self._rows(c)
        s = 2.0 * (x @ x.T) - np.outer(x[:, 0], x[:, 0])
        w, z = np.linalg.eigh(s)
        a = (z * (1.0 / np.sqrt(w))) @ z.T
        out = self._from_rows(a.T @ x)
        out[0] = out[0].real
        return out

    def rnlsm1(self, c0: np.ndarray, beta: np.ndarray) -> np.ndarray:
        """fnl(ia, i) = 2 Re sum_G conj(beta(G, ia)) c(G, i), G = 0 halved (one projector set beta)."""
        bb = np.array(beta, dtype=np.complex128, copy=True)
        bb[0] *= 0.5
        return 2.0 * (self._rows(bb) @ self._rows(c0).T)

    def fnonloc(self, c2: np.ndarray, beta: np.ndarray, dd: np.ndarray) -> np.ndarray:
        """c2 += beta dd (the scaled projector sum of fnonloc as one real GEMM)."""
        return self._from_rows(self._rows(c2) + np.asarray(dd, dtype=np.float64).T @ self._rows(beta))
