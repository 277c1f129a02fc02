"""ORACLE -- test infrastructure and the CPU baseline of bench.py only.

A multi-threaded CPU restatement of the SCF step content that the GPU path
runs, used to time CPMD's CPU algorithm on the host's cores (the product
library never calls it):

* rhoofr (src/rhoofr_utils.mod.F90:305-409) and vpsi
  (src/vpsi_utils.mod.F90:376-675) over Gamma state pairs with CPMD's
  stick-sparse 3D transforms (src/fftmain_utils.mod.F90:71-157, sparse=.TRUE.):
  pack two real states into one complex stick array at +G / -G
  (src/state_utils.mod.F90:171-189), z pass over the sticks, y pass over the
  x-frequency columns, x pass over the full mesh (inverse: e^{+iGr},
  unscaled; forward: e^{-iGr} scaled by 1/(n1 n2 n3)); the decode
  fp = psi(+G), fm = psi(-G) of vpsi:661-671 with the f/2 = 0 -> 1 rule of
  vpsi:627-633.  The 1D
  transforms are scipy.fft (pocketfft) over all requested threads, batched
  over several pairs.
* the projector contractions rnlsm1 (src/rnlsm1_utils.mod.F90:84-163) and
  fnonloc (src/fnonloc_utils.mod.F90:209-300) and the Loewdin
  orthonormalisation (src/lowdin_utils.mod.F90:27-192: S = ovlap(C, C),
  S = Z W Z^T, C <- C Z W^-1/2 Z^T) as threaded BLAS / LAPACK calls through
  numpy / scipy.

Its results are checked against the plain-C oracle in tests/test_oracle.py.
"""
from __future__ import annotations

import numpy as np
import scipy.fft as sfft
import scipy.linalg as sla


class CpuStep:
    def __init__(self, cell, gset, ngw: int | None = None, workers: int = 1):
        self.cell = cell
        self.ngw = ngw = gset.ngw if ngw is None else ngw
        self.workers = workers
        n1, n2, n3 = self.nr = tuple(int(x) for x in cell.nr)
        self.kr = tuple(int(x) for x in cell.kr)
        m = np.asarray(gset.miller[:ngw], dtype=np.int64)
        self.hg = np.asarray(gset.hg[:ngw], dtype=np.float64)
        self.tpiba2, self.omega = cell.tpiba2, cell.omega
        # sticks: the distinct (m1, m2) of the sphere (+G and -G)
        xp, yp, zp = m[:, 0] % n1, m[:, 1] % n2, m[:, 2] % n3
        xm, ym, zm = (-m[:, 0]) % n1, (-m[:, 1]) % n2, (-m[:, 2]) % n3
        keys = np.concatenate([xp * n2 + yp, xm * n2 + ym])
        uniq, inv = np.unique(keys, return_inverse=True)
        self.nstick = len(uniq)
        self.sp, self.sm = inv[:ngw], inv[ngw:]
        self.zp, self.zm = zp, zm
        sx, sy = uniq // n2, uniq % n2
        cols, self.scol = np.unique(sx, return_inverse=True)   # x-frequency columns
        self.xcol = cols
        self.sy = sy
        self.ncol = len(cols)

    # ---------------------------------------------------------------- transforms
    def _pack(self, a: np.ndarray, b: np.ndarray) -> np.ndarray:
        """Sticks A[p, s, z] of the pairs (a[:, p], b[:, p]) after the z pass."""
        P = a.shape[1]
        A = np.zeros((P, self.nstick, self.nr[2]), dtype=np.complex128)
        A[:, self.sp, self.zp] = (a + 1j * b).T
        A[:, self.sm[1:], self.zm[1:]] = (np.conj(a[1:]) + 1j * np.conj(b[1:])).T
        return sfft.ifft(A, axis=2, norm="forward", workers=self.workers, overwrite_x=True)

    def _to_real(self, A: np.ndarray) -> np.ndarray:
        """Sticks -> psi(p, x, y, z) through the y and x passes."""
        P = A.shape[0]
        n1, n2, n3 = self.nr
        B = np.zeros((P, self.ncol, n2, n3), dtype=np.complex128)
        B[:, self.scol, self.sy, :] = A
        B = sfft.ifft(B, axis=2, norm="forward", workers=self.workers, overwrite_x=True)
        R = np.zeros((P, n1, n2, n3), dtype=np.complex128)
        R[:, self.xcol] = B
        return sfft.ifft(R, axis=1, norm="forward", workers=self.workers, overwrite_x=True)

    def _to_sticks(self, R: np.ndarray) -> np.ndarray:
        """psi(p, x, y, z) -> forward transform values at +G and -G (x, y, z passes)."""
        R = sfft.fft(R, axis=1, norm="forward", workers=self.workers, overwrite_x=True)
        B = sfft.fft(R[:, self.xcol], axis=2, norm="forward", workers=self.workers, overwrite_x=True)
        A = sfft.fft(B[:, self.scol, self.sy, :], axis=2, norm="forward", workers=self.workers, overwrite_x=True)
        return A[:, self.sp, self.zp].T, A[:, self.sm, self.zm].T

    def _real_layout(self, g: np.ndarray) -> np.ndarray:
        """(n1, n2, n3) -> CPMD's v(ix + kr1 (iy + kr2 iz)) with zero padding."""
        n1, n2, n3 = self.nr
        k1, k2, k3 = self.kr
        out = np.zeros((k3, k2, k1))
        out[:n3, :n2, :n1] = g.transpose(2, 1, 0)
        return out.ravel()

    def _pairs(self, c: np.ndarray, p0: int, p1: int):
        n = c.shape[1]
        ia = np.arange(2 * p0, 2 * p1, 2)
        ib = ia + 1
        a = c[:, ia]
        b = np.where(ib < n, 1, 0)[None, :] * c[:, np.minimum(ib, n - 1)]
        return ia, ib, a, b

    def rhoofr(self, c0: np.ndarray, f: np.ndarray, batch: int = 8, pairs: int | None = None):
        """rho (CPMD layout), ekin, csumg, csumr over the first `pairs` pairs (all by default)."""
        c0 = np.asarray(c0, dtype=np.complex128)
        f = np.asarray(f, dtype=np.float64)
        n = c0.shape[1]
        npairs = (n + 1) // 2 if pairs is None else pairs
        n1, n2, n3 = self.nr
        rho = np.zeros((n1, n2, n3))
        for p0 in range(0, npairs, batch):
            p1 = min(npairs, p0 + batch)
            ia, ib, a, b = self._pairs(c0, p0, p1)
            psi = self._to_real(self._pack(a, b))
            fa = f[ia] / self.omega
            fb = np.where(ib < n, f[np.minimum(ib, n - 1)], 0.0) / self.omega
            rho += np.einsum("p,pxyz->xyz", fa, psi.real ** 2) + np.einsum("p,pxyz->xyz", fb, psi.imag ** 2)
        ns = min(n, 2 * npairs)
        cc = c0[:, :ns]
        nrm = 2.0 * (cc.real ** 2 + cc.imag ** 2).sum(axis=0) - cc[0].real ** 2
        kin = (self.hg[:, None] * (cc.real ** 2 + cc.imag ** 2)).sum(axis=0)
        csumg = float(f[:ns] @ nrm)
        ekin = float(f[:ns] @ kin) * self.tpiba2
        csumr = float(rho.sum()) * self.omega / (n1 * n2 * n3)
        return self._real_layout(rho), ekin, csumg, csumr

    def vpsi(self, c0: np.ndarray, c2: np.ndarray, f: np.ndarray, vpot: np.ndarray, batch: int = 8,
             pairs: int | None = None) -> np.ndarray:
        """c2 += -(f_i/2)(tpiba2 hg c0_i + 2 P_G[V psi_i]) (f_i/2 = 0 -> 1) over the first `pairs` pairs."""
        c0 = np.asarray(c0, dtype=np.complex128)
        out = np.array(c2, dtype=np.complex128, order="F", copy=True)
        f = np.asarray(f, dtype=np.float64)
        n = c0.shape[1]
        npairs = (n + 1) // 2 if pairs is None else pairs
        n1, n2, n3 = self.nr
        k1, k2, k3 = self.kr
        v = np.asarray(vpot, dtype=np.float64).reshape(k3, k2, k1)[:n3, :n2, :n1].transpose(2, 1, 0)
        for p0 in range(0, npairs, batch):
            p1 = min(npairs, p0 + batch)
            ia, ib, a, b = self._pairs(c0, p0, p1)
            psi = self._to_real(self._pack(a, b))
            psi *= v[None]
            fp, fm = self._to_sticks(psi)
            va = 0.5 * (fp + np.conj(fm))
            vb = -0.5j * (fp - np.conj(fm))
            for k, i in enumerate(ia):
                fi = f[i] if f[i] != 0.0 else 2.0
                out[:, i] += -0.5 * fi * self.tpiba2 * self.hg * c0[:, i] - fi * va[:, k]
                j = ib[k]
                if j < n:
                    fj = f[j] if f[j] != 0.0 else 2.0
                    out[:, j] += -0.5 * fj * self.tpiba2 * self.hg * c0[:, j] - fj * vb[:, k]
        return out

    # ---------------------------------------------------------------- contractions
    @staticmethod
    def _real(c: np.ndarray) -> np.ndarray:
        """(ngw, n) complex -> its (2 ngw, n) real view (re, im interleaved down each column)."""
        return np.ascontiguousarray(np.asarray(c, dtype=np.complex128).T).view(np.float64).T

    @staticmethod
    def _cplx(x: np.ndarray) -> np.ndarray:
        return np.ascontiguousarray(x.T).view(np.complex128).T

    def ovlap(self, c1: np.ndarray, c2: np.ndarray) -> np.ndarray:
        x1, x2 = self._real(c1), self._real(c2)
        return 2.0 * (x1.T @ x2) - np.outer(x1[0], x2[0])

    def lowdin(self, c: np.ndarray) -> np.ndarray:
        """C S^-1/2 through the eigendecomposition of S (CPMD's lowdin), Im c(G=0) = 0."""
        s = self.ovlap(c, c)
        w, z = sla.eigh(s)
        a = (z * (1.0 / np.sqrt(w))) @ z.T
        out = self._cplx(self._real(c) @ a)
        out[0] = out[0].real
        return out

    def rnlsm1(self, c0: np.ndarray, beta: np.ndarray) -> np.ndarray:
        """fnl(ia, i) = 2 Re sum_G conj(beta(G, ia)) c(G, i), G = 0 halved (one projector set beta)."""
        bb = np.array(beta, dtype=np.complex128, copy=True)
        bb[0] *= 0.5
        return 2.0 * (self._real(bb).T @ self._real(c0))

    def fnonloc(self, c2: np.ndarray, beta: np.ndarray, dd: np.ndarray) -> np.ndarray:
        """c2 += beta dd (the scaled projector sum of fnonloc as one GEMM)."""
        out = np.array(c2, dtype=np.complex128, order="F", copy=True)
        out += np.asfortranarray(beta, dtype=np.complex128) @ dd
        return out
