"""ORACLE -- test infrastructure only.

A restatement of the arithmetic structure of the device's compile-time
transform engine (cpmd_amd/csrc/fft_device.h: the in-register DFTs of
length 2..16 recursing into radix 2/3/4/5/7 butterflies with the constants of
twiddle_consts.h, and the four-step N = A B composition with the N-point
twiddle table) with every constant and twiddle rounded to the engine's
arithmetic type, evaluated in FP64.  It explains the charge the FP32
transforms lose: the rounded constants make each transform slightly
contractive, by a fixed amount per output frequency.

column_gains(N) gives d(j) = |F e_j|^2 / N - 1 for the inverse transform of
length N; the real-space norm of a state with coefficients c_G after the
three inverse passes is sum_G |c_G|^2 (1 + d1(m1) + d2(m2) + d3(m3)) to first
order.  tests/test_fp32_charge.py pins the restatement (FP64 constants give
the exact DFT) and compares its prediction with the device's charge.
"""
from __future__ import annotations

import functools

import numpy as np

SHAPES = {16: (4, 4), 18: (6, 3), 24: (6, 4), 54: (9, 6), 128: (16, 8), 192: (16, 12), 224: (16, 14),
          240: (16, 15)}

S3 = 0.86602540378443864676
C51, C52 = 0.30901699437494742410, -0.80901699437494742410
S51, S52 = 0.95105651629515357212, 0.58778525229247312917


def _split(r: int):
    if r % 4 == 0 and r != 4:
        return 4, r // 4
    if r % 3 == 0 and r != 3:
        return 3, r // 3
    if r % 2 == 0 and r != 2:
        return 2, r // 2
    if r % 5 == 0 and r != 5:
        return 5, r // 5
    return 7, r // 7


def _rt(x: float, t) -> float:
    return float(t(x))


def _di(a, d):     # a * d * i (exact)
    return -1j * a if d < 0 else 1j * a


def dftn(r: int, d: int, v: list, t=np.float32) -> list:
    """The register DFT of length r (dftn<R, DIR>) on r arrays of columns."""
    if r == 1:
        return list(v)
    if r == 2:
        return [v[0] + v[1], v[0] - v[1]]
    if r == 4:
        t0, t1, t2, t3 = v[0] + v[2], v[0] - v[2], v[1] + v[3], _di(v[1] - v[3], d)
        return [t0 + t2, t1 + t3, t0 - t2, t1 - t3]
    if r == 3:
        s3 = _rt(S3, t)
        t1, t2 = v[1] + v[2], v[1] - v[2]
        m, q = v[0] - 0.5 * t1, _di(t2 * s3, d)
        return [v[0] + t1, m + q, m - q]
    if r == 5:
        c1, c2, s1, s2 = (_rt(x, t) for x in (C51, C52, S51, S52))
        t1, t2, t3, t4 = v[1] + v[4], v[2] + v[3], v[1] - v[4], v[2] - v[3]
        a1, a2 = v[0] + c1 * t1 + c2 * t2, v[0] + c2 * t1 + c1 * t2
        b1, b2 = _di(s1 * t3 + s2 * t4, d), _di(s2 * t3 - s1 * t4, d)
        return [v[0] + t1 + t2, a1 + b1, a2 + b2, a2 - b2, a1 - b1]
    if r == 7:      # bfly7: cos / sin of 2 pi j / 7 rounded to t
        cj = [_rt(np.cos(2 * np.pi * j / 7), t) for j in range(7)]
        sj = [_rt(np.sin(2 * np.pi * j / 7), t) for j in range(7)]
        tt = [v[k] + v[7 - k] for k in (1, 2, 3)]
        uu = [v[k] - v[7 - k] for k in (1, 2, 3)]
        out = [v[0] + tt[0] + tt[1] + tt[2]] + [None] * 6
        for m in (1, 2, 3):
            a = v[0] + sum(cj[(k * m) % 7] * tt[k - 1] for k in (1, 2, 3))
            b = _di(sum(sj[(k * m) % 7] * uu[k - 1] for k in (1, 2, 3)), d)
            out[m], out[7 - m] = a + b, a - b
        return out
    p, q = _split(r)
    e = np.arange(r)
    cc, ss = np.cos(2 * np.pi * e / r), np.sin(2 * np.pi * e / r)   # W_r^e = c - i s (twiddle_consts.h)
    y = [None] * r
    for n2 in range(q):
        u = dftn(p, d, [v[q * n1 + n2] for n1 in range(p)], t)
        for k1 in range(p):
            k = (n2 * k1) % r
            w = complex(_rt(cc[k], t), (-1 if d < 0 else 1) * _rt(ss[k], t))
            y[n2 * p + k1] = u[k1] if k == 0 else u[k1] * w
    out = [None] * r
    for k1 in range(p):
        u = dftn(q, d, [y[n2 * p + k1] for n2 in range(q)], t)
        for k2 in range(q):
            out[k1 + p * k2] = u[k2]
    return out


def engine_matrix(n: int, d: int = 1, t=np.float32) -> np.ndarray:
    """The n x n matrix of the four-step engine of length n = A B (fft_fast /
    F4): an A-point DFT over x[B n1 + n2], the table twiddle w_n^{n2 k1}
    (rounded to t), a B-point DFT; output X[k1 + A k2]."""
    a, b = SHAPES[n]
    j = np.arange(n)
    tab = np.exp(-2j * np.pi * j / n)
    tab = np.array([complex(_rt(z.real, t), _rt(z.imag, t)) for z in tab])
    x = np.eye(n, dtype=complex)
    ys = []
    for n2 in range(b):
        v = dftn(a, d, [x[b * n1 + n2] for n1 in range(a)], t)
        for k1 in range(1, a):
            w = tab[(n2 * k1) % n]
            v[k1] = v[k1] * (w if d < 0 else np.conj(w))
        ys.append(v)
    out = np.zeros((n, n), complex)
    for k1 in range(a):
        u = dftn(b, d, [ys[n2][k1] for n2 in range(b)], t)
        for k2 in range(b):
            out[k1 + a * k2] = u[k2]
    return out


@functools.lru_cache(maxsize=None)
def column_gains(n: int, t=np.float32) -> np.ndarray:
    """d(j) = |F e_j|^2 / n - 1 of the inverse engine of length n."""
    m = engine_matrix(n, 1, t)
    return (np.abs(m) ** 2).sum(axis=0) / n - 1.0


def predicted_charge_loss(cell, miller: np.ndarray, c: np.ndarray, f: np.ndarray) -> float:
    """(csumr - csumg) / csumg of the FP32 rhoofr to first order: the
    occupation-weighted mean over the G vectors (and their mirrors) of
    d1(m1) + d2(m2) + d3(m3)."""
    n1, n2, n3 = cell.nr
    d1, d2, d3 = column_gains(n1), column_gains(n2), column_gains(n3)
    m = np.asarray(miller[:c.shape[0]], dtype=np.int64)
    dp = d1[m[:, 0] % n1] + d2[m[:, 1] % n2] + d3[m[:, 2] % n3]
    dm = d1[(-m[:, 0]) % n1] + d2[(-m[:, 1]) % n2] + d3[(-m[:, 2]) % n3]
    w = (np.abs(c) ** 2) * f[None, :]
    g0 = w[0].sum() * dp[0]                                  # G = 0 once
    num = g0 + ((w[1:] * (dp[1:] + dm[1:])[:, None]).sum())
    den = w[0].sum() + 2.0 * w[1:].sum()
    return float(num / den)
