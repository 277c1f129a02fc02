"""The charge the FP32 transforms lose, explained: a restatement of the device
engine's arithmetic structure (oracle/fp32_engine.py) with its constants in
FP64 is the exact DFT, and with the constants and twiddles rounded to FP32 it
contracts the norm by a fixed amount per frequency -- 5.9e-8 on average for
the 192-point engine, 1.8e-7 over the three inverse passes of a 192^3 mesh,
which is the csumr - csumg the mixed path shows before its rescale
(tests/test_gpu_mixed.py compares the prediction with the device)."""
import numpy as np
import pytest

from oracle import fp32_engine as E


@pytest.mark.parametrize("n", [16, 24, 54, 128, 192, 224])
def test_engine_restatement_is_the_dft(n):
    j = np.arange(n)
    ref = np.exp(2j * np.pi * np.outer(j, j) / n)         # inverse: e^{+i 2 pi j k / n}
    m = E.engine_matrix(n, 1, np.float64)
    assert np.abs(m - ref).max() < 1e-12
    assert np.abs(E.engine_matrix(n, -1, np.float64) - np.conj(ref)).max() < 1e-12
    assert np.abs(E.column_gains(n, np.float64)).max() < 1e-13


def test_fp32_constants_contract_the_norm():
    d = E.column_gains(192)
    assert d.mean() == pytest.approx(-5.88e-8, rel=0.02)
    assert d.max() <= 1e-12                                # no frequency gains norm
    assert 3 * d.mean() == pytest.approx(-1.76e-7, rel=0.02)
    # the in-register DFTs alone (no table twiddles) carry most of it
    for r, ref in ((16, -2.28e-8), (12, -2.09e-8), (3, -1.20e-8)):
        m = np.array(E.dftn(r, 1, list(np.eye(r, dtype=complex)))).T
        assert (np.abs(m) ** 2).sum() / r ** 2 - 1 == pytest.approx(ref, rel=0.03)


@pytest.mark.parametrize("r", [7, 14, 21, 28, 35, 49])
def test_register_dft_radix7(r):
    """The register DFTs with a factor 7 (bfly7 and the recursion through it)
    in FP64 are the DFT, both directions."""
    j = np.arange(r)
    for d in (1, -1):
        m = np.array(E.dftn(r, d, list(np.eye(r, dtype=complex)), np.float64)).T
        assert np.abs(m - np.exp(d * 2j * np.pi * np.outer(j, j) / r)).max() < 1e-13
