"""The MFMA issue-rate probe (mc_probe_mfma) runs for every instruction it
knows and returns rates between a tenth of and a little above the published
dense peaks (bf16 2500, f64 78.6, f32 157.3 TFLOP/s), and refuses an unknown
kind."""
import ctypes as C

import pytest

from cpmd_amd import ops
from cpmd_amd._lib import McError, check, lib

pytestmark = pytest.mark.gpu

PEAKS = {"bf16_16x16x32": 2500.0, "bf16_32x32x16": 2500.0, "f64_16x16x4": 78.6, "f32_16x16x4": 157.3}


@pytest.mark.parametrize("kind", sorted(PEAKS))
def test_probe_mfma(gpu, kind):
    r = ops.probe_mfma(kind)
    print(f"{kind}: {r:.1f} TFLOP/s", flush=True)
    assert 0.1 * PEAKS[kind] < r < 1.1 * PEAKS[kind]


def test_probe_mfma_refuses_unknown_kind(gpu):
    r = C.c_double()
    with pytest.raises(McError):
        check(lib().mc_probe_mfma(7, C.byref(r)))
