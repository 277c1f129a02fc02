"""The C-ABI library: it loads, exports every entry point include/micpmd.h
declares, and its host-only entry points behave (no GPU needed)."""
import ctypes as C
import re
from pathlib import Path

import numpy as np
import pytest

from cpmd_amd import _lib, ops
from cpmd_amd import system as S

ROOT = Path(__file__).resolve().parent.parent


def declared_functions():
    names = set()
    for h in (ROOT / "include").glob("*.h"):
        txt = h.read_text()
        txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
        for m in re.finditer(r"\b(mc_[a-z0-9_]+)\s*\(", txt):
            names.add(m.group(1))
    return sorted(names)


def test_library_exports_header():
    L = _lib.lib()
    names = declared_functions()
    assert "mc_vpsi" in names and "mc_rhoofr" in names
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing
    # every declared function is bound with a signature on the Python side
    assert set(names) <= set(_lib.SIGNATURES), set(names) - set(_lib.SIGNATURES)


def test_layout_stats_host_only():
    cell = S.cubic_cell(12.0, 40.0)
    gs = S.gvectors(cell, with_density=False)
    assert ops.layout_stats(cell, gs) == (457, 25, 229)


def test_layout_rejects_bad_input():
    cell = S.cubic_cell(12.0, 40.0)
    gs = S.gvectors(cell, with_density=False)
    # a G set with both +G and -G is not a CPMD half sphere
    bad = S.GSet(miller=np.vstack([gs.miller[:10], -gs.miller[1:3]]).astype(np.int32),
                 hg=np.concatenate([gs.hg[:10], gs.hg[1:3]]), ngw=12, nhg=12)
    with pytest.raises(_lib.McError) as e:
        ops.layout_stats(cell, bad)
    assert e.value.code == 5 and "half sphere" in str(e.value)
    # G vector outside the mesh
    big = S.GSet(miller=np.array([[0, 0, 0], [40, 0, 0]], np.int32), hg=np.array([0.0, 1600.0]), ngw=2, nhg=2)
    with pytest.raises(_lib.McError) as e:
        ops.layout_stats(cell, big)
    assert e.value.code == 1
    # transform lengths with a factor 7 are refused at context creation only;
    # the layout itself is fine
    assert _lib.lib().mc_last_error() is not None


def test_null_arguments():
    L = _lib.lib()
    assert L.mc_layout_stats(None, None, None, None) == 1
    assert b"null" in L.mc_last_error()
    assert L.mc_ctx_destroy(None) == 0
