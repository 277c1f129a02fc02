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


# Entry points that never write device memory other than rho, V, c2 and the
# projections: they leave the wavefunction epoch unchanged (MC_API_BEGIN_RO),
# so the intermediates kept under column reuse survive them.  A writing entry
# point added to this list would let a kept intermediate go stale.
READ_ONLY = {
    "mc_rhoofr_dev", "mc_vofrho_dev", "mc_ewald", "mc_set_ions", "mc_phfac",
    "mc_rnlsm_dev", "mc_rnlsm1_dev", "mc_rnlsm2_dev", "mc_rnlfor", "mc_rnlrh",
    "mc_forces_c2_dev", "mc_vpsi_dev",
    "mc_ctx_synchronize", "mc_mem_stats", "mc_memcpy_d2h", "mc_wfn_download", "mc_rhoofr_raw_charge",
    "mc_timer_enable", "mc_timer_reset", "mc_timer_get", "mc_timer_select", "mc_probe_mfma", "mc_dev_alloc",
}


def entry_kinds():
    """{entry point: 'ro' | 'rw'} from the library sources."""
    kinds = {}
    for src in (ROOT / "cpmd_amd" / "csrc").glob("*.hip"):
        txt = src.read_text()
        for m in re.finditer(r"\bint\s+(mc_[a-z0-9_]+)\s*\([^)]*\)\s*\{\s*(MC_API_BEGIN(?:_RO)?)\b", txt, flags=re.S):
            kinds[m.group(1)] = "ro" if m.group(2).endswith("_RO") else "rw"
    return kinds


def test_read_only_entry_points():
    kinds = entry_kinds()
    assert len(kinds) > 80
    ro = {k for k, v in kinds.items() if v == "ro"}
    assert ro == READ_ONLY, (ro - READ_ONLY, READ_ONLY - ro)
    # every entry point that writes wavefunctions counts an epoch
    for name in ("mc_posupa_dev", "mc_cp_velupa_rortv_dev", "mc_rortv_dev", "mc_wfn_axpy", "mc_wfn_axpby",
                 "mc_wfn_upload", "mc_memcpy_h2d", "mc_memset_d", "mc_dev_free", "mc_rgs_dev", "mc_lowdin_dev",
                 "mc_lagrange_dev", "mc_sd_update_dev", "mc_odiis_dev", "mc_set_precision",
                 "mc_ctx_set_column_reuse", "mc_vpsi_dev_noalloc", "mc_restart_read_dev"):
        assert kinds.get(name) == "rw", name
