"""RESTART files (CPMD's wv30 / rv30 sections 1-10 and 23) through the
library's host entry points, checked three ways:

* byte-exact against the oracle's plain-Python restatement of the layout
  wv30 writes (record framing, section order, markers, w_wfnio / wr30wfn
  headers, the izamax / idamax scales);
* read back by the library from the oracle's bytes (every value bit-exact),
  and the library's write -> read round trip;
* read by the Fortran runtime (amdflang) with the READ statements of rv30,
  which proves the record framing CPMD's own reader sees.
The reader's refusals (wrong header, compressed records, k points) are
checked on edited files."""
import shutil
import struct
import subprocess
from pathlib import Path

import numpy as np
import pytest

from cpmd_amd import restart as R
from cpmd_amd import system as S
from cpmd_amd._lib import McError
from oracle import restart as OR

ROOT = Path(__file__).resolve().parent.parent


def _case(nstate=6, seed=3):
    sysm = S.make_system(12.0, 8, 30.0, nstate=nstate, with_density=True)
    rng = np.random.default_rng(seed)
    c0 = S.initial_coefficients(sysm, seed=seed)
    cm = np.asfortranarray(1e-3 * (rng.standard_normal(c0.shape) + 1j * rng.standard_normal(c0.shape)))
    cm[0] = cm[0].real
    velp = 1e-4 * rng.standard_normal(sysm.tau.shape)
    taui = sysm.tau + 0.01
    return sysm, R.header_for(sysm), c0, cm, velp, taui


@pytest.mark.parametrize("which", ["all", "c0_only", "no_wfn"])
def test_bytes_equal_oracle_layout(tmp_path, which):
    sysm, h, c0, cm, velp, taui = _case()
    kw = dict(c0=c0, cm=cm, velp=velp, taui=taui, f=sysm.f)
    if which == "c0_only":
        kw = dict(c0=c0)
    elif which == "no_wfn":
        kw = dict(velp=velp)
    p = tmp_path / "RESTART.1"
    R.write_restart(p, h, sysm.tau, **kw)
    ref = OR.wv30_bytes(h, sysm.tau, **kw)
    data = p.read_bytes()
    assert data == ref
    recs = OR.records(data)
    assert recs[0].startswith(b" STUTTGART VERSION 3.0") and len(recs[0]) == 80
    # every section 2..99 has its marker, in order: 98 markers with a 12-byte payload
    n, nat = h["nstate"], sysm.nat
    nwfn = (3 + 2 * n if "c0" in kw else 0) + (1 + 2 * n if "cm" in kw else 0)
    ndata = 2 + 2 + nat * (1 + ("velp" in kw) + ("taui" in kw)) + 1 + 1 + nwfn + ("f" in kw)
    assert len(recs) == 1 + 98 + ndata


def test_read_oracle_bytes(tmp_path):
    sysm, h, c0, cm, velp, taui = _case(nstate=5)
    p = tmp_path / "RESTART"
    p.write_bytes(OR.wv30_bytes(h, sysm.tau, c0=c0, cm=cm, velp=velp, taui=taui, f=sysm.f))
    r = R.read_restart(p)
    assert r["sections"] == R.VELP | R.TAUI | R.C0 | R.CM | R.OCC
    for k in ("ibrav", "indpg", "ecut", "cdual", "dual", "nel", "nr", "nstate", "ngwks", "ngwls", "nhgs",
              "nhgls"):
        assert r[k] == (tuple(h[k]) if k == "nr" else h[k]), k
    assert list(r["celldm"]) == list(h["celldm"]) and list(r["na"]) == list(h["na"])
    assert np.array_equal(r["tau0"], sysm.tau) and np.array_equal(r["velp"], velp)
    assert np.array_equal(r["taui"], taui) and np.array_equal(r["f"], sysm.f[:5])
    assert np.array_equal(r["c0"], c0) and np.array_equal(r["cm"], cm)


def test_round_trip_and_header_only(tmp_path):
    sysm, h, c0, cm, velp, taui = _case(nstate=7)
    p = tmp_path / "RESTART.2"
    R.write_restart(p, h, sysm.tau, c0=c0, cm=cm)
    hh = R.read_header(p)
    assert hh["sections"] == R.C0 | R.CM and hh["nat"] == sysm.nat and hh["nstate"] == 7
    r = R.read_restart(p)
    assert np.array_equal(r["c0"], c0) and np.array_equal(r["cm"], cm)
    assert "velp" not in r and "f" not in r


def test_fortran_runtime_reads_it(tmp_path):
    fc = shutil.which("amdflang") or "/opt/rocm/bin/amdflang"
    if not Path(fc).exists():
        pytest.skip("amdflang not available")
    exe = tmp_path / "restart_read.x"
    subprocess.run([fc, "-O0", str(ROOT / "tests" / "fortran" / "restart_read.F90"), "-o", str(exe)], check=True,
                   cwd=tmp_path, capture_output=True)
    sysm, h, c0, cm, velp, taui = _case(nstate=4)
    p = tmp_path / "RESTART.1"
    R.write_restart(p, h, sysm.tau, c0=c0, cm=cm, velp=velp, f=sysm.f)
    out = tmp_path / "read.bin"
    subprocess.run([str(exe), str(p), str(out)], check=True, capture_output=True)
    n, ngw, nat = 4, h["ngwks"], sysm.nat
    scale9 = abs(c0.T.ravel()[np.argmax(np.abs(c0.T.ravel().real) + np.abs(c0.T.ravel().imag))])
    exp = struct.pack("<ii6d", h["ibrav"], h["indpg"], *h["celldm"]) + struct.pack("<ii", 1, nat)
    exp += struct.pack("<i", nat) + sysm.tau.astype("<f8").tobytes()
    exp += struct.pack("<i", nat) + velp.astype("<f8").tobytes() + struct.pack("<i", 0)
    exp += struct.pack("<ddiiiii", h["ecut"], h["cdual"], 0, h["nel"], *h["nr"])
    exp += struct.pack("<6i", n, 0, ngw, h["ngwls"], h["nhgs"], h["nhgls"])
    exp += struct.pack("<iiiid", -(2 * n + 3), n, ngw, 1, scale9) + struct.pack("<iii", 0, 0, 0)
    exp += np.asfortranarray(c0).T.tobytes()
    exp += struct.pack("<iiiid", 2 * n + 1, n, ngw, 1, float(max(np.abs(cm.real).max(), np.abs(cm.imag).max())))
    exp += np.asfortranarray(cm).T.tobytes()
    exp += struct.pack("<i", 1) + sysm.f[:n].astype("<f8").tobytes() + struct.pack("<i", -12345)
    assert out.read_bytes() == exp


def test_reader_refusals(tmp_path):
    sysm, h, c0, cm, velp, taui = _case(nstate=2)
    good = OR.wv30_bytes(h, sysm.tau, c0=c0)
    p = tmp_path / "bad"
    p.write_bytes(good.replace(b"STUTTGART", b"STUTTGARX", 1))
    with pytest.raises(McError, match="not a CPMD RESTART"):
        R.read_header(p)
    # a compressed C0 (WCOMPB 2): the icompr field of the w_wfnio header
    n, ngw = 2, h["ngwks"]
    hdr = OR.record(struct.pack("<iiid", n, ngw, 1, 0.0))
    i = good.index(hdr[:16])            # length word, n, ngwks, icompr
    bad = bytearray(good)
    bad[i + 4 + 8:i + 4 + 12] = struct.pack("<i", 2)
    p.write_bytes(bytes(bad))
    with pytest.raises(McError, match="compressed"):
        R.read_restart(p)
    # a k-point file: nkpts0 != 0 in section 8
    s8 = OR.record(struct.pack("<6i", n, 0, ngw, h["ngwls"], h["nhgs"], h["nhgls"]))
    p.write_bytes(good.replace(s8, OR.record(struct.pack("<6i", n, 1, ngw, h["ngwls"], h["nhgs"], h["nhgls"]))))
    with pytest.raises(McError, match="Gamma"):
        R.read_header(p)
    # a truncated file
    p.write_bytes(good[:len(good) // 2])
    with pytest.raises(McError):
        R.read_restart(p)
