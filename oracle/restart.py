"""ORACLE -- test infrastructure only.

Plain-Python restatement of the byte layout of CPMD's RESTART file as
zhwwf / wv30 write it for a serial Gamma-point run, used by the tests as the
checker of the library's writer and reader:

* every Fortran WRITE of a sequential unformatted file is one record framed
  by its 4-byte little-endian length before and after the payload;
* section 1: CHARACTER(len=80) ' STUTTGART VERSION 3.0'
  (src/wv30_utils.mod.F90:304-311);
* sections 2..99 start with the marker record (irecord INTEGER(4), fpos
  INTEGER(8) = 0) followed by irecord records; an absent section is the
  marker with irecord = 0 (:330-466, :1220-1266);
* section 2: (ibrav, indpg), (celldm(1:6)); 3: (nsp), (na(1:nsp));
  4/5/6: one record of 3 REAL(8) per atom; 7: (ecut, cdual, dual LOGICAL(4),
  nel, nr1s, nr2s, nr3s); 8: (n, nkpts0 = 0, ngwks, ngwls, nhgs, nhgls);
* section 9 through w_wfnio (src/wfnio_utils.mod.F90:124-139): marker
  (-(2 n + 3), pos = -1), (iwfntype = 1), (n, ngwks, icompr = 1, scale),
  (tlsd = .FALSE., nsup = 0, nsdown = 0), then per state wrwfn's records
  (nx = 2 ngwks), (icmp(1:2 nx)) -- with icompr = 1, compress copies the
  REAL(8) words unchanged (src/compress.F90:33-36); scale = |c| of the BLAS
  izamax element;
* section 10 through wr30wfn (src/wr30wfn_utils.mod.F90:637-649): marker
  (2 n + 1, 0), (n, ngwks, icompr, scale = |x| of the idamax element of the
  real view), the states as above;
* section 23: marker (1, 0), (f(1:n)) (src/wv30_utils.mod.F90:854-871).
"""
from __future__ import annotations

import struct

import numpy as np


def record(payload: bytes) -> bytes:
    n = struct.pack("<i", len(payload))
    return n + payload + n


def records(data: bytes) -> list[bytes]:
    """Split a sequential unformatted file into its record payloads, checking
    that the leading and trailing length words agree."""
    out, pos = [], 0
    while pos < len(data):
        (n,) = struct.unpack_from("<i", data, pos)
        assert n >= 0 and pos + 8 + n <= len(data), "bad record framing"
        payload = data[pos + 4:pos + 4 + n]
        (m,) = struct.unpack_from("<i", data, pos + 4 + n)
        assert m == n, "record length words differ"
        out.append(payload)
        pos += 8 + n
    return out


def _marker(irecord: int, fpos: int = 0) -> bytes:
    return record(struct.pack("<iq", irecord, fpos))


def _states(c: np.ndarray) -> bytes:
    out = b""
    ngw = c.shape[0]
    for i in range(c.shape[1]):
        out += record(struct.pack("<i", 2 * ngw))
        out += record(np.ascontiguousarray(c[:, i], dtype=np.complex128).tobytes())
    return out


def wv30_bytes(h: dict, tau0, c0=None, cm=None, velp=None, taui=None, f=None) -> bytes:
    """The bytes zhwwf writes for header h (the fields of sections 2, 3, 7, 8)."""
    na = [int(x) for x in h["na"]]
    nat = sum(na)
    n, ngw = int(h["nstate"]), int(h["ngwks"])
    out = record(b" STUTTGART VERSION 3.0".ljust(80, b" "))
    out += _marker(2) + record(struct.pack("<ii", h["ibrav"], h["indpg"]))
    out += record(struct.pack("<6d", *h["celldm"]))
    out += _marker(2) + record(struct.pack("<i", len(na))) + record(struct.pack(f"<{len(na)}i", *na))
    for t in (tau0, velp, taui):
        if t is None:
            out += _marker(0)
        else:
            t = np.asarray(t, dtype=np.float64).reshape(nat, 3)
            out += _marker(nat) + b"".join(record(struct.pack("<3d", *t[a])) for a in range(nat))
    nr = h["nr"]
    out += _marker(1) + record(struct.pack("<ddiiiii", h["ecut"], h["cdual"], 1 if h["dual"] else 0, h["nel"],
                                           nr[0], nr[1], nr[2]))
    out += _marker(1) + record(struct.pack("<6i", n, 0, ngw, h["ngwls"], h["nhgs"], h["nhgls"]))
    if c0 is None:
        out += _marker(0, -1)
    else:
        c = np.asarray(c0, dtype=np.complex128)[:ngw, :n]
        flat = c.T.ravel()
        k = int(np.argmax(np.abs(flat.real) + np.abs(flat.imag)))   # izamax: first maximum
        out += _marker(-(2 * n + 3), -1) + record(struct.pack("<i", 1))
        out += record(struct.pack("<iiid", n, ngw, 1, abs(flat[k])))
        out += record(struct.pack("<iii", 0, 0, 0)) + _states(c)
    if cm is None:
        out += _marker(0)
    else:
        c = np.asarray(cm, dtype=np.complex128)[:ngw, :n]
        scale = float(max(np.abs(c.real).max(), np.abs(c.imag).max())) if c.size else 0.0
        out += _marker(2 * n + 1) + record(struct.pack("<iiid", n, ngw, 1, scale)) + _states(c)
    for _ in range(11, 23):
        out += _marker(0)
    if f is None:
        out += _marker(0)
    else:
        out += _marker(1) + record(np.ascontiguousarray(f[:n], dtype=np.float64).tobytes())
    for _ in range(24, 100):
        out += _marker(0)
    return out
