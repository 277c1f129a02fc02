"""CPMD RESTART files through the library (mc_restart_* of the C ABI).

``write_restart`` / ``read_restart`` take and return host arrays in CPMD's
layout: coefficients (ngw, nstate) complex128 in CPMD G order, positions
(nat, 3) in bohr in isa order.  The header fields are those of wv30's
sections 2, 3, 7 and 8 (src/wv30_utils.mod.F90:322-466):

    ibrav, indpg, celldm (6), na (per species), ecut, cdual, dual, nel,
    nr (mesh), ngwks, ngwls, nhgs, nhgls.

``header_for(sysm)`` fills them for a synthetic system.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from ._lib import check, lib

VELP, TAUI, C0, CM, OCC = 1, 2, 4, 8, 16
_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)


def _p(a, t=_dp):
    return None if a is None else a.ctypes.data_as(t)


def shell_count(hg: np.ndarray, delta: float = 1e-6) -> int:
    """Number of |G|^2 shells (rggen's nhgl, src/rggen_utils.mod.F90:175-183)."""
    hg = np.asarray(hg)
    if hg.size == 0:
        return 0
    return 1 + int(np.count_nonzero(np.abs(np.diff(hg)) > delta))


def header_for(sysm, nstate: int | None = None) -> dict:
    """wv30 header fields of a synthetic cubic system (ibrav 1, celldm (a, 1, 1, 0, 0, 0))."""
    gs = sysm.gset
    cell = sysm.cell
    n = sysm.nstate if nstate is None else nstate
    ngwl = shell_count(gs.hg[:gs.ngw])
    return dict(ibrav=1, indpg=1, celldm=[cell.alat, 1.0, 1.0, 0.0, 0.0, 0.0], na=[sysm.nat],
                ecut=cell.ecut, cdual=cell.dual, dual=False, nel=int(round(float(np.sum(sysm.f[:n])))),
                nr=tuple(cell.nr), ngwks=gs.ngw, ngwls=ngwl, nhgs=gs.nhg, nhgls=shell_count(gs.hg), nstate=n)


def _header(h: dict, sections: int) -> _lib.mc_restart_header:
    H = _lib.mc_restart_header()
    H.ibrav, H.indpg = h["ibrav"], h["indpg"]
    for k in range(6):
        H.celldm[k] = h["celldm"][k]
    H.nsp = len(h["na"])
    H.nat = int(np.sum(h["na"]))
    H.ecut, H.cdual, H.dual, H.nel = h["ecut"], h["cdual"], int(bool(h["dual"])), h["nel"]
    H.nr1s, H.nr2s, H.nr3s = h["nr"]
    H.nstate, H.ngwks, H.ngwls, H.nhgs, H.nhgls = h["nstate"], h["ngwks"], h["ngwls"], h["nhgs"], h["nhgls"]
    H.sections = sections
    return H


def _to_dict(H) -> dict:
    return dict(ibrav=H.ibrav, indpg=H.indpg, celldm=list(H.celldm), nsp=H.nsp, nat=H.nat, ecut=H.ecut,
                cdual=H.cdual, dual=bool(H.dual), nel=H.nel, nr=(H.nr1s, H.nr2s, H.nr3s), nstate=H.nstate,
                ngwks=H.ngwks, ngwls=H.ngwls, nhgs=H.nhgs, nhgls=H.nhgls, sections=H.sections)


def _arrays(h, tau0, velp, taui, f, c0, cm):
    na = np.ascontiguousarray(h["na"], dtype=np.int32)
    t = np.ascontiguousarray(tau0, dtype=np.float64)
    v = None if velp is None else np.ascontiguousarray(velp, dtype=np.float64)
    ti = None if taui is None else np.ascontiguousarray(taui, dtype=np.float64)
    ff = None if f is None else np.ascontiguousarray(f, dtype=np.float64)
    a = None if c0 is None else np.asfortranarray(c0, dtype=np.complex128)
    b = None if cm is None else np.asfortranarray(cm, dtype=np.complex128)
    sec = (VELP if v is not None else 0) | (TAUI if ti is not None else 0) | (OCC if ff is not None else 0)
    sec |= (C0 if a is not None else 0) | (CM if b is not None else 0)
    return na, t, v, ti, ff, a, b, sec


def write_restart(path: str, h: dict, tau0, c0=None, cm=None, velp=None, taui=None, f=None):
    na, t, v, ti, ff, a, b, sec = _arrays(h, tau0, velp, taui, f, c0, cm)
    H = _header(h, sec)
    ldc = a.shape[0] if a is not None else (b.shape[0] if b is not None else h["ngwks"])
    check(lib().mc_restart_write(str(path).encode(), C.byref(H), _p(na, _ip), _p(t), _p(v), _p(ti), _p(ff),
                                 _p(a), _p(b), ldc))


def read_header(path: str) -> dict:
    H = _lib.mc_restart_header()
    check(lib().mc_restart_read_header(str(path).encode(), C.byref(H)))
    return _to_dict(H)


def read_restart(path: str) -> dict:
    """Header fields plus 'na', 'tau0', and whichever of 'velp', 'taui', 'f',
    'c0', 'cm' the file holds."""
    h = read_header(path)
    n, nat, ngw = h["nstate"], h["nat"], h["ngwks"]
    sec = h["sections"]
    na = np.zeros(max(h["nsp"], 1), np.int32)
    tau0 = np.zeros((max(nat, 1), 3))
    velp = np.zeros((max(nat, 1), 3)) if sec & VELP else None
    taui = np.zeros((max(nat, 1), 3)) if sec & TAUI else None
    f = np.zeros(max(n, 1)) if sec & OCC else None
    c0 = np.zeros((ngw, max(n, 1)), np.complex128, order="F") if sec & C0 else None
    cm = np.zeros((ngw, max(n, 1)), np.complex128, order="F") if sec & CM else None
    H = _lib.mc_restart_header()
    check(lib().mc_restart_read(str(path).encode(), C.byref(H), _p(na, _ip), _p(tau0), _p(velp), _p(taui), _p(f),
                                _p(c0), _p(cm), ngw))
    out = _to_dict(H)
    out.update(na=na[:h["nsp"]], tau0=tau0[:nat])
    for k, v in (("velp", velp), ("taui", taui)):
        if v is not None:
            out[k] = v[:nat]
    if f is not None:
        out["f"] = f[:n]
    if c0 is not None:
        out["c0"] = c0[:, :n]
    if cm is not None:
        out["cm"] = cm[:, :n]
    return out


def write_restart_dev(ctx, path: str, h: dict, tau0, c0_dev=None, cm_dev=None, velp=None, taui=None, f=None):
    """As write_restart with device wavefunctions (ops.DeviceArray) of ctx."""
    na, t, v, ti, ff, _, _, sec = _arrays(h, tau0, velp, taui, f, None, None)
    sec |= (C0 if c0_dev is not None else 0) | (CM if cm_dev is not None else 0)
    H = _header(h, sec)
    check(lib().mc_restart_write_dev(ctx.h, str(path).encode(), C.byref(H), _p(na, _ip), _p(t), _p(v), _p(ti),
                                     _p(ff), c0_dev.ptr if c0_dev is not None else None,
                                     cm_dev.ptr if cm_dev is not None else None))


def read_restart_dev(ctx, path: str, c0_dev=None, cm_dev=None) -> dict:
    """Reads the header, positions, velocities and occupations to the host and
    uploads C0 / CM into the given device arrays of ctx."""
    h = read_header(path)
    nat, n = h["nat"], h["nstate"]
    na = np.zeros(max(h["nsp"], 1), np.int32)
    tau0 = np.zeros((max(nat, 1), 3))
    velp = np.zeros((max(nat, 1), 3))
    taui = np.zeros((max(nat, 1), 3))
    f = np.zeros(max(n, 1))
    H = _lib.mc_restart_header()
    check(lib().mc_restart_read_dev(ctx.h, str(path).encode(), C.byref(H), _p(na, _ip), _p(tau0), _p(velp),
                                    _p(taui), _p(f), c0_dev.ptr if c0_dev is not None else None,
                                    cm_dev.ptr if cm_dev is not None else None))
    out = _to_dict(H)
    out.update(na=na[:h["nsp"]], tau0=tau0[:nat], velp=velp[:nat], taui=taui[:nat], f=f[:n])
    return out
