"""EDI (SEG Electrical Data Interchange) files: MT responses of one station
per file, read and written by the host library (include/m3d_host.h,
m3dh_edi_*; the reference's io::read_edi / write_edi / load_edi_directory /
save_edi_directory, src/io/edi_io.cpp:156-619).

Impedances are held in ohm (EDI files carry mV/km/nT; the factor is
mu0 * 1e3 = 4 pi 1e-4), errors are standard deviations (EDI files carry
variances), and a component without data has weight 0.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field

import numpy as np

from .native import _check_host as _check, _ptr, lib

NAME_LEN = 128


@dataclass
class EDIData:
    """Stations x frequencies of MT responses (index [f, s] for the arrays)."""
    names: list
    geo: np.ndarray          # [ns, 4]: lat, lon, elevation, has_geo
    freqs: np.ndarray        # [nf] Hz
    z: np.ndarray            # [nf, ns, 4] complex, ohm (Zxx, Zxy, Zyx, Zyy)
    z_err: np.ndarray        # [nf, ns, 4]
    z_w: np.ndarray          # [nf, ns, 4]
    t: np.ndarray = field(default=None)      # [nf, ns, 2] complex (Tx, Ty)
    t_err: np.ndarray = field(default=None)
    t_w: np.ndarray = field(default=None)


def _bind():
    L = lib()
    vp, i, cp = C.c_void_p, C.c_int, C.c_char_p
    L.m3dh_edi_read.argtypes = [cp, C.POINTER(i), vp, i] + [vp] * 8
    L.m3dh_edi_write.argtypes = [cp, cp, vp, i] + [vp] * 7
    L.m3dh_edi_load_directory.argtypes = [cp, C.POINTER(i), C.POINTER(i), vp, i] + [vp] * 8
    L.m3dh_edi_save_directory.argtypes = [cp, i, i, vp, i] + [vp] * 8
    return L


def _alloc(ns, nf):
    return dict(geo=np.zeros((ns, 4)), freqs=np.zeros(nf), z=np.zeros((nf, ns, 4), np.complex128),
                z_err=np.zeros((nf, ns, 4)), z_w=np.zeros((nf, ns, 4)), t=np.zeros((nf, ns, 2), np.complex128),
                t_err=np.zeros((nf, ns, 2)), t_w=np.zeros((nf, ns, 2)))


def read_edi(path: str) -> EDIData:
    """One station (ns = 1), frequencies in file order."""
    L = _bind()
    nf = C.c_int()
    _check(L.m3dh_edi_read(str(path).encode(), C.byref(nf), None, 0, *([None] * 8)))
    a = _alloc(1, nf.value)
    name = C.create_string_buffer(NAME_LEN)
    _check(L.m3dh_edi_read(str(path).encode(), C.byref(nf), name, NAME_LEN, _ptr(a["geo"]), _ptr(a["freqs"]),
                           _ptr(a["z"]), _ptr(a["z_err"]), _ptr(a["z_w"]), _ptr(a["t"]), _ptr(a["t_err"]),
                           _ptr(a["t_w"])))
    return EDIData([name.value.decode()], **a)


def write_edi(path: str, d: EDIData, station: int = 0) -> None:
    """Station `station` of d to one file."""
    L = _bind()
    nf = len(d.freqs)
    s = station
    geo = np.ascontiguousarray(d.geo[s:s + 1], dtype=np.float64)
    z = np.ascontiguousarray(d.z[:, s], dtype=np.complex128)
    ze = np.ascontiguousarray(d.z_err[:, s], dtype=np.float64)
    zw = np.ascontiguousarray(d.z_w[:, s], dtype=np.float64)
    tip = d.t is not None
    t = np.ascontiguousarray(d.t[:, s], dtype=np.complex128) if tip else None
    te = np.ascontiguousarray(d.t_err[:, s], dtype=np.float64) if tip else None
    tw = np.ascontiguousarray(d.t_w[:, s], dtype=np.float64) if tip else None
    fr = np.ascontiguousarray(d.freqs, dtype=np.float64)
    _check(L.m3dh_edi_write(str(path).encode(), d.names[s].encode(), _ptr(geo), nf, _ptr(fr), _ptr(z), _ptr(ze),
                            _ptr(zw), _ptr(t) if tip else None, _ptr(te) if tip else None,
                            _ptr(tw) if tip else None))


def load_directory(path: str) -> EDIData:
    """Every *.edi of a directory: the union of the stations' frequencies,
    highest first, weight 0 where a station lacks a frequency."""
    L = _bind()
    ns, nf = C.c_int(), C.c_int()
    _check(L.m3dh_edi_load_directory(str(path).encode(), C.byref(ns), C.byref(nf), None, 0, *([None] * 8)))
    a = _alloc(ns.value, nf.value)
    names = C.create_string_buffer(NAME_LEN * max(ns.value, 1))
    _check(L.m3dh_edi_load_directory(str(path).encode(), C.byref(ns), C.byref(nf), names, NAME_LEN,
                                     _ptr(a["geo"]), _ptr(a["freqs"]), _ptr(a["z"]), _ptr(a["z_err"]),
                                     _ptr(a["z_w"]), _ptr(a["t"]), _ptr(a["t_err"]), _ptr(a["t_w"])))
    raw = names.raw
    nm = [raw[k * NAME_LEN:(k + 1) * NAME_LEN].split(b"\0", 1)[0].decode() for k in range(ns.value)]
    return EDIData(nm, **a)


def save_directory(path: str, d: EDIData) -> None:
    """One <name>.edi per station (station_<i>.edi when unnamed)."""
    L = _bind()
    ns, nf = len(d.names), len(d.freqs)
    names = C.create_string_buffer(NAME_LEN * max(ns, 1))
    for k, n in enumerate(d.names):
        b = n.encode()[:NAME_LEN - 1]
        C.memmove(C.addressof(names) + k * NAME_LEN, b, len(b))
    tip = d.t is not None
    arr = lambda x, dt: np.ascontiguousarray(x, dtype=dt)
    geo, fr = arr(d.geo, np.float64), arr(d.freqs, np.float64)
    z, ze, zw = arr(d.z, np.complex128), arr(d.z_err, np.float64), arr(d.z_w, np.float64)
    t = arr(d.t, np.complex128) if tip else None
    te = arr(d.t_err, np.float64) if tip else None
    tw = arr(d.t_w, np.float64) if tip else None
    _check(L.m3dh_edi_save_directory(str(path).encode(), ns, nf, names, NAME_LEN, _ptr(geo), _ptr(fr), _ptr(z),
                                     _ptr(ze), _ptr(zw), _ptr(t) if tip else None, _ptr(te) if tip else None,
                                     _ptr(tw) if tip else None))
