"""Readers/writers of the reference's file formats.

All binary files are Fortran unformatted sequential with 4-byte little-endian record markers
(the gfortran/amdflang default):

* ``.bbb`` -- one record per realisation: 6N doubles, particle-major (x, y, z, u, v, w)
  (IC/bbb_maker_REV.f90:71, read at bia/bia.f90:64);
* ``.DIM`` -- record ``[t, 6N doubles]`` (dim/dim_3D.f90:32);
* ``.BIA`` -- record ``[dt, 6N doubles of deltas dX..dW]`` (bia/bia.f90:206);
* text ``BIA_10^k.NNNNNNN`` / ``DIM_10^k.NNNNNNN`` -- ``(es23.16,6es24.16)`` of particle 1
  (bia/bia.f90:207, dim/dim_3D.f90:31).

Arrays returned here are structure-of-arrays ``[6, N]``.  The typed readers and writers go
through the library's C functions (``bia_records_scan``, ``bia_record_read``,
``bia_record_write``, csrc/bia_io.cpp), the ones C and C++ callers use; a malformed file raises
ValueError.  ``read_records`` / ``write_records`` handle raw records of any payload.
"""
from __future__ import annotations

import ctypes as C
import os
import struct

import numpy as np


def read_records(path: str) -> list[bytes]:
    out = []
    with open(path, "rb") as f:
        data = f.read()
    p = 0
    while p < len(data):
        if p + 4 > len(data):
            raise ValueError(f"{path}: truncated record marker at byte {p}")
        (n,) = struct.unpack_from("<i", data, p)
        if n < 0 or p + 8 + n > len(data):
            raise ValueError(f"{path}: record of {n} bytes at byte {p} does not fit the file ({len(data)} bytes)")
        rec = data[p + 4:p + 4 + n]
        (n2,) = struct.unpack_from("<i", data, p + 4 + n)
        if n2 != n:
            raise ValueError(f"{path}: record marker mismatch at byte {p}")
        out.append(rec)
        p += 8 + n
    return out


def write_records(path: str, records: list[bytes]) -> None:
    with open(path, "wb") as f:
        for r in records:
            f.write(struct.pack("<i", len(r)))
            f.write(r)
            f.write(struct.pack("<i", len(r)))


def _native(status: int, path: str, what: str) -> None:
    from . import _lib
    if status == _lib.BIA_OK:
        return
    if status == _lib.BIA_ERR_IO and not os.path.exists(path):
        raise FileNotFoundError(path)
    raise ValueError(f"{path}: {what}: {_lib.load().bia_error_string(status).decode()}")


def _read(path: str, nhead: int) -> list[tuple[np.ndarray, np.ndarray]]:
    from . import _lib
    L, p = _lib.load(), os.fsencode(path)
    nrec, nmax = C.c_longlong(), C.c_longlong()
    _native(L.bia_records_scan(p, nhead, C.byref(nrec), C.byref(nmax)), path, "scan")
    out = []
    for k in range(nrec.value):
        n = C.c_longlong()
        _native(L.bia_record_read(p, nhead, k, None, None, 0, C.byref(n)), path, f"record {k}")
        head, soa = np.zeros(max(nhead, 1)), np.empty((6, n.value))
        _native(L.bia_record_read(p, nhead, k, head.ctypes.data, soa.ctypes.data, max(n.value, 1), C.byref(n)),
                path, f"record {k}")
        out.append((head[:nhead], soa))
    return out


def _write(path: str, recs: list[tuple[np.ndarray, np.ndarray]], nhead: int) -> None:
    from . import _lib
    L, p = _lib.load(), os.fsencode(path)
    if not recs:
        open(path, "wb").close()
        return
    for k, (head, s) in enumerate(recs):
        s = np.ascontiguousarray(np.asarray(s, dtype=np.float64).reshape(6, -1))
        head = np.ascontiguousarray(np.asarray(head, dtype=np.float64).reshape(-1))
        n = s.shape[1]
        _native(L.bia_record_write(p, 0 if k == 0 else 1, nhead, head.ctypes.data if nhead else None, s.ctypes.data,
                                   max(n, 1), n), path, f"record {k}")


def read_bbb(path: str) -> list[np.ndarray]:
    """All realisations of a ``.bbb`` file as ``[6, N]`` arrays."""
    return [soa for _, soa in _read(path, 0)]


def write_bbb(path: str, states: list[np.ndarray]) -> None:
    _write(path, [(np.zeros(0), s) for s in states], 0)


def read_timed(path: str) -> list[tuple[float, np.ndarray]]:
    """``.DIM`` / ``.BIA`` records: list of (t or dt, ``[6, N]``)."""
    return [(float(h[0]), soa) for h, soa in _read(path, 1)]


def write_timed(path: str, recs: list[tuple[float, np.ndarray]]) -> None:
    _write(path, [(np.array([t], dtype=np.float64), s) for t, s in recs], 1)


def format_particle_line(t: float, p6) -> str:
    """One line in the reference's ``(es23.16,6es24.16)`` text format."""

    def es(v: float, w: int) -> str:
        m, e = f"{v:.16E}".split("E")
        return f"{m}E{int(e):+03d}".rjust(w)

    return es(t, 23) + "".join(es(float(x), 24) for x in p6)


def read_particle_line(path: str) -> np.ndarray:
    """First data line of a ``BIA_10^k`` / ``DIM_10^k`` text file: [t, x, y, z, u, v, w]."""
    with open(path) as f:
        return np.array([float(x) for x in f.readline().split()])
