"""Readers/writers of the reference's file formats.

All binary files are Fortran unformatted sequential with 4-byte little-endian record markers
(the gfortran/amdflang default):

* ``.bbb`` -- one record per realisation: 6N doubles, particle-major (x, y, z, u, v, w)
  (IC/bbb_maker_REV.f90:71, read at bia/bia.f90:64);
* ``.DIM`` -- record ``[t, 6N doubles]`` (dim/dim_3D.f90:32);
* ``.BIA`` -- record ``[dt, 6N doubles of deltas dX..dW]`` (bia/bia.f90:206);
* text ``BIA_10^k.NNNNNNN`` / ``DIM_10^k.NNNNNNN`` -- ``(es23.16,6es24.16)`` of particle 1
  (bia/bia.f90:207, dim/dim_3D.f90:31).

Arrays returned here are structure-of-arrays ``[6, N]``.
"""
from __future__ import annotations

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


def _aos_to_soa(v: np.ndarray) -> np.ndarray:
    return np.ascontiguousarray(v.reshape(-1, 6).T)


def _soa_to_aos(s: np.ndarray) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(s, dtype="<f8").reshape(6, -1).T).reshape(-1)


def read_bbb(path: str) -> list[np.ndarray]:
    """All realisations of a ``.bbb`` file as ``[6, N]`` arrays."""
    return [_aos_to_soa(np.frombuffer(r, dtype="<f8").astype(np.float64)) for r in read_records(path)]


def write_bbb(path: str, states: list[np.ndarray]) -> None:
    write_records(path, [_soa_to_aos(s).tobytes() for s in states])


def read_timed(path: str) -> list[tuple[float, np.ndarray]]:
    """``.DIM`` / ``.BIA`` records: list of (t or dt, ``[6, N]``)."""
    out = []
    for r in read_records(path):
        v = np.frombuffer(r, dtype="<f8").astype(np.float64)
        out.append((float(v[0]), _aos_to_soa(v[1:])))
    return out


def write_timed(path: str, recs: list[tuple[float, np.ndarray]]) -> None:
    write_records(path, [np.concatenate([[t], _soa_to_aos(s)]).astype("<f8").tobytes() for t, s in recs])


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
