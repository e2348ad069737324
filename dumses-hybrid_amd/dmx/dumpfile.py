"""Reference binary dump format (output/restart), read and written in numpy.

The reference writes each rank's state as a Fortran unformatted sequential
file ``output_NNNNNN/data.RRRRRR.bin`` (src/modules/file.f90:21-70) with five
records (writer src/subroutines/output.f90:224-235, reader
src/subroutines/restart.f90:212-223):

  1. int32[6]   nx, ny, nz, nxslice, nyslice, nzslice
  2. float64[5] time, dt, dx, dy, dz
  3. float64[6] xmin, xmax, ymin, ymax, zmin, zmax
  4. float64[2] gamma, ciso
  5. float64   uin(-2:nx+3, -2:ny+3, -2:nz+3, 1:11), column-major

Each record is framed by 4-byte little-endian length markers.  In numpy the
uin record is viewed as an array ``u[var, k, j, i]`` (C order), which is the
same memory as the Fortran array.
"""
from __future__ import annotations

import os
import struct
from dataclasses import dataclass, field

import numpy as np


@dataclass
class Dump:
    nx: int
    ny: int
    nz: int
    time: float
    dt: float
    dx: float
    dy: float
    dz: float
    bounds: tuple  # xmin, xmax, ymin, ymax, zmin, zmax
    gamma: float
    ciso: float
    uin: np.ndarray  # (11, nz+6, ny+6, nx+6) float64
    slices: tuple = field(default=(1, 1, 1))


def _record(payload: bytes) -> bytes:
    n = struct.pack("<i", len(payload))
    return n + payload + n


def write_dump(path: str, d: Dump) -> None:
    """Writes ``d`` in the reference's binary restart layout."""
    u = np.ascontiguousarray(d.uin, dtype="<f8")
    if u.shape != (11, d.nz + 6, d.ny + 6, d.nx + 6):
        raise ValueError(f"uin shape {u.shape} does not match ({d.nx},{d.ny},{d.nz})")
    os.makedirs(os.path.dirname(os.path.abspath(path)), exist_ok=True)
    with open(path, "wb") as f:
        f.write(_record(struct.pack("<6i", d.nx, d.ny, d.nz, *d.slices)))
        f.write(_record(struct.pack("<5d", d.time, d.dt, d.dx, d.dy, d.dz)))
        f.write(_record(struct.pack("<6d", *d.bounds)))
        f.write(_record(struct.pack("<2d", d.gamma, d.ciso)))
        f.write(_record(u.tobytes()))


def read_dump(path: str) -> Dump:
    """Reads one rank's binary dump written by the reference or by write_dump."""
    with open(path, "rb") as f:
        raw = f.read()
    recs = []
    pos = 0
    while pos < len(raw):
        (n,) = struct.unpack_from("<i", raw, pos)
        recs.append(raw[pos + 4: pos + 4 + n])
        (n2,) = struct.unpack_from("<i", raw, pos + 4 + n)
        if n2 != n:
            raise ValueError(f"{path}: record marker mismatch at byte {pos}")
        pos += 8 + n
    if len(recs) != 5:
        raise ValueError(f"{path}: expected 5 records, found {len(recs)}")
    nx, ny, nz, sx, sy, sz = struct.unpack("<6i", recs[0])
    time, dt, dx, dy, dz = struct.unpack("<5d", recs[1])
    bounds = struct.unpack("<6d", recs[2])
    gamma, ciso = struct.unpack("<2d", recs[3])
    u = np.frombuffer(recs[4], dtype="<f8").reshape(11, nz + 6, ny + 6, nx + 6).copy()
    return Dump(nx, ny, nz, time, dt, dx, dy, dz, bounds, gamma, ciso, u, (sx, sy, sz))


def dump_path(ndump: int, rank: int = 0, root: str = ".") -> str:
    """The reference's file name for dump number ``ndump`` of rank ``rank``."""
    return os.path.join(root, f"output_{ndump:06d}", f"data.{rank:06d}.bin")
