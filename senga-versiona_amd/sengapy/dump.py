"""Restart dump (dmpi) files: the solution register in conservative form.

Unformatted layout, one sequential record (src/output.f:306-316 writes it,
src/indata.f:1430-1448 reads it):
    NX, NY, NZ, NSPEC (default INTEGER),
    DRUN, URUN, VRUN, WRUN, ERUN (each NX*NY*NZ FP64, x fastest),
    YRUN (NX*NY*NZ*NSPEC FP64, species last),
    ETIME, TSTEP, ERROLD, ERRLDR (FP64)
framed by 4-byte record-length markers.  The formatted variant
(src/output.f:318-333) writes the same values list-directed, one point per line.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np


@dataclass
class Dump:
    state: np.ndarray      # (5+N, NZ, NY, NX) float64: rho, rho u, rho v, rho w, rho E, rho Y_k
    etime: float = 0.0
    tstep: float = 0.0
    errold: float = 1.0e-4
    errldr: float = 1.0e-4

    @property
    def shape(self):
        return self.state.shape[3], self.state.shape[2], self.state.shape[1]


def write_dump(path: str, d: Dump) -> None:
    nv, nz, ny, nx = d.state.shape
    head = np.array([nx, ny, nz, nv - 5], dtype="<i4").tobytes()
    body = np.ascontiguousarray(d.state, dtype="<f8").tobytes()
    tail = np.array([d.etime, d.tstep, d.errold, d.errldr], dtype="<f8").tobytes()
    n = len(head) + len(body) + len(tail)
    if n >= 2 ** 31:
        raise ValueError("dump record exceeds the 2 GiB single-record limit")
    mark = np.array([n], dtype="<i4").tobytes()
    with open(path, "wb") as f:
        f.write(mark + head)
        f.write(body)
        f.write(tail + mark)


def read_dump(path: str) -> Dump:
    raw = open(path, "rb").read()
    n = int(np.frombuffer(raw[:4], "<i4")[0])
    if int(np.frombuffer(raw[4 + n:8 + n], "<i4")[0]) != n:
        raise ValueError("bad record markers")
    nx, ny, nz, ns = (int(v) for v in np.frombuffer(raw[4:20], "<i4"))
    npts = nx * ny * nz
    body = np.frombuffer(raw[20:20 + 8 * npts * (5 + ns)], "<f8")
    tail = np.frombuffer(raw[20 + 8 * npts * (5 + ns):20 + 8 * npts * (5 + ns) + 32], "<f8")
    st = body.reshape(5 + ns, nz, ny, nx).copy()
    return Dump(st, *(float(v) for v in tail))
