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


def read_dump_formatted(path: str) -> Dump:
    """The formatted (list-directed) dump of src/output.f:318-333, read as
    src/indata.f:1449-1473 reads it: NX NY NZ NSPEC; then for each point in
    Fortran order (x fastest) its 5+NSPEC conserved values; then ETIME, TSTEP,
    ERROLD, ERRLDR.  Values may wrap over several lines (list-directed)."""
    toks = open(path).read().replace("D", "E").replace("d", "e").split()
    nx, ny, nz, ns = (int(t) for t in toks[:4])
    npts = nx * ny * nz
    nv = 5 + ns
    vals = np.array([float(t) for t in toks[4:4 + npts * nv + 4]])
    if vals.size != npts * nv + 4:
        raise ValueError("formatted dump is short")
    st = vals[:npts * nv].reshape(nz, ny, nx, nv).transpose(3, 0, 1, 2).copy()
    return Dump(st, *(float(v) for v in vals[npts * nv:]))


def write_dump_formatted(path: str, d: Dump) -> None:
    """Formatted dump: the layout read_dump_formatted reads, one point per
    line, values in full precision (repr)."""
    nv, nz, ny, nx = d.state.shape
    pts = d.state.reshape(nv, -1).T
    with open(path, "w") as f:
        f.write(f" {nx} {ny} {nz} {nv - 5}\n")
        for row in pts:
            f.write(" " + " ".join(repr(float(v)) for v in row) + "\n")
        f.write(" " + " ".join(repr(float(v)) for v in (d.etime, d.tstep, d.errold, d.errldr)) + "\n")


OUT_FIELDS = ("rho", "u", "v", "w", "e", "T", "p")


def read_out_res(raw, nx: int, ny: int, nz: int, ns: int) -> dict:
    """The post-processing file outIIIIIPPPP.res (src/output.f:241-265): one
    unformatted sequential record of rho, u, v, w, E per unit mass, T, p (each
    NX*NY*NZ FP64), Y_k and the reaction rates (NX*NY*NZ*NSPEC each, species
    last) and ETIME.  raw is the file's bytes or its path.  Arrays come back
    as (nz, ny, nx) or (nspec, nz, ny, nx)."""
    if isinstance(raw, str):
        raw = open(raw, "rb").read()
    n = int(np.frombuffer(raw[:4], "<i4")[0])
    npts = nx * ny * nz
    if n != 8 * (7 * npts + 2 * ns * npts + 1) or int(np.frombuffer(raw[4 + n:8 + n], "<i4")[0]) != n:
        raise ValueError("out*.res record does not match the grid and species count")
    v = np.frombuffer(raw[4:4 + n], "<f8")
    out = {}
    for k, nm in enumerate(OUT_FIELDS):
        out[nm] = v[k * npts:(k + 1) * npts].reshape(nz, ny, nx)
    o = 7 * npts
    out["Y"] = v[o:o + ns * npts].reshape(ns, nz, ny, nx)
    out["rate"] = v[o + ns * npts:o + 2 * ns * npts].reshape(ns, nz, ny, nx)
    out["etime"] = float(v[-1])
    return out
