"""Seeded initial-condition generators (host side).

These replace the reference's cold-start generators (src/turbin.f, src/flamin.f),
whose fields depend on the MPI decomposition and its per-rank RNG seeding
(src/turbin.f:126-180).  Every case here starts from a restart dump written by
these functions, so the GPU path and the reference read the identical state.

Grid convention of the reference: x_i = (i-1)*L/(N-1), i = 1..N
(src/dfinit.f:42-50); a periodic direction therefore has period N*L/(N-1).
"""
from __future__ import annotations

import math

import numpy as np


def grid(n: int, length: float) -> np.ndarray:
    return np.arange(n) * (length / float(n - 1))


def solenoidal_field(shape, spacing, urms, k0_len, seed, spectrum="passot-pouquet"):
    """Random solenoidal velocity on a periodic box (numpy FFT, host side).

    shape=(nz,ny,nx); spacing=(dz,dy,dx); the period in each direction is
    n*d.  Energy spectrum with peak wavenumber k0 = 2*pi/k0_len:
      passot-pouquet:     E(k) ~ (k/k0)^4 exp(-2 (k/k0)^2)
      batchelor-townsend: E(k) ~ (k/k0)^4 exp(-2 (k/k0)^2) (1 + (k/k0)^2)^(-1/2)
    Random phases from numpy's PCG64 with the given seed; the field is scaled so
    that each component has RMS urms.  Returns (u, v, w) arrays of `shape`.
    """
    rng = np.random.default_rng(seed)
    nz, ny, nx = shape
    dz, dy, dx = spacing
    kz = 2 * np.pi * np.fft.fftfreq(nz, d=dz)
    ky = 2 * np.pi * np.fft.fftfreq(ny, d=dy)
    kx = 2 * np.pi * np.fft.fftfreq(nx, d=dx)
    KZ, KY, KX = np.meshgrid(kz, ky, kx, indexing="ij")
    k2 = KX ** 2 + KY ** 2 + KZ ** 2
    k = np.sqrt(k2)
    k0 = 2 * np.pi / k0_len
    r = np.where(k > 0, k / k0, 0.0)
    ek = r ** 4 * np.exp(-2.0 * r ** 2)
    if spectrum == "batchelor-townsend":
        ek = ek / np.sqrt(1.0 + r ** 2)
    amp = np.where(k > 0, np.sqrt(ek / np.maximum(4 * np.pi * k2, 1e-300)), 0.0)
    comps = []
    for _ in range(3):
        ph = rng.standard_normal(shape) + 1j * rng.standard_normal(shape)
        comps.append(amp * ph)
    ux, uy, uz = comps
    # project out the dilatational part
    with np.errstate(invalid="ignore", divide="ignore"):
        dot = (KX * ux + KY * uy + KZ * uz) / np.where(k2 > 0, k2, 1.0)
    ux = ux - KX * dot
    uy = uy - KY * dot
    uz = uz - KZ * dot
    out = []
    for c in (ux, uy, uz):
        f = np.real(np.fft.ifftn(c))
        f -= f.mean()
        s = math.sqrt(float((f ** 2).mean()))
        out.append(f * (urms / s if s > 0 else 0.0))
    return out[0], out[1], out[2]


def erf_profile(x, x0, delta):
    """Progress variable 0.5*(1+erf((x-x0)/delta)) using math.erf."""
    return 0.5 * (1.0 + np.vectorize(math.erf)((x - x0) / delta))


def enthalpy(tables, T, Y):
    """Mixture enthalpy per unit mass from the derived tables (AMASCH),
    interval chosen as the reference does (first interval with T <= TINTHI)."""
    h = np.zeros_like(T)
    for i in range(tables.nspec):
        it = np.zeros(T.shape, dtype=int)
        for t in range(int(tables.ntint[i]) - 1):
            it = np.where(T > tables.tinthi[it, i], np.minimum(it + 1, tables.ntint[i] - 1), it)
        hk = np.zeros_like(T)
        np5 = tables.ncpoly[0, i]
        a = tables.amasch
        f = a[np5 - 1, it, i]
        for c in range(np5 - 2, -1, -1):
            f = f * T + a[c, it, i]
        hk = a[tables.ncenth[0, i] - 1, it, i] + f * T
        h += hk * Y[i]
    return h


def conservative_state(tables, T, p, Y, u, v, w):
    """(rho, rho u, rho v, rho w, rho E, rho Y_k) with E = h - p/rho + |u|^2/2."""
    rmix = np.zeros_like(T)
    for i in range(tables.nspec):
        rmix += tables.rgspec[i] * Y[i]
    rho = p / (rmix * T)
    h = enthalpy(tables, T, Y)
    E = h - rmix * T + 0.5 * (u * u + v * v + w * w)
    out = [rho, rho * u, rho * v, rho * w, rho * E] + [rho * Y[i] for i in range(tables.nspec)]
    return np.stack(out)


def adiabatic_temperature(tables, Yu, Yb, Tu):
    """Burnt temperature with the same enthalpy as the unburnt mixture at Tu."""
    hu = float(enthalpy(tables, np.array([Tu]), [np.array([y]) for y in Yu])[0])
    lo, hi = Tu, 3500.0
    for _ in range(200):
        mid = 0.5 * (lo + hi)
        hb = float(enthalpy(tables, np.array([mid]), [np.array([y]) for y in Yb])[0])
        if hb > hu:
            hi = mid
        else:
            lo = mid
    return 0.5 * (lo + hi)
