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


def _wavenumbers(shape, spacing):
    nz, ny, nx = shape
    dz, dy, dx = spacing
    kz = 2 * np.pi * np.fft.fftfreq(nz, d=dz)
    ky = 2 * np.pi * np.fft.fftfreq(ny, d=dy)
    kx = 2 * np.pi * np.fft.fftfreq(nx, d=dx)
    return np.meshgrid(kz, ky, kx, indexing="ij")


def espect(k, k0):
    """Energy spectrum shape of the reference's turbulence generator,
    E(k) ~ (k/k0)^4 exp(-2 (k/k0)^2) (src/espect.f:62-65, which cites
    Batchelor and Townsend 1948 for it; Passot and Pouquet 1987 use the same
    form).  Unnormalised: the field is scaled to its r.m.s. afterwards."""
    r = k / k0
    return r ** 4 * np.exp(-2.0 * r ** 2)


def solenoidal_field(shape, spacing, urms, k0_len, seed, spectrum="passot-pouquet"):
    """Random solenoidal velocity on a periodic box (numpy FFT, host side).

    shape=(nz,ny,nx); spacing=(dz,dy,dx); the period in each direction is
    n*d.  Energy spectrum espect() with peak wavenumber k0 = 2*pi/k0_len
    (spectrum "passot-pouquet" and "batchelor-townsend" name the same form,
    src/espect.f).  Each Fourier mode gets amplitude sqrt(E(k) / (4 pi k^2))
    and random complex Gaussian components (numpy PCG64, seed); the
    dilatational part is projected out, the real part taken (which keeps the
    field solenoidal: the projection is even in k), and all three components
    are scaled by ONE factor so that u' = sqrt(<u^2+v^2+w^2>/3) = urms (a
    per-component factor would reintroduce divergence).  Returns (u, v, w).
    """
    if spectrum not in ("passot-pouquet", "batchelor-townsend"):
        raise ValueError(spectrum)
    rng = np.random.default_rng(seed)
    KZ, KY, KX = _wavenumbers(shape, spacing)
    k2 = KX ** 2 + KY ** 2 + KZ ** 2
    k = np.sqrt(k2)
    k0 = 2 * np.pi / k0_len
    ek = np.where(k > 0, espect(k, k0), 0.0)
    amp = np.where(k > 0, np.sqrt(ek / np.maximum(4 * np.pi * k2, 1e-300)), 0.0)
    # no energy in Nyquist modes (even n): a real field's Nyquist mode has no
    # sine part, so its derivative -- and its divergence -- is undefined there
    for ax, nn in enumerate(shape):
        if nn % 2 == 0:
            sl = [slice(None)] * 3
            sl[ax] = nn // 2
            amp[tuple(sl)] = 0.0
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
    out = [np.real(np.fft.ifftn(c)) for c in (ux, uy, uz)]
    for f in out:
        f -= f.mean()
    s = math.sqrt(float(sum((f ** 2).mean() for f in out)) / 3.0)
    return tuple(f * (urms / s if s > 0 else 0.0) for f in out)


def integral_scale(u, v, w, spacing):
    """Longitudinal integral length scale of a periodic field from its
    discrete energy spectrum: l = (pi / (2 u'^2)) int E(k)/k dk
    = (3 pi / 4) sum_k e(k)/|k| / sum_k e(k), e(k) the modal energy."""
    KZ, KY, KX = _wavenumbers(u.shape, spacing)
    k = np.sqrt(KX ** 2 + KY ** 2 + KZ ** 2)
    e = sum(np.abs(np.fft.fftn(f)) ** 2 for f in (u, v, w))
    m = k > 0
    return 0.75 * np.pi * float(np.sum(e[m] / k[m]) / np.sum(e[m]))


def expected_integral_scale(shape, spacing, k0_len):
    """integral_scale() of the mean modal energy of solenoidal_field's
    ensemble: e(k) ~ E(k) / (4 pi k^2) (the projection removes the same share
    of every mode)."""
    KZ, KY, KX = _wavenumbers(shape, spacing)
    k = np.sqrt(KX ** 2 + KY ** 2 + KZ ** 2)
    m = k > 0
    e = espect(k[m], 2 * np.pi / k0_len) / (4 * np.pi * k[m] ** 2)
    return 0.75 * np.pi * float(np.sum(e / k[m]) / np.sum(e))


def k0_len_for_integral_scale(shape, spacing, target):
    """Peak wavelength 2 pi / k0 whose expected discrete integral scale is
    `target` (bisection; on a fine grid k0_len -> target * sqrt(2 pi))."""
    lo, hi = 1e-3 * target, 1e3 * target
    for _ in range(200):
        mid = math.sqrt(lo * hi)
        if expected_integral_scale(shape, spacing, mid) > target:
            hi = mid
        else:
            lo = mid
    return math.sqrt(lo * hi)


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
