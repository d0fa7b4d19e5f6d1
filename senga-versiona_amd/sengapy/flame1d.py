"""Laminar flame speed and thickness of a single-step premixed flame (host side).

Used to set the turbulence of the single-step periodic case from flame
quantities (u'/S_L, l/delta).  The flame is the steady, planar, isobaric
solution of the reference's equations for a mechanism R -> P with one
irreversible step, equal molar masses, unity Lewis number and the constant-
Lewis transport law lambda/c_p = ALAMDC (T/TLAMDA)^RLAMDA
(src/rhscal.f:515-525, src/com_senga2.h:112-115):

    mdot dc/dx = d/dx(Gamma dc/dx) + omega(c),   Gamma = lambda/c_p = rho D,

with c the product mass fraction, omega = A T^n exp(-E/(R0 T)) rho (1 - c)
the product mass source (src/chrate.f:227-242, 451-470, 731-747; the molar
masses cancel), T(c) from constant total enthalpy (unity Lewis number and equal
c_p polynomials of R and P), and rho = p W / (R0 T).  With the diffusive flux
F = Gamma dc/dx as unknown and c as independent variable the problem is the
phase-plane eigenvalue problem (Williams, Combustion Theory, 2nd ed., 5.3)

    dF/dc = mdot - Gamma(c) omega(c) / F,   F(0) = 0, F(1) = 0,

solved by shooting from the burnt end (F ~ b (1 - c), b^2 + mdot b = Gamma a
with omega ~ a (1 - c)) and bisection on mdot.  S_L = mdot / rho_u; the
thermal thickness is delta = (T_b - T_u) / max dT/dx with dT/dx = T'(c) F / Gamma.
"""
from __future__ import annotations

import math

import numpy as np

from . import chem as _chem

ALAMDC, TLAMDA, RLAMDA = 2.58e-5, 298.0, 0.7


def _sensible_h(tables, T, s):
    """Enthalpy per unit mass of species s at T (AMASCH, interval by T)."""
    it = 0
    if tables.ntint[s] > 1 and T > tables.tinthi[0, s]:
        it = 1
    a = tables.amasch
    np5 = tables.ncpoly[0, s]
    f = a[np5 - 1, it, s]
    for c in range(np5 - 2, -1, -1):
        f = f * T + a[c, it, s]
    return a[tables.ncenth[0, s] - 1, it, s] + f * T


def single_step_flame(mech: _chem.Mechanism, Tu: float, p: float = 1.0e5, npts: int = 4001):
    """(S_L [m/s], thermal thickness [m], T_b [K], mdot [kg/m^2/s]) of the
    single-step mechanism `mech` (species 0 -> species 1) at unburnt
    temperature Tu, pressure p."""
    t = _chem.derive(mech)
    if mech.nspec != 2 or mech.nstep != 1:
        raise ValueError("single-step, two-species mechanisms only")
    W = float(mech.wmolar[0])
    if abs(mech.wmolar[1] - W) > 1e-12 * W:
        raise ValueError("equal molar masses assumed")
    A, nexp, E = (float(v) for v in mech.rparam[:, 0])
    Eovr = E / _chem.RGUNIV
    hu = _sensible_h(t, Tu, 0)

    def T_of_c(c):
        # total enthalpy (1-c) h_R(T) + c h_P(T) = h_R(Tu), by bisection
        lo, hi = 200.0, 5000.0
        for _ in range(100):
            mid = 0.5 * (lo + hi)
            h = (1.0 - c) * _sensible_h(t, mid, 0) + c * _sensible_h(t, mid, 1)
            if h > hu:
                hi = mid
            else:
                lo = mid
        return 0.5 * (lo + hi)

    c = np.linspace(0.0, 1.0, npts)
    T = np.array([T_of_c(x) for x in c])
    rho = p * W / (_chem.RGUNIV * T)
    gam = ALAMDC * (T / TLAMDA) ** RLAMDA
    om = A * T ** nexp * np.exp(-Eovr / T) * rho * (1.0 - c)
    dTdc = np.gradient(T, c)
    a_end = A * T[-1] ** nexp * math.exp(-Eovr / T[-1]) * rho[-1]   # omega ~ a (1 - c) at c -> 1

    def too_large(mdot):
        """Shoot from the burnt end towards the cold end: with mdot too large F
        reaches zero before c = 0; too small leaves F > mdot c near c = 0."""
        b = 0.5 * (-mdot + math.sqrt(mdot * mdot + 4.0 * gam[-1] * a_end))
        F = b * (c[-1] - c[-2])
        for i in range(npts - 3, 0, -1):
            # midpoint step of dF/dc = mdot - Gamma omega / F from c_i+1 down to c_i
            h = c[i + 1] - c[i]
            Fm = F - 0.5 * h * (mdot - gam[i + 1] * om[i + 1] / F)
            if Fm <= 0.0:
                return True
            F = F - h * (mdot - 0.5 * (gam[i] + gam[i + 1]) * 0.5 * (om[i] + om[i + 1]) / Fm)
            if F <= 0.0:
                return True
        return F < mdot * c[1]

    lo, hi = 1e-6, 1e3
    for _ in range(200):
        mid = math.sqrt(lo * hi)
        if too_large(mid):
            hi = mid
        else:
            lo = mid
        if hi / lo < 1 + 1e-12:
            break
    mdot = math.sqrt(lo * hi)
    # flux profile at the converged mdot, for the thickness
    F = np.zeros(npts)
    b = 0.5 * (-mdot + math.sqrt(mdot * mdot + 4.0 * gam[-1] * a_end))
    F[-2] = b * (c[-1] - c[-2])
    for i in range(npts - 3, -1, -1):
        h = c[i + 1] - c[i]
        Fi = max(F[i + 1], 1e-300)
        F[i] = max(F[i + 1] - h * (mdot - gam[i + 1] * om[i + 1] / Fi), 0.0)
    dTdx = dTdc * F / gam
    delta = (T[-1] - T[0]) / float(np.max(dTdx))
    return mdot / rho[0], delta, float(T[-1]), mdot
