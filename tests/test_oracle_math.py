"""The oracle against what the mathematics fixes (independent of the reference).

Derivative operators: closed forms on polynomials (exactness orders of each
row), the modified wavenumber of sin(kx), and the structure of the diagonal
mixed stencil.  Point physics: T -> E(T) -> T round trip, single-step
Arrhenius closed form, mass and element conservation of the 21-step H2
chemistry.  RHS: exact mass conservation on a periodic box and
sum_k Y_k-RHS = rho-RHS.
"""
import math

import numpy as np
import pytest

import golden_cases
from conftest import rel_err
from pyoracle import Oracle
from sengapy import case as C
from sengapy import ic

A = [5.0 / 3.0, -20.0 / 21.0, 5.0 / 14.0, -5.0 / 63.0, 1.0 / 126.0]   # a_m of the 10th-order scheme


def bounded_case(nx=24, ny=12, nz=11):
    return C.h2_nscbc(n=(nx, ny, nz), length=(2.3e-3, 1.1e-3, 1.0e-3), mixavg=False, nsteps=1)


def periodic_case(nx=20, ny=16, nz=12):
    return C.h2_nscbc(n=(nx, ny, nz), length=(1.9e-3, 1.5e-3, 1.1e-3), mixavg=False, periodic_x=True,
                      nsteps=1)


# exactness degree of each x row of the bounded first / second derivative
D1_DEG = [4, 4, 4, 6, 8]
D2_DEG = [4, 4, 4, 6, 8]   # one-sided second derivative rows are 4th order


@pytest.mark.parametrize("p", range(0, 11))
def test_first_second_derivative_polynomial_exactness(p, oracle_built):
    c = bounded_case()
    o = Oracle(c)
    nx, ny, nz = c.shape
    h = c.control.xgdlen / (nx - 1)
    x = np.arange(nx) * h
    s = x / x[-1]                       # scaled coordinate keeps values O(1)
    f = np.broadcast_to(s ** p, (nz, ny, nx)).copy()
    out = o.derivatives(f)
    d1 = p * s ** max(p - 1, 0) / x[-1] if p > 0 else 0 * s
    d2 = p * (p - 1) * s ** max(p - 2, 0) / x[-1] ** 2 if p > 1 else 0 * s
    scale1 = max(1.0, p) / x[-1]
    scale2 = max(1.0, p * p) / x[-1] ** 2
    for i in range(nx):
        row = min(i, nx - 1 - i)
        deg = 10 if row >= 5 else D1_DEG[row]
        if p <= deg:
            assert abs(out[0, 3, 4, i] - d1[i]) <= 1e-9 * scale1, (p, i)
        deg2 = 10 if row >= 5 else D2_DEG[row]
        if p <= deg2:
            assert abs(out[3, 3, 4, i] - d2[i]) <= 1e-7 * scale2, (p, i)


def test_modified_wavenumber_periodic(oracle_built):
    c = periodic_case()
    o = Oracle(c)
    nx, ny, nz = c.shape
    hs = [c.control.xgdlen / (nx - 1), c.control.ygdlen / (ny - 1), c.control.zgdlen / (nz - 1)]
    per = [nx * hs[0], ny * hs[1], nz * hs[2]]
    kx, ky, kz = 2 * np.pi * 3 / per[0], 2 * np.pi * 2 / per[1], 2 * np.pi * 1 / per[2]
    x = np.arange(nx) * hs[0]
    y = np.arange(ny) * hs[1]
    z = np.arange(nz) * hs[2]
    Z, Y, X = np.meshgrid(z, y, x, indexing="ij")
    f = np.sin(kx * X) * np.sin(ky * Y) * np.sin(kz * Z)
    out = o.derivatives(f)

    def g1(k, h):   # modified wavenumber of the first derivative
        return sum(A[m - 1] / m * math.sin(m * k * h) for m in range(1, 6)) / h

    def g2(k, h):   # of the second derivative: sum a_m/m^2 * 2(cos(mkh)-1)/h^2
        return sum(A[m - 1] / (m * m) * 2.0 * (math.cos(m * k * h) - 1.0) for m in range(1, 6)) / h / h

    sx, sy, sz = np.sin(kx * X), np.sin(ky * Y), np.sin(kz * Z)
    cx, cy, cz = np.cos(kx * X), np.cos(ky * Y), np.cos(kz * Z)
    exp = [g1(kx, hs[0]) * cx * sy * sz, g1(ky, hs[1]) * sx * cy * sz, g1(kz, hs[2]) * sx * sy * cz,
           g2(kx, hs[0]) * f, g2(ky, hs[1]) * f, g2(kz, hs[2]) * f]

    def gm(k1, h1, k2, h2):   # diagonal mixed stencil: sum a_m/(4m^2) * 4 sin(m k1 h1) sin(m k2 h2)
        return sum(A[m - 1] / (m * m) * math.sin(m * k1 * h1) * math.sin(m * k2 * h2)
                   for m in range(1, 6)) / (h1 * h2)

    exp += [gm(kx, hs[0], ky, hs[1]) * cx * cy * sz, gm(kx, hs[0], kz, hs[2]) * cx * sy * cz,
            gm(ky, hs[1], kz, hs[2]) * sx * cy * cz]
    for q in range(9):
        assert rel_err(out[q], exp[q]) < 1e-11, q


def test_mixed_derivative_of_x_times_g_bounded(oracle_built):
    """d2/dxdy of x*g(y) with a bounded x direction equals the centred y
    difference of g of the order the x row uses (4th at rows 0-2, 6th at row 3,
    8th at row 4, 10th inside), because the stencil is exact for linear x."""
    c = bounded_case()
    o = Oracle(c)
    nx, ny, nz = c.shape
    hx, hy = c.control.xgdlen / (nx - 1), c.control.ygdlen / (ny - 1)
    x = np.arange(nx) * hx
    y = np.arange(ny) * hy
    k = 2 * np.pi * 2 / (ny * hy)
    g = np.sin(k * y)
    f = np.broadcast_to(x[None, None, :] * g[None, :, None], (nz, ny, nx)).copy()
    out = o.derivatives(f)
    coef = {4: [2.0 / 3.0, -1.0 / 12.0], 6: [3.0 / 4.0, -3.0 / 20.0, 1.0 / 60.0],
            8: [4.0 / 5.0, -1.0 / 5.0, 4.0 / 105.0, -1.0 / 280.0], 10: [a / (2 * (m + 1)) for m, a in enumerate(A)]}

    def dy(order):
        cs = coef[order]
        return sum(cs[m] * (np.roll(g, -(m + 1)) - np.roll(g, m + 1)) for m in range(len(cs))) / hy

    for i in range(nx):
        row = min(i, nx - 1 - i)
        order = 10 if row >= 5 else (4 if row <= 2 else (6 if row == 3 else 8))
        assert np.max(np.abs(out[6, 2, :, i] - dy(order))) < 1e-8 * k, (i, order)


def test_temperature_round_trip(oracle_built):
    c = periodic_case()
    nx, ny, nz = c.shape
    rng = np.random.default_rng(3)
    t = c.tables
    Tt = rng.uniform(320.0, 2400.0, (nz, ny, nx))
    Yr = rng.uniform(0.0, 1.0, (t.nspec, nz, ny, nx)) ** 3
    Yr /= Yr.sum(axis=0)
    u = rng.normal(0, 5, (3, nz, ny, nx))
    st = ic.conservative_state(t, Tt, np.full(Tt.shape, 1.0e5), list(Yr), u[0], u[1], u[2])
    o = Oracle(c)
    o.set_state(st, 0.0)
    _, T0, _, _ = o.get()
    assert np.max(np.abs(T0 - Tt)) < 1e-5           # bisection on the residual to 1e-4
    o.rhs(1)                                         # one Newton solve from that guess
    _, T1, _, _ = o.get()
    assert np.max(np.abs(T1 - Tt) / Tt) < 1e-12


def test_single_step_arrhenius_closed_form(oracle_built):
    c = C.periodic_1step(n=(12, 11, 11))
    o = Oracle(c)
    o.set_state(c.dump.state, 0.0)
    o.rhs(1)
    _, T, _, rate = o.get()
    rho = c.dump.state[0]
    ry = c.dump.state[5]
    A_, n_, E_ = c.mech.rparam[:, 0]
    W = c.mech.wmolar[0]
    exp = -A_ * T ** n_ * np.exp(-E_ / 8314.2 / T) * np.maximum(ry / W, 0.0) * W
    assert rel_err(rate[0], exp) < 1e-12
    assert rel_err(rate[1], -exp) < 1e-12
    del rho


def test_h2_chemistry_conserves_mass_and_elements(oracle_built):
    c = periodic_case(12, 11, 11)
    o = Oracle(c)
    o.set_state(c.dump.state, 0.0)
    o.rhs(1)
    _, _, _, rate = o.get()
    W = c.mech.wmolar
    scale = np.max(np.abs(rate))
    assert np.max(np.abs(rate.sum(axis=0))) < 1e-12 * scale
    comp = {"H2": (2, 0), "O2": (0, 2), "H2O": (2, 1), "O": (0, 1), "OH": (1, 1), "H": (1, 0),
            "HO2": (1, 2), "H2O2": (2, 2), "N2": (0, 0)}
    for e in range(2):
        tot = sum(comp[s][e] * rate[i] / W[i] for i, s in enumerate(c.mech.species))
        assert np.max(np.abs(tot)) < 1e-11 * np.max(np.abs(rate / W[:, None, None, None]))


def test_rhs_invariants_periodic(oracle_built):
    c = periodic_case(12, 12, 11)
    o = Oracle(c)
    o.set_state(c.dump.state, 0.0)
    rhs = o.rhs(1)
    # mass: sum of -div(rho u) over the periodic box vanishes
    assert abs(rhs[0].sum()) < 1e-12 * np.abs(rhs[0]).sum()
    # sum_k Y_k-RHS = rho-RHS when sum_k Y_k = 1
    ysum = rhs[5:].sum(axis=0)
    assert rel_err(ysum, rhs[0]) < 1e-9


def test_uniform_state_has_zero_rhs(oracle_built):
    c = C.periodic_1step(n=(12, 11, 11))
    st = c.dump.state.copy()
    for v in range(st.shape[0]):
        st[v] = st[v].flat[0]
    st[1:4] = 0.0
    st[6] = 0.0
    st[5] = st[0]
    # a cold reactant at rest: no reaction to speak of, no gradients
    o = Oracle(c)
    o.set_state(st, 0.0)
    rhs = o.rhs(1)
    for v in range(5):
        assert np.max(np.abs(rhs[v])) < 1e-9 * max(1.0, np.max(np.abs(st[v])))
