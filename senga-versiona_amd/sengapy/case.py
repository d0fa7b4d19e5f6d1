"""Case builders: control data, mechanism, transport data and a seeded initial dump.

Two case families match the workloads this library is measured on:

* ``periodic_1step``: triply periodic box, 2-species single-step irreversible
  Arrhenius chemistry (R -> P, Zeldovich number 6, heat-release parameter 0.75,
  unity Lewis number, Pr = 0.7), a planar flame slab in progress variable along
  x (two erf fronts so the profile is periodic) and seeded solenoidal
  turbulence with u'/S_L = 4 and l/delta = 2 from the computed laminar flame.
  At 32^3 with L = 2*pi mm this is the CPU-runnable workload.
* ``h2_nscbc``: 9-species 21-step H2-air mechanism (the reference's
  input/chem.dat), NSCBC inflow (type 13, specified density/velocity/species) at
  x-left and partially non-reflecting outflow (type 21) at x-right, periodic in
  y and z, premixed phi = 0.7 at T_u = 300 K and 1 atm, a planar erf flame with
  radical bumps, and seeded Passot-Pouquet turbulence.  At 128^3 this is the
  FP64 GPU workload.

The time step is fixed from the initial field: dt = CFL * min(dx) / max(|u|+a)
with CFL = 0.5 (the reference has no CFL logic; it takes a fixed TSTEP from
cont.dat when the step switch is 0).
"""
from __future__ import annotations

import math
import os
from dataclasses import dataclass
from typing import Optional

import numpy as np

from . import chem as _chem
from . import diff as _diff
from .cont import Control
from .dump import Dump, write_dump
from . import ic as _ic

DATA = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "data")


@dataclass
class Case:
    name: str
    control: Control
    mech: _chem.Mechanism
    tables: _chem.ChemTables
    diff: Optional[_diff.DiffTables]
    dump: Dump
    chem_path: str
    diff_path: Optional[str]
    radn_path: Optional[str] = None   # radn.dat, read when NFRADN = 1

    @property
    def nspec(self):
        return self.tables.nspec

    @property
    def radn(self):
        """The radiation data (sengapy.radn) when radn.dat is given, else None."""
        from .radn import read_radn
        return read_radn(self.radn_path) if self.radn_path else None

    @property
    def shape(self):
        return (self.control.nx, self.control.ny, self.control.nz)

    def write_inputs(self, rundir: str) -> None:
        """Lay out input/ and output/ as the reference expects (src/indata.f:54-136)."""
        os.makedirs(os.path.join(rundir, "input"), exist_ok=True)
        os.makedirs(os.path.join(rundir, "output"), exist_ok=True)
        self.control.write(os.path.join(rundir, "input", "cont.dat"))
        _chem.write_chem(self.mech, os.path.join(rundir, "input", "chem.dat"))
        if self.diff_path:
            with open(self.diff_path) as f, open(os.path.join(rundir, "input", "diff.dat"), "w") as g:
                g.write(f.read())
        if self.radn_path:
            with open(self.radn_path) as f, open(os.path.join(rundir, "input", "radn.dat"), "w") as g:
                g.write(f.read())
        write_dump(os.path.join(rundir, "output", "dmpi0000000.dat"), self.dump)


def sound_speed_max(tables, state, T=None):
    """max(|u|+a) over the field, a^2 = gamma R T with gamma from the mixture cp.
    T is recovered from the state by bisection when not given."""
    rho = state[0]
    u = state[1] / rho; v = state[2] / rho; w = state[3] / rho
    Y = [state[5 + i] / rho for i in range(tables.nspec)]
    if T is None:
        T = _case_temperature(tables, state)
    rmix = sum(tables.rgspec[i] * Y[i] for i in range(tables.nspec))
    cp = np.zeros_like(T)
    for i in range(tables.nspec):
        it = np.where(T > tables.tinthi[0, i], 1, 0) if tables.ntint[i] > 1 else np.zeros(T.shape, int)
        f = tables.amascp[4, it, i]
        for c in range(3, -1, -1):
            f = f * T + tables.amascp[c, it, i]
        cp += f * Y[i]
    a = np.sqrt(cp / (cp - rmix) * rmix * T)
    return float(np.max(np.maximum(np.abs(u), np.maximum(np.abs(v), np.abs(w))) + a))


def _case_temperature(tables, state):
    """Temperature from the conservative state by bisection on E(T) (host side,
    used only to size the time step)."""
    rho = state[0]
    Y = [state[5 + i] / rho for i in range(tables.nspec)]
    ke = 0.5 * (state[1] ** 2 + state[2] ** 2 + state[3] ** 2) / rho ** 2
    e = state[4] / rho - ke
    rmix = sum(tables.rgspec[i] * Y[i] for i in range(tables.nspec))
    lo = np.full(rho.shape, 200.0); hi = np.full(rho.shape, 3500.0)
    for _ in range(60):
        mid = 0.5 * (lo + hi)
        em = _ic.enthalpy(tables, mid, Y) - rmix * mid
        hi = np.where(em > e, mid, hi)
        lo = np.where(em > e, lo, mid)
    return 0.5 * (lo + hi)


def _f32(x: float) -> float:
    """Round a time step to the nearest float32 value.  On a restart with a
    fixed step the reference keeps cont.dat's TSTEP through an implicitly typed
    REAL variable (TSOLD, src/indata.f:1423,1481), so only steps exactly
    representable in single precision reach its RK coefficients unchanged."""
    return float(np.float32(x))


def one_step_mechanism() -> _chem.Mechanism:
    """R -> P, single irreversible step.  Thermo: the reference's 1-step case
    (src/casefiles/1step/chem.dat, W = 28.84 for both species); the product's
    enthalpy constant is set so that T_b = T_u/(1-alpha) = 1200 K from T_u = 300 K
    (alpha = 0.75), and E/R0 = Ze*T_b^2/(T_b-T_u) = 9600 K with Ze = 6."""
    base = _chem.read_chem(os.path.join(DATA, "onestep_base", "chem.dat"))
    m = base
    t = _chem.derive(m)
    Tu, Tb = 300.0, 1200.0
    # sensible enthalpy rise of the reactant from Tu to Tb sets the heat release
    hR_u = float(_ic.enthalpy(t, np.array([Tu]), [np.array([1.0]), np.array([0.0])])[0])
    hR_b = float(_ic.enthalpy(t, np.array([Tb]), [np.array([1.0]), np.array([0.0])])[0])
    hP_b = float(_ic.enthalpy(t, np.array([Tb]), [np.array([0.0]), np.array([1.0])])[0])
    # want h_P(Tb) = h_R(Tu): shift the product's a6 (both intervals) by dh*W/R0
    dh = hR_u - hP_b
    shift = dh * m.wmolar[1] / _chem.RGUNIV
    m.amolcp[5, :, 1] += shift
    m.clewis[:] = 1.0
    ze = 6.0
    m.rparam[:, 0] = [2.75e7, 0.0, ze * Tb * Tb / (Tb - Tu) * _chem.RGUNIV]
    del hR_b
    return m


def _fmt_roundtrip(m: _chem.Mechanism, path: str) -> _chem.Mechanism:
    """Write the mechanism to chem.dat and read it back, so that every consumer
    (reference, oracle, GPU) sees the values as printed in the file."""
    _chem.write_chem(m, path)
    return _chem.read_chem(path)


_FLAME_1STEP = {}


def one_step_flame(mech=None, Tu=300.0):
    """(S_L, thermal thickness delta, T_b) of the single-step mechanism's
    planar laminar flame at Tu and 1 atm (sengapy.flame1d), cached."""
    mech = mech or one_step_mechanism()
    key = (tuple(mech.rparam[:, 0]), float(mech.amolcp[5, 0, 1]), Tu)
    if key not in _FLAME_1STEP:
        from . import flame1d
        sl, delta, tb, _ = flame1d.single_step_flame(mech, Tu, npts=2001)
        _FLAME_1STEP[key] = (sl, delta, tb)
    return _FLAME_1STEP[key]


def periodic_1step(n=(32, 32, 32), length=2.0e-3 * math.pi, seed=1, uprime_over_sl=4.0, l_over_delta=2.0,
                   workdir=None, cfl=0.5, nsteps=10) -> Case:
    """The single-step flame slab in turbulence with u'/S_L and l/delta given
    (defaults 4 and 2): S_L and the thermal thickness delta = (T_b - T_u) /
    max dT/dx come from the mechanism's planar laminar flame (one_step_flame),
    u' = sqrt(<u^2+v^2+w^2>/3) is set to uprime_over_sl * S_L, the spectrum's
    peak wavelength so that the expected integral length scale of the discrete
    field is l_over_delta * delta, and the erf fronts have the laminar
    flame's thermal thickness (erf width delta / sqrt(pi))."""
    nx, ny, nz = n
    workdir = workdir or os.path.join(DATA, "_gen")
    os.makedirs(workdir, exist_ok=True)
    mech = _fmt_roundtrip(one_step_mechanism(), os.path.join(workdir, f"onestep_{nx}.chem.dat"))
    tab = _chem.derive(mech)
    Tu, Tb = 300.0, 1200.0
    sl, dth, _ = one_step_flame(mech, Tu)
    L = (length, length, length)
    d = tuple(L[i] / (n[i] - 1) for i in range(3))
    x = _ic.grid(nx, L[0])
    period = nx * d[0]
    width = dth / math.sqrt(math.pi)
    c = 0.5 * (np.vectorize(math.erf)((x - 0.3 * period) / width)
               - np.vectorize(math.erf)((x - 0.7 * period) / width))
    shape = (nz, ny, nx)
    C = np.broadcast_to(c, shape)
    T = Tu + (Tb - Tu) * C
    Y = [1.0 - C, C.copy()]
    spacing = (d[2], d[1], d[0])
    k0_len = _ic.k0_len_for_integral_scale(shape, spacing, l_over_delta * dth)
    u, v, w = _ic.solenoidal_field(shape, spacing, uprime_over_sl * sl, k0_len, seed,
                                   spectrum="batchelor-townsend")
    st = _ic.conservative_state(tab, T, np.full(shape, 1.0e5), Y, u, v, w)
    dt = _f32(cfl * min(d) / sound_speed_max(tab, st, T))
    ctl = Control(L[0], L[1], L[2], nx, ny, nz, dt, 1, nsteps, 0, prin=1.0e5, trin=Tu,
                  yrin=[1.0, 0.0])
    ctl.bc = [(1, [0, 0, 0, 0], [0.0] * 4)] * 6
    chem_path = os.path.join(workdir, f"onestep_{nx}.chem.dat")
    return Case(f"periodic_1step_{nx}x{ny}x{nz}", ctl, mech, tab, None,
                Dump(st, 0.0, dt), chem_path, None)


def h2_mechanism() -> _chem.Mechanism:
    return _chem.read_chem(os.path.join(DATA, "h2_9sp", "chem.dat"))


def _h2_mixtures(mech, phi):
    """Unburnt H2 + air (O2 + 3.76 N2) at equivalence ratio phi and its complete
    combustion products (H2O, O2 or H2 in excess, N2) as mass fractions."""
    W = mech.wmolar
    nO2 = 0.5 / phi
    mol_u = {"H2": 1.0, "O2": nO2, "N2": 3.76 * nO2}
    if phi <= 1.0:
        mol_b = {"H2O": 1.0, "O2": nO2 - 0.5, "N2": 3.76 * nO2}
    else:
        mol_b = {"H2O": 2.0 * nO2, "H2": 1.0 - 2.0 * nO2, "N2": 3.76 * nO2}
    idx = {s: i for i, s in enumerate(mech.species)}

    def massfrac(mol):
        y = np.zeros(mech.nspec)
        for sp, nm in mol.items():
            y[idx[sp]] = nm * W[idx[sp]]
        return y / y.sum()

    return massfrac(mol_u), massfrac(mol_b), idx


def h2_nscbc(n=(128, 128, 128), length=(5.0e-3, 5.0e-3, 5.0e-3), seed=7, phi=0.7,
             uin=1.25, uprime_over_sl=5.0, sl=1.2, mixavg=True, cfl=0.5, nsteps=100,
             periodic_x=False, radicals=True, Tu=300.0, falloff_variant=False) -> Case:
    """falloff_variant: the mechanism of h2_falloff_mechanism() (Troe and SRI
    steps) instead of the reference's H2-air mechanism."""
    nx, ny, nz = n
    chem_path = os.path.join(DATA, "h2_9sp", "chem.dat")
    if falloff_variant:
        gen = os.path.join(DATA, "_gen")
        os.makedirs(gen, exist_ok=True)
        chem_path = os.path.join(gen, "h2_falloff.chem.dat")
        mech = _fmt_roundtrip(h2_falloff_mechanism(), chem_path)
    else:
        mech = h2_mechanism()
    tab = _chem.derive(mech)
    dif = _diff.read_diff(os.path.join(DATA, "h2_9sp", "diff.dat"), mech.wmolar) if mixavg else None
    Yu, Yb, idx = _h2_mixtures(mech, phi)
    Tb = _ic.adiabatic_temperature(tab, Yu, Yb, Tu)
    d = tuple(length[i] / (n[i] - 1) for i in range(3))
    x = _ic.grid(nx, length[0])
    shape = (nz, ny, nx)
    delta = 0.35e-3
    if periodic_x:
        period = nx * d[0]
        c = 0.5 * (np.vectorize(math.erf)((x - 0.3 * period) / delta)
                   - np.vectorize(math.erf)((x - 0.7 * period) / delta))
    else:
        c = _ic.erf_profile(x, 0.5 * length[0], delta)
    C = np.broadcast_to(c, shape)
    T = Tu + (Tb - Tu) * C
    Y = [Yu[i] * (1.0 - C) + Yb[i] * C for i in range(mech.nspec)]
    if radicals:
        bump = np.exp(-((C - 0.5) / 0.2) ** 2)
        for s, amp in (("O", 2e-3), ("OH", 5e-3), ("H", 3e-4), ("HO2", 1e-4), ("H2O2", 2e-5)):
            Y[idx[s]] = Y[idx[s]] + amp * bump
        tot = sum(Y)
        # remove the radical mass from N2 so that sum(Y) = 1
        Y[idx["N2"]] = Y[idx["N2"]] - (tot - 1.0)
    up = uprime_over_sl * sl
    u, v, w = _ic.solenoidal_field(shape, (d[2], d[1], d[0]), up, 1.0e-3, seed,
                                   spectrum="passot-pouquet")
    if not periodic_x:
        # taper the fluctuation to zero at the x boundaries (inflow/outflow planes)
        taper = np.sin(np.pi * x / length[0]) ** 2
        u = u * taper; v = v * taper; w = w * taper
    u = u + uin
    st = _ic.conservative_state(tab, T, np.full(shape, 1.0e5), Y, u, v, w)
    dt = _f32(cfl * min(d) / sound_speed_max(tab, st, T))
    ctl = Control(length[0], length[1], length[2], nx, ny, nz, dt, 1, nsteps, 0, prin=1.0e5,
                  trin=Tu, urin=uin, yrin=[float(y) for y in Yu])
    per = (1, [0, 0, 0, 0], [0.0] * 4)
    if periodic_x:
        ctl.bc = [per] * 6
    else:
        ctl.bc = [(13, [1, 0, 0, 0], [uin, 0.0, 0.0, 0.0]),
                  (21, [0, 0, 0, 0], [1.0e5, 0.287, 1.0e-3, 0.0]), per, per, per, per]
    ctl.nfmavt = 1 if mixavg else 0
    name = f"h2_{'per' if periodic_x else 'nscbc'}_{'mix' if mixavg else 'cle'}_{nx}x{ny}x{nz}"
    if falloff_variant:
        name = "h2_falloff" + name[2:]
    return Case(name, ctl, mech, tab, dif, Dump(st, 0.0, dt), chem_path,
                os.path.join(DATA, "h2_9sp", "diff.dat") if mixavg else None,
                os.path.join(DATA, "h2_9sp", "radn.dat"))


def h2_kernel(n=(64, 64, 64), length=(4.0e-3, 4.0e-3, 4.0e-3), seed=11, phi=0.8, Tu=300.0,
              t_peak=1800.0, radius=1.0e-3, uprime_over_sl=4.0, sl=1.2, mixavg=True, cfl=0.5,
              nsteps=200) -> Case:
    """Flame kernel: a spherical Gaussian hot spot T = Tu + (T_peak - Tu) exp(-(r/radius)^2)
    at the box centre in an unburnt H2-air mixture, seeded Passot-Pouquet
    turbulence, triply periodic, 1 atm."""
    nx, ny, nz = n
    mech = h2_mechanism()
    tab = _chem.derive(mech)
    dif = _diff.read_diff(os.path.join(DATA, "h2_9sp", "diff.dat"), mech.wmolar) if mixavg else None
    Yu, _, _ = _h2_mixtures(mech, phi)
    d = tuple(length[i] / (n[i] - 1) for i in range(3))
    shape = (nz, ny, nx)
    X = [_ic.grid(n[i], length[i]) - 0.5 * n[i] * d[i] for i in range(3)]
    r2 = X[2][:, None, None] ** 2 + X[1][None, :, None] ** 2 + X[0][None, None, :] ** 2
    T = Tu + (t_peak - Tu) * np.exp(-r2 / radius ** 2)
    Y = [np.full(shape, Yu[i]) for i in range(mech.nspec)]
    u, v, w = _ic.solenoidal_field(shape, (d[2], d[1], d[0]), uprime_over_sl * sl, 1.0e-3, seed,
                                   spectrum="passot-pouquet")
    st = _ic.conservative_state(tab, T, np.full(shape, 1.0e5), Y, u, v, w)
    dt = _f32(cfl * min(d) / sound_speed_max(tab, st, T))
    ctl = Control(length[0], length[1], length[2], nx, ny, nz, dt, 1, nsteps, 0, prin=1.0e5,
                  trin=Tu, yrin=[float(y) for y in Yu])
    ctl.bc = [(1, [0, 0, 0, 0], [0.0] * 4)] * 6
    ctl.nfmavt = 1 if mixavg else 0
    return Case(f"h2_kernel_{nx}x{ny}x{nz}", ctl, mech, tab, dif, Dump(st, 0.0, dt),
                os.path.join(DATA, "h2_9sp", "chem.dat"),
                os.path.join(DATA, "h2_9sp", "diff.dat") if mixavg else None,
                os.path.join(DATA, "h2_9sp", "radn.dat"))


def h2_falloff_mechanism() -> _chem.Mechanism:
    """The H2-air mechanism with its two pressure-dependent steps moved to the
    other falloff forms CHRATE supports (src/chrate.f:312-436):

    * step 9, H + O2 (+M) = HO2 (+M), in Troe form with its Lindemann
      low-pressure Arrhenius data and Troe constants alpha = 0.7346,
      T*** = 94 K, T** = 5182 K, T* = 1756 K and the standard broadening
      constants c = -0.4 - 0.67 log Fc, n = 0.75 - 1.27 log Fc, d = 0.14
      (file layout of src/chemin.f:300-322; conversions src/chemin.f:843-850);
    * step 15, H2O2 (+M) = 2 OH (+M), in SRI form with a = 0.45, b = 797 K,
      c = 979 K, d = 1, e = 0.  Its low-pressure pre-exponential is raised by
      1e6 so that the reduced pressure stays above 10 over 300-2500 K: the
      reference's SRI exponent is 1/(1 + log10(log10 Pr)) (src/chrate.f:415),
      which has no real value for Pr < 1.

    Test data for the Troe and SRI branches, not a validated mechanism."""
    m = h2_mechanism()
    ist_troe, ist_sri = 8, 14                       # 0-based steps 9 and 15
    lt, ls = m.mllist[ist_troe] - 1, m.mllist[ist_sri] - 1
    a0t, n0t, e0t, _ = m.rclind[:, lt]
    a0s, n0s, e0s, _ = m.rclind[:, ls]
    m.rctroe = np.array([[a0t, n0t, e0t, 0.7346, 94.0, 5182.0, 1756.0, -0.4, -0.67, 0.75, -1.27, 0.14]]).T
    m.rcsrif = np.array([[a0s * 1.0e6, n0s, e0s, 0.45, 797.0, 979.0, 1.0, 0.0]]).T
    m.mllist = [0] * m.nstep
    m.rclind = np.zeros((4, 0))
    m.mtlist = [0] * m.nstep
    m.mtlist[ist_troe] = 1
    m.mslist = [0] * m.nstep
    m.mslist[ist_sri] = 1
    return m


def ch4_mechanism() -> _chem.Mechanism:
    return _chem.read_chem(os.path.join(DATA, "ch4_4step", "chem.dat"))


def ch4_flame(n=(24, 24, 24), length=None, seed=3, phi=0.8, Tu=300.0, uprime=1.0,
              cfl=0.5, nsteps=10, workdir=None) -> Case:
    """Triply periodic methane-air flame slab with the reference's 4-step
    mechanism (data/ch4_4step/chem.dat): two erf fronts along x between
    unburnt CH4 + (2/phi)(O2 + 3.76 N2) at Tu and its complete-combustion
    products at the adiabatic temperature, Gaussian bumps of H2, H and CO in
    the fronts, seeded Passot-Pouquet turbulence, constant-Lewis transport
    with the mechanism's Lewis numbers, 1 atm."""
    nx, ny, nz = n
    workdir = workdir or os.path.join(DATA, "_gen")
    os.makedirs(workdir, exist_ok=True)
    mech = ch4_mechanism()
    tab = _chem.derive(mech)
    idx = {s: i for i, s in enumerate(mech.species)}
    W = mech.wmolar
    nO2 = 2.0 / phi

    def massfrac(mol):
        y = np.zeros(mech.nspec)
        for sp, nm in mol.items():
            y[idx[sp]] = nm * W[idx[sp]]
        return y / y.sum()

    Yu = massfrac({"CH4": 1.0, "O2": nO2, "N2": 3.76 * nO2})
    Yb = massfrac({"CO2": 1.0, "H2O": 2.0, "O2": max(nO2 - 2.0, 0.0), "N2": 3.76 * nO2})
    Tb = _ic.adiabatic_temperature(tab, Yu, Yb, Tu)
    if length is None:
        length = tuple((m - 1) * H2_SPACING for m in n)
    d = tuple(length[i] / (n[i] - 1) for i in range(3))
    x = _ic.grid(nx, length[0])
    period = nx * d[0]
    delta = 3.0 * d[0]
    c = 0.5 * (np.vectorize(math.erf)((x - 0.3 * period) / delta)
               - np.vectorize(math.erf)((x - 0.7 * period) / delta))
    shape = (nz, ny, nx)
    C = np.broadcast_to(c, shape)
    T = Tu + (Tb - Tu) * C
    Y = [Yu[i] * (1.0 - C) + Yb[i] * C for i in range(mech.nspec)]
    bump = np.exp(-((C - 0.5) / 0.2) ** 2)
    for s, amp in (("H2", 2e-3), ("H", 1e-4), ("CO", 5e-3)):
        Y[idx[s]] = Y[idx[s]] + amp * bump
    Y[idx["N2"]] = Y[idx["N2"]] - (sum(Y) - 1.0)
    u, v, w = _ic.solenoidal_field(shape, (d[2], d[1], d[0]), uprime, 4.0 * delta, seed,
                                   spectrum="passot-pouquet")
    st = _ic.conservative_state(tab, T, np.full(shape, 1.0e5), Y, u, v, w)
    dt = _f32(cfl * min(d) / sound_speed_max(tab, st, T))
    ctl = Control(length[0], length[1], length[2], nx, ny, nz, dt, 1, nsteps, 0, prin=1.0e5,
                  trin=Tu, yrin=[float(y) for y in Yu])
    ctl.bc = [(1, [0, 0, 0, 0], [0.0] * 4)] * 6
    ctl.nfmavt = 0
    return Case(f"ch4_4step_{nx}x{ny}x{nz}", ctl, mech, tab, None, Dump(st, 0.0, dt),
                os.path.join(DATA, "ch4_4step", "chem.dat"), None)


def ch4_16sp_mechanism() -> _chem.Mechanism:
    return _chem.read_chem(os.path.join(DATA, "ch4_16sp", "chem.dat"))


def ch4_16sp_nscbc(n=(12, 11, 11), length=(1.5e-3, 1.4e-3, 1.4e-3), seed=5, phi=0.8, Tu=300.0,
                   uin=0.39, uprime=0.5, cfl=0.3, nsteps=10, periodic_x=False, transport=True,
                   radiation=True) -> Case:
    """Methane-air flame front with the reference's 16-species, 35-step
    mechanism (data/ch4_16sp/chem.dat, src/casefiles/25step: third-body and
    Lindemann steps), its mixture-averaged transport fits with thermal
    diffusion ratios and its radiation data (four radiating species): an erf
    front along x between unburnt CH4 + (2/phi)(O2 + 3.76 N2) at Tu and its
    complete-combustion products at the adiabatic temperature, Gaussian bumps
    of the intermediates in the front, seeded Passot-Pouquet turbulence, 1 atm.
    An NSCBC duct (inflow at x-left with the unburnt state, outflow at
    x-right), or triply periodic with two fronts.  transport: the four
    molecular-transport options of the reference's own cont.dat for this case
    (molar-mass and pressure diffusion, Soret, Dufour, WMLTDR = 5); radiation:
    NFRADN = 1."""
    nx, ny, nz = n
    mech = ch4_16sp_mechanism()
    tab = _chem.derive(mech)
    data = os.path.join(DATA, "ch4_16sp")
    dif = _diff.read_diff(os.path.join(data, "diff.dat"), mech.wmolar)
    idx = {s: i for i, s in enumerate(mech.species)}
    W = mech.wmolar
    nO2 = 2.0 / phi

    def massfrac(mol):
        y = np.zeros(mech.nspec)
        for sp, nm in mol.items():
            y[idx[sp]] = nm * W[idx[sp]]
        return y / y.sum()

    Yu = massfrac({"CH4": 1.0, "O2": nO2, "N2": 3.76 * nO2})
    Yb = massfrac({"CO2": 1.0, "H2O": 2.0, "O2": max(nO2 - 2.0, 0.0), "N2": 3.76 * nO2})
    Tb = _ic.adiabatic_temperature(tab, Yu, Yb, Tu)
    d = tuple(length[i] / (n[i] - 1) for i in range(3))
    x = _ic.grid(nx, length[0])
    shape = (nz, ny, nx)
    delta = 0.3e-3
    if periodic_x:
        period = nx * d[0]
        c = 0.5 * (np.vectorize(math.erf)((x - 0.3 * period) / delta)
                   - np.vectorize(math.erf)((x - 0.7 * period) / delta))
    else:
        c = _ic.erf_profile(x, 0.5 * length[0], delta)
    C = np.broadcast_to(c, shape)
    T = Tu + (Tb - Tu) * C
    Y = [Yu[i] * (1.0 - C) + Yb[i] * C for i in range(mech.nspec)]
    bump = np.exp(-((C - 0.5) / 0.2) ** 2)
    for sp, amp in (("H2", 1e-3), ("O", 1e-3), ("OH", 3e-3), ("H", 1e-4), ("HO2", 1e-4), ("H2O2", 2e-5),
                    ("CO", 5e-3), ("CH2O", 1e-4), ("HCO", 1e-5), ("CH3", 5e-4), ("CH3O", 1e-5)):
        Y[idx[sp]] = Y[idx[sp]] + amp * bump
    Y[idx["N2"]] = Y[idx["N2"]] - (sum(Y) - 1.0)
    u, v, w = _ic.solenoidal_field(shape, (d[2], d[1], d[0]), uprime, 4.0 * delta, seed,
                                   spectrum="passot-pouquet")
    if not periodic_x:
        taper = np.sin(np.pi * x / length[0]) ** 2
        u = u * taper; v = v * taper; w = w * taper
        u = u + uin
    st = _ic.conservative_state(tab, T, np.full(shape, 1.0e5), Y, u, v, w)
    dt = _f32(cfl * min(d) / sound_speed_max(tab, st, T))
    ctl = Control(length[0], length[1], length[2], nx, ny, nz, dt, 1, nsteps, 0, prin=1.0e5,
                  trin=Tu, urin=uin if not periodic_x else 0.0, yrin=[float(y) for y in Yu])
    per = (1, [0, 0, 0, 0], [0.0] * 4)
    if periodic_x:
        ctl.bc = [per] * 6
    else:
        ctl.bc = [(13, [1, 0, 0, 0], [uin, 0.0, 0.0, 0.0]),
                  (21, [0, 0, 0, 0], [1.0e5, 0.287, 1.0e-3, 0.0]), per, per, per, per]
    ctl.nfmavt = 1
    if transport:
        ctl.nfmixw = ctl.nfmixp = ctl.nfmsor = ctl.nfmduf = 1
    ctl.wmltdr = 5.0
    ctl.nfradn = 1 if radiation else 0
    name = f"ch4_16sp_{'per' if periodic_x else 'nscbc'}_{nx}x{ny}x{nz}"
    return Case(name, ctl, mech, tab, dif, Dump(st, 0.0, dt), os.path.join(data, "chem.dat"),
                os.path.join(data, "diff.dat"), os.path.join(data, "radn.dat"))


# ---------------------------------------------------------------------------
# Case sets for batched and served runs.  Each spec is a small dict drawn from
# a seeded generator; make_case(spec) builds the initial dump on demand.
H2_SPACING = 5.0e-3 / 127   # grid spacing of the 128^3, 5 mm box


def batch_specs(count=64, seed=2024, sizes=(48, 64, 96, 128)):
    """Uneven H2-air inflow/outflow flames: n^3 with n drawn from sizes,
    phi ~ U(0.5, 1.2), u'/S_L ~ U(1, 10), T_u ~ U(300, 700) K, each with its
    own turbulence seed; the box is (n - 1) grid spacings of the 128^3 case."""
    rng = np.random.default_rng(seed)
    out = []
    for q in range(count):
        n = int(rng.choice(sizes))
        out.append(dict(kind="nscbc", n=(n, n, n), phi=float(rng.uniform(0.5, 1.2)),
                        uprime_over_sl=float(rng.uniform(1.0, 10.0)), Tu=float(rng.uniform(300.0, 700.0)),
                        seed=int(rng.integers(1, 2 ** 31 - 1))))
    return out


def kernel_specs(count=32, seed=4242, rate=4.0, sizes=(48, 64, 80, 96, 112, 128, 144, 160), nsteps=200):
    """Flame-kernel requests: Poisson arrivals of mean rate per second (arrival
    times in seconds from 0), n^3 with n drawn from sizes, T_peak ~ U(1500,
    2000) K, radius ~ U(0.5, 2) mm, u'/S_L ~ U(1, 10), phi ~ U(0.5, 1.2)."""
    rng = np.random.default_rng(seed)
    t = 0.0
    out = []
    for q in range(count):
        t += float(rng.exponential(1.0 / rate))
        n = int(rng.choice(sizes))
        out.append(dict(kind="kernel", n=(n, n, n), arrival=t, t_peak=float(rng.uniform(1500.0, 2000.0)),
                        radius=float(rng.uniform(0.5e-3, 2.0e-3)), phi=float(rng.uniform(0.5, 1.2)),
                        uprime_over_sl=float(rng.uniform(1.0, 10.0)), seed=int(rng.integers(1, 2 ** 31 - 1)),
                        nsteps=nsteps))
    return out


def make_case(spec) -> Case:
    n = tuple(spec["n"])
    length = tuple((m - 1) * H2_SPACING for m in n)
    if spec["kind"] == "nscbc":
        return h2_nscbc(n=n, length=length, seed=spec["seed"], phi=spec["phi"],
                        uprime_over_sl=spec["uprime_over_sl"], Tu=spec["Tu"], nsteps=spec.get("nsteps", 100))
    if spec["kind"] == "kernel":
        return h2_kernel(n=n, length=length, seed=spec["seed"], phi=spec["phi"], t_peak=spec["t_peak"],
                         radius=spec["radius"], uprime_over_sl=spec["uprime_over_sl"],
                         nsteps=spec.get("nsteps", 200))
    raise ValueError(spec["kind"])
