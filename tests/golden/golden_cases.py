"""Seeded cases behind the golden fixtures (shared by make_golden.py and tests)."""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "senga-versiona_amd"))

from sengapy import case as C  # noqa: E402


def build(name):
    if name == "periodic_1step_16x12x11":
        return C.periodic_1step(n=(16, 12, 11))
    if name == "h2_nscbc_mix_12x11x11":
        return C.h2_nscbc(n=(12, 11, 11), length=(1.5e-3, 1.4e-3, 1.4e-3), mixavg=True)
    if name == "h2_nscbc_cle_12x11x11":
        return C.h2_nscbc(n=(12, 11, 11), length=(1.5e-3, 1.4e-3, 1.4e-3), mixavg=False)
    if name == "h2_per_mix_11x12x11":
        return C.h2_nscbc(n=(11, 12, 11), length=(1.4e-3, 1.5e-3, 1.4e-3), mixavg=True, periodic_x=True)
    if name == "h2_falloff_nscbc_mix_12x11x11":
        # CHRATE's Troe and SRI branches (src/chrate.f:312-436)
        return C.h2_nscbc(n=(12, 11, 11), length=(1.5e-3, 1.4e-3, 1.4e-3), mixavg=True, falloff_variant=True)
    if name == "ch4_4step_16x12x11":
        # the reference's 4-step methane mechanism: reactant coefficient lists
        # with non-integer and negative orders (src/chrate.f:475-516)
        return C.ch4_flame(n=(16, 12, 11))
    if name == "h2_nscbc_mixwp_12x11x11":
        # molar-mass and pressure diffusion (src/rhscal.f:945-992,1970-2759)
        c = C.h2_nscbc(n=(12, 11, 11), length=(1.5e-3, 1.4e-3, 1.4e-3), mixavg=True)
        c.control.nfmixw = c.control.nfmixp = 1
        return c
    if name == "h2_per_mixwp_11x12x11":
        c = C.h2_nscbc(n=(11, 12, 11), length=(1.4e-3, 1.5e-3, 1.4e-3), mixavg=True, periodic_x=True)
        c.control.nfmixw = c.control.nfmixp = 1
        return c
    if name == "h2_nscbc_sordf_12x11x11":
        # Soret and Dufour for the light species H2, H (src/rhscal.f:1353-1410,
        # 1549-1566, 2757-3758)
        c = C.h2_nscbc(n=(12, 11, 11), length=(1.5e-3, 1.4e-3, 1.4e-3), mixavg=True)
        c.control.nfmsor = c.control.nfmduf = 1
        return c
    if name == "h2_per_duf_11x12x11":
        c = C.h2_nscbc(n=(11, 12, 11), length=(1.4e-3, 1.5e-3, 1.4e-3), mixavg=True, periodic_x=True)
        c.control.nfmduf = 1
        return c
    if name == "h2_nscbc_alltr_12x11x11":
        # all four molecular-transport options together
        c = C.h2_nscbc(n=(12, 11, 11), length=(1.5e-3, 1.4e-3, 1.4e-3), mixavg=True)
        c.control.nfmixw = c.control.nfmixp = c.control.nfmsor = c.control.nfmduf = 1
        return c
    if name == "h2_nscbc_radn_12x11x11":
        # optically thin radiation from H2O (src/radcal.f:36-111)
        c = C.h2_nscbc(n=(12, 11, 11), length=(1.5e-3, 1.4e-3, 1.4e-3), mixavg=True)
        c.control.nfradn = 1
        return c
    if name == "h2_per_radn_cle_11x12x11":
        # radiation with constant-Lewis transport, periodic x
        c = C.h2_nscbc(n=(11, 12, 11), length=(1.4e-3, 1.5e-3, 1.4e-3), mixavg=False, periodic_x=True)
        c.control.nfradn = 1
        return c
    if name == "ch4_16sp_nscbc_12x11x11":
        # the reference's 16-species, 35-step methane mechanism (src/casefiles/25step)
        # with every option of that case's cont.dat: mixture-averaged transport,
        # molar-mass and pressure diffusion, Soret, Dufour, and radiation from
        # four species; NSCBC duct
        return C.ch4_16sp_nscbc(n=(12, 11, 11), length=(1.5e-3, 1.4e-3, 1.4e-3))
    if name == "ch4_16sp_per_11x12x11":
        return C.ch4_16sp_nscbc(n=(11, 12, 11), length=(1.4e-3, 1.5e-3, 1.4e-3), periodic_x=True)
    if name == ADAPT:
        c = C.h2_nscbc(n=(12, 11, 11), length=(1.5e-3, 1.4e-3, 1.4e-3), mixavg=True)
        c.control.nstpsw = 1
        return c
    raise KeyError(name)


ADAPT = "h2_nscbc_mix_adapt_12x11x11"


NAMES = ["periodic_1step_16x12x11", "h2_nscbc_mix_12x11x11", "h2_nscbc_cle_12x11x11", "h2_per_mix_11x12x11",
         "h2_falloff_nscbc_mix_12x11x11", "ch4_4step_16x12x11", "h2_nscbc_mixwp_12x11x11", "h2_per_mixwp_11x12x11",
         "h2_nscbc_sordf_12x11x11", "h2_per_duf_11x12x11", "h2_nscbc_alltr_12x11x11",
         "h2_nscbc_radn_12x11x11", "h2_per_radn_cle_11x12x11", "ch4_16sp_nscbc_12x11x11",
         "ch4_16sp_per_11x12x11"]
OPS = {"ops_periodic_16x12x11": "periodic_1step_16x12x11", "ops_boundedx_12x11x11": "h2_nscbc_cle_12x11x11"}


def load(name):
    """Case with the fixture's exact initial state and time step, plus the fixture."""
    fx = dict(np.load(os.path.join(HERE, name + ".npz")))
    c = build(name)
    c.dump.state = fx["state0"]
    c.dump.tstep = float(fx["tstep"])
    c.control.tstep = c.dump.tstep
    return c, fx
