"""The C-ABI library on the CPU: it loads, exports what include/senga_gpu.h
declares, agrees with the Python struct layout, and rejects unsupported
configurations before touching a device.  Host-side readers round-trip."""
import ctypes as C
import os
import re

import numpy as np
import pytest

from sengapy import abi, chem, cont
from sengapy import case as CS

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def lib():
    if not os.path.exists(abi.LIBPATH):
        import build
        build.build_library()
    return abi.load_library()


def declared_functions():
    text = open(os.path.join(ROOT, "include", "senga_gpu.h")).read()
    return sorted(set(re.findall(r"^\s*(?:int|size_t)\s+(sg_\w+)\s*\(", text, re.M)))


def test_exports_every_declared_symbol(lib):
    names = declared_functions()
    assert len(names) >= 18
    for n in names:
        assert hasattr(lib, n), n
    assert set(names) == set(abi.ABI_FUNCS)


def test_abi_version_and_layout(lib):
    assert lib.sg_abi_version() == 1
    assert lib.sg_config_size() == C.sizeof(abi.Config)


def test_rejects_unsupported_configs(lib):
    c = CS.periodic_1step(n=(12, 11, 11))
    cfg = abi.config_from_case(c)
    h = C.c_void_p()
    bad = abi.Config.from_buffer_copy(cfg)
    bad.nspec = 1
    assert lib.sg_create(C.byref(bad), C.byref(h)) == 1
    bad = abi.Config.from_buffer_copy(cfg)
    bad.bc_type[2] = 21                     # non-periodic y
    assert lib.sg_create(C.byref(bad), C.byref(h)) == 1
    bad = abi.Config.from_buffer_copy(cfg)
    bad.nx = 8                               # shorter than the 11-point stencil
    assert lib.sg_create(C.byref(bad), C.byref(h)) == 1
    bad = abi.Config.from_buffer_copy(cfg)
    bad.precision = 3
    assert lib.sg_create(C.byref(bad), C.byref(h)) == 1
    assert lib.sg_create(None, C.byref(h)) == 1


def test_chem_dat_round_trip(tmp_path):
    m = CS.h2_mechanism()
    p = tmp_path / "chem.dat"
    chem.write_chem(m, str(p))
    m2 = chem.read_chem(str(p))
    assert m2.species == m.species
    assert np.array_equal(m2.amolcp, m.amolcp)
    assert np.array_equal(m2.rparam, m.rparam)
    assert m2.nsspec == m.nsspec and m2.diffmu == m.diffmu and m2.mllist == m.mllist
    assert np.array_equal(m2.rclind, m.rclind) and np.array_equal(m2.effy3b, m.effy3b)


def test_cont_dat_round_trip(tmp_path):
    c = CS.h2_nscbc(n=(12, 11, 11), length=(1e-3, 1e-3, 1e-3))
    p = tmp_path / "cont.dat"
    c.control.write(str(p))
    r = cont.read_cont(str(p))
    assert (r.nx, r.ny, r.nz) == (12, 11, 11)
    assert r.tstep == c.control.tstep
    assert r.bc == c.control.bc
    assert r.yrin == c.control.yrin
    assert r.nfmavt == 1


def test_derived_tables_match_reference_recipe():
    """chemin.f:896-961 identities: AMASCT(1) = AMASCP(1) - R_k, AMASCT(m) =
    AMASCP(m)/m, enthalpy polynomial derivative equals cp."""
    t = chem.derive(CS.h2_mechanism())
    for s in range(t.nspec):
        for it in range(2):
            assert t.amasct[0, it, s] == t.amascp[0, it, s] - t.rgspec[s]
            T = 1500.0
            h = lambda TT: t.amasch[5, it, s] + TT * sum(t.amasch[m, it, s] * TT ** m for m in range(5))
            cp = sum(t.amascp[m, it, s] * T ** m for m in range(5))
            dh = (h(T + 1e-3) - h(T - 1e-3)) / 2e-3
            assert abs(dh - cp) < 1e-6 * cp
