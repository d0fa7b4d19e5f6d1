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
    assert lib.sg_abi_version() == 4
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


def test_radiation_switch_and_data_checks(lib, tmp_path):
    """Radiation (cont.dat NFRADN, src/contin.f:194-209; radn.dat,
    src/radcin.f:76-126) is computed: the switch reaches sg_config through the
    reader with the H2O absorption polynomial, sg_footprint accepts it, and
    the radiation data is range-checked (species index, coefficient count,
    radiating-species count) before it goes to the device."""
    c = CS.h2_nscbc(n=(12, 11, 11), length=(1e-3, 1e-3, 1e-3), mixavg=True)
    c.control.nfradn = 1
    p = tmp_path / "cont.dat"
    c.control.write(str(p))
    assert cont.read_cont(str(p)).nfradn == 1
    cfg = abi.config_from_case(c)
    assert cfg.physics == abi.PHYS_RADN
    assert cfg.nsprad == 1 and cfg.nsprid[0] == 2 and cfg.nkprad[0] == 6
    assert cfg.akprad[0][0] == 6.869480e-04 and cfg.akprad[0][5] == -1.373456e-20
    assert cfg.trefrn == c.control.trefrn
    out = C.c_size_t(0)
    assert lib.sg_footprint(C.byref(cfg), C.byref(out)) == 0
    for field, value in (("nsprid", 9), ("nsprid", -1), ("nkprad", 0), ("nkprad", abi.NKPRAD + 1)):
        bad = abi.Config.from_buffer_copy(cfg)
        getattr(bad, field)[0] = value
        assert lib.sg_footprint(C.byref(bad), C.byref(out)) == 1, (field, value)
    bad = abi.Config.from_buffer_copy(cfg)
    bad.nsprad = 10
    assert lib.sg_footprint(C.byref(bad), C.byref(out)) == 1
    c.radn_path = None
    with pytest.raises(ValueError):
        abi.config_from_case(c)


def test_transport_options_need_mixture_averaging():
    """The four molecular-transport options act only with mixture-averaged
    transport on (src/diffin.f:66,384-407); radiation acts either way."""
    c = CS.h2_nscbc(n=(12, 11, 11), length=(1e-3, 1e-3, 1e-3), mixavg=False)
    c.control.nfmsor = c.control.nfmduf = c.control.nfmixw = c.control.nfmixp = 1
    assert abi.physics_bits(c.control) == 0
    c.control.nfradn = 1
    assert abi.physics_bits(c.control) == abi.PHYS_RADN


@pytest.mark.parametrize("field,value", [
    ("nsslen", 11), ("nrslen", -1), ("npslen", 11), ("nrclen", 11), ("npclen", 11),
    ("nsspec", 9), ("nrspec", -1), ("npspec", 9), ("nrcpec", 9), ("npcpec", 9),
    ("mblist", 17), ("mllist", 33), ("mtlist", -1), ("mslist", 33)])
def test_rejects_out_of_range_mechanism_tables(lib, field, value):
    """Every list length and index the kernels dereference is checked before
    the tables go to the device: a malformed configuration gets SG_EINVAL."""
    c = CS.h2_nscbc(n=(12, 11, 11), length=(1e-3, 1e-3, 1e-3), mixavg=True)
    cfg = abi.config_from_case(c)
    out = C.c_size_t(0)
    assert lib.sg_footprint(C.byref(cfg), C.byref(out)) == 0
    st = 4
    arr = getattr(cfg, field)
    if field.endswith("len") or field.endswith("list"):
        arr[st] = value
    else:
        lenf = {"nsspec": "nsslen", "nrspec": "nrslen", "npspec": "npslen", "nrcpec": "nrclen",
                "npcpec": "npclen"}[field]
        getattr(cfg, lenf)[st] = max(getattr(cfg, lenf)[st], 1)
        arr[st][0] = value
    assert lib.sg_footprint(C.byref(cfg), C.byref(out)) == 1
    h = C.c_void_p()
    assert lib.sg_create(C.byref(cfg), C.byref(h)) == 1


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


def test_footprint_without_device(lib):
    """sg_footprint sizes a case on the host: registers (2(5+N) FP64 fields)
    plus working fields (FP64, or FP32 in the mixed form), and rejects what
    sg_create rejects."""
    c = CS.periodic_1step(n=(16, 12, 11))
    ns, npts = c.nspec, 16 * 12 * 11
    b64 = abi.footprint(c, 0, lib)
    b32 = abi.footprint(c, 1, lib)
    work = (9 + 3 * ns) + ns + (5 + ns) + 3       # primitives, rates, accumulators, viscosity gradient
    assert b64 >= npts * (8 * 2 * (5 + ns) + 8 * work)
    assert b32 >= npts * (8 * 2 * (5 + ns) + 4 * work)
    assert b64 - b32 >= npts * 4 * work
    big = CS.periodic_1step(n=(32, 24, 22))        # 8x the points: the per-point part grows 8x
    assert abi.footprint(big, 0, lib) - b64 >= 7 * npts * (8 * 2 * (5 + ns) + 8 * work)
    cfg = abi.config_from_case(c)
    out = C.c_size_t(0)
    bad = abi.Config.from_buffer_copy(cfg)
    bad.nx = 8
    assert lib.sg_footprint(C.byref(bad), C.byref(out)) == 1
    bad = abi.Config.from_buffer_copy(cfg)
    bad.nspec = 1
    assert lib.sg_footprint(C.byref(bad), C.byref(out)) == 1


def test_batch_entry_points_reject_null(lib):
    h = C.c_void_p()
    n = C.c_int(0)
    assert lib.sg_batch_create(None) == 1
    assert lib.sg_batch_destroy(None) == 1
    assert lib.sg_batch_add(None, None) == 1
    assert lib.sg_batch_remove(None, None) == 1
    assert lib.sg_batch_size(None, C.byref(n)) == 1
    assert lib.sg_batch_member(None, 0, C.byref(h)) == 1
    assert lib.sg_batch_step(None, 1) == 1
    assert lib.sg_batch_sync(None, None) == 1


def test_case_sets_follow_their_distributions():
    b = CS.batch_specs(200, seed=3)
    assert {s["n"][0] for s in b} == {48, 64, 96, 128}
    assert all(0.5 <= s["phi"] <= 1.2 and 1.0 <= s["uprime_over_sl"] <= 10.0 and 300.0 <= s["Tu"] <= 700.0
               for s in b)
    assert CS.batch_specs(5, seed=3) == b[:5]          # seeded
    k = CS.kernel_specs(4000, seed=5, rate=4.0)
    t = np.array([s["arrival"] for s in k])
    assert np.all(np.diff(t) > 0)
    assert abs(len(k) / t[-1] - 4.0) < 0.3              # Poisson rate (standard error 0.06)
    assert all(1500.0 <= s["t_peak"] <= 2000.0 and 0.5e-3 <= s["radius"] <= 2e-3 for s in k)
    assert all(48 <= s["n"][0] <= 160 for s in k)


def test_serving_summary():
    from sengapy.serve import Result, summary
    rs = [Result(q, arrival=0.5 * q, admitted=0.5 * q, finished=0.5 * q + 1.0 + 0.1 * q, nsteps=10,
                 points=1000) for q in range(10)]
    s = summary(rs)
    lat = np.array([1.0 + 0.1 * q for q in range(10)])
    assert s["cases"] == 10
    assert abs(s["makespan_s"] - (4.5 + 1.9)) < 1e-12
    assert abs(s["p50_s"] - np.percentile(lat, 50)) < 1e-12
    assert abs(s["p99_s"] - np.percentile(lat, 99)) < 1e-12
    assert abs(s["mpoint_substages_per_s"] - 10 * 1000 * 50 / 6.4 / 1e6) < 1e-9


@pytest.mark.parametrize("switch,bit", [("nfmixw", abi.PHYS_MIXW), ("nfmixp", abi.PHYS_MIXP),
                                        ("nfmsor", abi.PHYS_SORET), ("nfmduf", abi.PHYS_DUFOUR)])
def test_accepts_molecular_transport_options(lib, switch, bit, tmp_path):
    """The four molecular-transport options are computed: the switch reaches
    sg_config through the reader, and sg_footprint accepts it.  Soret and
    Dufour need the thermal diffusion ratio fits."""
    c = CS.h2_nscbc(n=(12, 11, 11), length=(1e-3, 1e-3, 1e-3), mixavg=True)
    setattr(c.control, switch, 1)
    p = tmp_path / "cont.dat"
    c.control.write(str(p))
    assert getattr(cont.read_cont(str(p)), switch) == 1
    cfg = abi.config_from_case(c)
    assert cfg.physics == bit
    out = C.c_size_t(0)
    assert lib.sg_footprint(C.byref(cfg), C.byref(out)) == 0
    if bit in (abi.PHYS_SORET, abi.PHYS_DUFOUR):
        assert cfg.ncotdr == 4
        cfg.physics = bit
        cfg.ncotdr = 0
        assert lib.sg_footprint(C.byref(cfg), C.byref(out)) == 1
