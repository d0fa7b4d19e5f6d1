"""The Fortran driver (senga-versiona_amd/fortran/senga_gpu_driver.f90).

CPU: reading input/cont.dat, input/chem.dat and input/diff.dat with the
Fortran readers builds an sg_config byte-identical to the one the Python
readers build (sengapy.abi.config_from_case) -- both follow
src/contin.f, src/chemin.f:155-344,822-983 and src/diffin.f.

GPU: the driver's time loop (src/senga2.f:96-194 shape: five substages and the
end-of-step call per step, dumps every NTDUMP steps) writes the same dump as
the library driven from Python, and agrees with the oracle."""
import os
import subprocess
import sys
import tempfile

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
PKG = os.path.join(ROOT, "senga-versiona_amd")
EXE = os.path.join(PKG, "lib", "senga_gpu")
FLANG = "/opt/rocm/lib/llvm/bin/flang"


def _driver():
    if not os.path.exists(EXE):
        if not os.path.exists(FLANG):
            pytest.skip("Fortran driver not built and flang not present")
        sys.path.insert(0, PKG)
        import build
        assert build.build_fortran_driver() == EXE
    return EXE


def _cases():
    from sengapy import case as C
    return [("h2_mix", C.h2_nscbc(n=(16, 12, 12), length=(2e-3, 1.5e-3, 1.5e-3), mixavg=True, nsteps=2)),
            ("h2_cle", C.h2_nscbc(n=(16, 12, 12), length=(2e-3, 1.5e-3, 1.5e-3), mixavg=False, nsteps=2)),
            ("periodic", C.periodic_1step(n=(16, 12, 11), nsteps=2))]


def test_fortran_config_equals_python_config(tmp_path):
    from sengapy import abi
    exe = _driver()
    for name, c in _cases():
        d = tmp_path / name
        c.write_inputs(str(d))
        out = d / "cfg.bin"
        r = subprocess.run([exe, "--config", str(d), str(out)], capture_output=True, text=True)
        assert r.returncode == 0, r.stdout + r.stderr
        assert out.read_bytes() == bytes(abi.config_from_case(c)), name


@pytest.mark.gpu
def test_fortran_driver_run_matches_library_and_oracle(tmp_path):
    from sengapy import case as C
    from sengapy.abi import Solver
    from sengapy.dump import read_dump
    from pyoracle import Oracle
    from conftest import rel_err
    exe = _driver()
    steps = 3
    c = C.h2_nscbc(n=(24, 16, 16), length=(3e-3, 2e-3, 2e-3), mixavg=True, nsteps=steps)
    c.control.ntdump = steps
    c.write_inputs(str(tmp_path))
    print("fortran driver", exe, flush=True)
    r = subprocess.run([exe, str(tmp_path)], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr, flush=True)
    assert r.returncode == 0
    # the driver's configuration is the binding's byte for byte, so the H2
    # mechanism it reads from chem.dat is the one compiled into the library
    assert "compiled mechanism" in r.stdout
    d = read_dump(str(tmp_path / "output" / "dmpi0000000.dat"))
    s = Solver(c)
    assert s.mechanism_compiled()
    s.set_timestep(c.dump.tstep)
    s.upload(c.dump.state, 0.0)
    s.step(steps)
    run_g, _, _, _ = s.download()
    assert np.array_equal(d.state, run_g)
    assert d.etime == pytest.approx(steps * c.dump.tstep, rel=1e-15)
    o = Oracle(c)
    o.set_timestep(c.dump.tstep)
    o.set_state(c.dump.state, 0.0)
    o.step(steps)
    run_o, _, _, _ = o.get()
    errs = [rel_err(d.state[v], run_o[v]) for v in range(run_o.shape[0])]
    print("driver vs oracle", " ".join(f"{e:.1e}" for e in errs), flush=True)
    assert max(errs) < 1e-10


@pytest.mark.gpu
def test_fortran_driver_adaptive_restart_keeps_dump_step(tmp_path):
    """With the adaptive step (NSTPSW = 1) a restart takes TSTEP from the dump,
    not from cont.dat (src/indata.f:1423,1481): cont.dat here holds a
    different step, and the run must follow the oracle started from the
    dump's step."""
    import dataclasses
    from sengapy import case as C
    from sengapy.dump import read_dump
    from pyoracle import Oracle
    from conftest import rel_err
    exe = _driver()
    steps = 2
    c = C.h2_nscbc(n=(24, 16, 16), length=(3e-3, 2e-3, 2e-3), mixavg=True, nsteps=steps)
    c.control.nstpsw = 1
    c.control.ntdump = steps
    c.control.tstep = 1.37 * c.dump.tstep      # must not be used
    c.dump = dataclasses.replace(c.dump, tstep=0.9 * c.dump.tstep, errold=1.0e-4, errldr=1.0e-4)
    c.write_inputs(str(tmp_path))
    print("fortran driver adaptive restart", exe, flush=True)
    r = subprocess.run([exe, str(tmp_path)], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr, flush=True)
    assert r.returncode == 0
    d = read_dump(str(tmp_path / "output" / "dmpi0000000.dat"))
    o = Oracle(c)
    o.set_timestep(c.dump.tstep)
    o.set_adaptive(True, c.dump.errold, c.dump.errldr)
    o.set_state(c.dump.state, 0.0)
    o.step(steps)
    run_o, _, _, _ = o.get()
    dt_o = o.timestep()[0]
    print("adapted step", d.tstep, "oracle", dt_o, flush=True)
    # tolerances of the adaptive-step parity test (the error norm amplifies
    # last-bit RHS differences to ~1e-8 relative in the step)
    assert abs(d.tstep - dt_o) <= 1e-7 * dt_o
    errs = [rel_err(d.state[v], run_o[v]) for v in range(run_o.shape[0])]
    assert max(errs) < 1e-9


@pytest.mark.parametrize("what", ["species", "body", "lindemann"])
def test_fortran_chem_reader_rejects_out_of_range_indices(tmp_path, what):
    """A chem.dat whose step lists name a species, third body or falloff
    entry that does not exist is refused by the Fortran reader (it would
    otherwise write outside sg_config's tables)."""
    import copy
    from sengapy import case as C
    from sengapy import chem
    exe = _driver()
    c = C.h2_nscbc(n=(16, 12, 12), length=(2e-3, 1.5e-3, 1.5e-3), mixavg=True)
    c.write_inputs(str(tmp_path))
    m = copy.deepcopy(c.mech)
    if what == "species":
        m.nsspec[0][0] = len(m.species) + 1
    elif what == "body":
        st = next(i for i, b in enumerate(m.mblist) if b)
        m.mblist[st] = len(m.bodies) + 1
    else:
        st = next(i for i, b in enumerate(m.mllist) if b)
        m.mllist[st] = int(max(m.mllist)) + 1
    chem.write_chem(m, str(tmp_path / "input" / "chem.dat"))
    r = subprocess.run([exe, "--config", str(tmp_path), str(tmp_path / "cfg.bin")], capture_output=True, text=True)
    assert r.returncode != 0
    assert "cannot read input/chem.dat" in r.stdout
