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
    alltr = C.h2_nscbc(n=(16, 12, 12), length=(2e-3, 1.5e-3, 1.5e-3), mixavg=True, nsteps=2)
    alltr.control.nfmixw = alltr.control.nfmixp = alltr.control.nfmsor = alltr.control.nfmduf = 1
    alltr.control.nfradn = 1
    return [("h2_mix", C.h2_nscbc(n=(16, 12, 12), length=(2e-3, 1.5e-3, 1.5e-3), mixavg=True, nsteps=2)),
            ("h2_cle", C.h2_nscbc(n=(16, 12, 12), length=(2e-3, 1.5e-3, 1.5e-3), mixavg=False, nsteps=2)),
            ("periodic", C.periodic_1step(n=(16, 12, 11), nsteps=2)),
            # the molecular-transport switches and the thermal diffusion ratios
            ("h2_alltr", alltr),
            ("ch4_16sp", C.ch4_16sp_nscbc(n=(16, 12, 12), length=(2e-3, 1.5e-3, 1.5e-3), nsteps=2))]


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
    d = read_dump(str(tmp_path / "output" / "dmpi0000001.dat"))
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
    d = read_dump(str(tmp_path / "output" / "dmpi0000001.dat"))
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


@pytest.mark.parametrize("fname,nspec,etime,tstep", [
    ("ref21step_dmpi0000000.dat", 9, 1.394146750614353e-2, 1.413965038003083e-8),
    ("ref25step_dmpi0000000.dat", 16, 8.576785249268155e-3, 1.8955182896609245e-8)])
def test_both_readers_read_the_reference_formatted_dump(tmp_path, fname, nspec, etime, tstep):
    """Dumps the reference itself wrote: src/casefiles/21step/dmpi0000000.dat
    and src/casefiles/25step/dmpi0000000.dat (formatted, 1001 x 1 x 1 points,
    9 and 16 species; kept as the fixtures tests/golden/ref21step_... and
    ref25step_dmpi0000000.dat).  The Python formatted reader and the Fortran
    reader (the driver's --convert-dump mode: formatted read as
    src/indata.f:1449-1473, unformatted write as src/output.f:306-316) give the
    same values bit for bit, and each state is a 1-D flame: sum_k rho Y_k =
    rho to the dump's precision, positive density."""
    from sengapy.dump import read_dump, read_dump_formatted
    exe = _driver()
    src = os.path.join(HERE, "golden", fname)
    d = read_dump_formatted(src)
    assert d.state.shape == (5 + nspec, 1, 1, 1001)
    out = tmp_path / "dmpi.bin"
    r = subprocess.run([exe, "--convert-dump", src, str(out)], capture_output=True, text=True)
    assert r.returncode == 0, r.stdout + r.stderr
    u = read_dump(str(out))
    assert np.array_equal(u.state, d.state)
    assert (u.etime, u.tstep, u.errold, u.errldr) == (d.etime, d.tstep, d.errold, d.errldr)
    rho = d.state[0]
    assert rho.min() > 0
    assert np.max(np.abs(d.state[5:].sum(axis=0) / rho - 1.0)) < 1e-11
    assert d.etime == pytest.approx(etime) and d.tstep == pytest.approx(tstep)


@pytest.mark.gpu
def test_fortran_driver_outputs_match_reference(tmp_path):
    """The run outputs of src/output.f against the reference's own, written by
    the reference build with OUTPUT called after every step (fixture
    output_h2_nscbc_mix_12x11x11.npz: 4 steps, reports every step, dumps and
    statistics every 2 steps): the report lines OUTPUT appends to rept.res
    (step number, elapsed and next time step in 1PE12.4, dump file names),
    stat.res, the out*.res file names, and their arrays -- rho, u, v, w, E, T,
    p, Y_k, the reaction rates and the elapsed time -- at 1e-10 max-norm."""
    import golden_cases
    from sengapy.dump import OUT_FIELDS, read_out_res
    from conftest import rel_err
    exe = _driver()
    fx = dict(np.load(os.path.join(HERE, "golden", "output_h2_nscbc_mix_12x11x11.npz")))
    c = golden_cases.build("h2_nscbc_mix_12x11x11")
    c.dump.state = fx["state0"]
    c.dump.tstep = float(fx["tstep"])
    c.control.tstep = c.dump.tstep
    c.control.ntime = int(fx["nsteps"])
    c.control.ntdump, c.control.ntrept, c.control.ntstat = 2, 1, 2
    c.write_inputs(str(tmp_path))
    print("fortran driver outputs", exe, flush=True)
    r = subprocess.run([exe, str(tmp_path)], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr, flush=True)
    assert r.returncode == 0
    od = tmp_path / "output"
    rept = [ln.rstrip() for ln in (od / "rept.res").read_text().splitlines()]
    first = next(i for i, ln in enumerate(rept) if ln.startswith("Time step number:"))
    assert rept[first:] == list(fx["rept_lines"])
    assert [ln.rstrip() for ln in (od / "stat.res").read_text().splitlines()] == list(fx["stat_lines"])
    names = sorted(f for f in os.listdir(od) if f.startswith("out"))
    assert names == list(fx["outres_names"])
    nx, ny, nz = c.shape
    for q, nm in enumerate(names):
        o = read_out_res(str(od / nm), nx, ny, nz, c.nspec)
        errs = {k: rel_err(o[k], fx[f"out{q}_{k}"]) for k in OUT_FIELDS + ("Y", "rate")}
        print(nm, " ".join(f"{k} {e:.1e}" for k, e in errs.items()), flush=True)
        assert max(errs.values()) < 1e-10
        assert o["etime"] == float(fx[f"out{q}_etime"])
