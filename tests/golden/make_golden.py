#!/usr/bin/env python3
"""Write the golden fixtures of this directory from the reference SENGA2 build.

Runs the reference Fortran (compiled in place from /root/reference by
oracle/ref/build_ref.py, driven by oracle/ref/refrun.f) on small seeded cases
and stores inputs and outputs as compressed .npz files
(usage: make_golden.py [CASE ...] for only the named state cases).  Needs /root/reference
and ROCm flang; only run in the build container.  The GPU tests never read the
reference: they compare against these files and against the C oracle.

Per case the fixture holds
  state0   initial solution register (5+N, nz, ny, nx), the exact input dump
  tstep    fixed time step (float32-exact, see sengapy.case._f32)
  T0, itndex0      temperature and interval index after TEMPIN
  rhs, T1, P1, rate1   one RHS evaluation at RK substage 1 (mode 1)
  run2, T2, P2, rate2  solution after NSTEPS full steps (mode 2)
and for the operator fixtures (mode 3) the input field and the nine outputs.
The adaptive-step fixture holds state0, tstep, errold, errldr, and after
nsteps steps run2, T2, etime2 and the adapted tstep2.
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "senga-versiona_amd"))
sys.path.insert(0, os.path.join(ROOT, "oracle", "ref"))
sys.path.insert(0, HERE)

import golden_cases  # noqa: E402
import run_ref  # noqa: E402

NSTEPS = 2


def cases(names=None):
    return {n: golden_cases.build(n) for n in (names or golden_cases.NAMES)}


def main(names=None):
    """All fixtures, or only the named state cases when names are given."""
    for name, c in cases(names).items():
        r1 = run_ref.run(c, 1)
        r2 = run_ref.run(c, 2, nsteps=NSTEPS)
        np.savez_compressed(os.path.join(HERE, name + ".npz"), state0=c.dump.state, tstep=c.dump.tstep,
                            nsteps=NSTEPS, T0=r1["T0"], itndex0=r1["itndex0"], rhs=r1["rhs"],
                            T1=r1["T"], P1=r1["P"], rate1=r1["rate"], run2=r2["run"], T2=r2["T"],
                            P2=r2["P"], rate2=r2["rate"])
        print("wrote", name)
    if names:
        return
    rng = np.random.default_rng(20240601)
    for name, cname in golden_cases.OPS.items():
        c = golden_cases.build(cname)
        nx, ny, nz = c.shape
        f = rng.standard_normal((nz, ny, nx))
        r = run_ref.run(c, 3, opin=f)
        np.savez_compressed(os.path.join(HERE, name + ".npz"), field=f, ops=r["ops"])
        print("wrote", name)


OUTPUT_CASE = "h2_nscbc_mix_12x11x11"
OUTPUT_STEPS = 4


def outputs():
    """The run outputs of src/output.f for the H2 NSCBC mixture-averaged case:
    OUTPUT_STEPS steps with reports every step, dumps every 2 steps and
    statistics every 2 steps.  Stores the report and statistics lines that
    OUTPUT appends (rept.res after its start-up header; stat.res whole) and
    the arrays of each out*.res post-processing file."""
    from sengapy.dump import OUT_FIELDS, read_out_res
    c = golden_cases.build(OUTPUT_CASE)
    c.control.ntdump, c.control.ntrept, c.control.ntstat = 2, 1, 2
    r = run_ref.run(c, 5, nsteps=OUTPUT_STEPS)
    rept = r["rept"].splitlines()
    first = next(i for i, ln in enumerate(rept) if ln.startswith("Time step number:"))
    nx, ny, nz = c.shape
    arrays = {}
    for q, (fname, raw) in enumerate(sorted(r["outres"].items())):
        o = read_out_res(raw, nx, ny, nz, c.nspec)
        for k in OUT_FIELDS + ("Y", "rate"):
            arrays[f"out{q}_{k}"] = o[k]
        arrays[f"out{q}_etime"] = o["etime"]
    np.savez_compressed(os.path.join(HERE, "output_" + OUTPUT_CASE + ".npz"),
                        state0=c.dump.state, tstep=c.dump.tstep, nsteps=OUTPUT_STEPS,
                        rept_lines=np.array([ln.rstrip() for ln in rept[first:]]),
                        stat_lines=np.array([ln.rstrip() for ln in r["stat"].splitlines()]),
                        outres_names=np.array(sorted(r["outres"])), **arrays)
    print("wrote output fixture:", sorted(r["outres"]), rept[first:])


ADAPT_STEPS = 4


def adaptive():
    """The H2 NSCBC mixture-averaged case with the adaptive time step
    (cont.dat step switch 1, ADAPTT after every step): state after ADAPT_STEPS
    steps, elapsed time and the adapted time step."""
    name = golden_cases.ADAPT
    c = golden_cases.build(name)
    r2 = run_ref.run(c, 2, nsteps=ADAPT_STEPS)
    np.savez_compressed(os.path.join(HERE, name + ".npz"), state0=c.dump.state, tstep=c.dump.tstep,
                        nsteps=ADAPT_STEPS, errold=c.dump.errold, errldr=c.dump.errldr,
                        run2=r2["run"], T2=r2["T"], etime2=r2["etime"], tstep2=r2["tstep"])
    print("wrote", name, "tstep", c.dump.tstep, "->", r2["tstep"])


if __name__ == "__main__":
    if sys.argv[1:] == ["adaptive"]:
        adaptive()
    elif sys.argv[1:] == ["outputs"]:
        outputs()
    else:
        main(sys.argv[1:])
        if not sys.argv[1:]:
            adaptive()
