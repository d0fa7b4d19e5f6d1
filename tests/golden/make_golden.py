#!/usr/bin/env python3
"""Write the golden fixtures of this directory from the reference SENGA2 build.

Runs the reference Fortran (compiled in place from /root/reference by
oracle/ref/build_ref.py, driven by oracle/ref/refrun.f) on small seeded cases
and stores inputs and outputs as compressed .npz files.  Needs /root/reference
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


def cases():
    return {n: golden_cases.build(n) for n in golden_cases.NAMES}


def main():
    for name, c in cases().items():
        r1 = run_ref.run(c, 1)
        r2 = run_ref.run(c, 2, nsteps=NSTEPS)
        np.savez_compressed(os.path.join(HERE, name + ".npz"), state0=c.dump.state, tstep=c.dump.tstep,
                            nsteps=NSTEPS, T0=r1["T0"], itndex0=r1["itndex0"], rhs=r1["rhs"],
                            T1=r1["T"], P1=r1["P"], rate1=r1["rate"], run2=r2["run"], T2=r2["T"],
                            P2=r2["P"], rate2=r2["rate"])
        print("wrote", name)
    rng = np.random.default_rng(20240601)
    for name, cname in golden_cases.OPS.items():
        c = golden_cases.build(cname)
        nx, ny, nz = c.shape
        f = rng.standard_normal((nz, ny, nx))
        r = run_ref.run(c, 3, opin=f)
        np.savez_compressed(os.path.join(HERE, name + ".npz"), field=f, ops=r["ops"])
        print("wrote", name)


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
    else:
        main()
        adaptive()
