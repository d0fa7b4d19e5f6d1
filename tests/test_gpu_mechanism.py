"""The compiled H2-air mechanism against the table-driven step loop (needs the GPU).

sg_create compares a configuration's chemistry with the mechanism compiled
into the library (csrc/sg_mech_h2.h, written by mechgen.py from the
reference's input/chem.dat) bit for bit; on a match the point kernels run the
CHRATE step loop (src/chrate.f:146-760) unrolled over its constants.  The
unrolled loop evaluates the same expressions on the same values as the
table-driven loop, so the bar is bit-identity of the state, temperature,
pressure and rates after several RK steps, FP64 and mixed.  Setting an unused
slot of the SRI table leaves the arithmetic unchanged but defeats the match,
which forces the table-driven loop on the same case.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _run(case, precision, cfg=None, steps=3):
    from sengapy.abi import Solver
    s = Solver(case, precision=precision, cfg=cfg)
    compiled = s.mechanism_compiled()
    s.set_timestep(case.dump.tstep)
    s.upload(case.dump.state, 0.0)
    s.step(steps)
    out = s.download()
    s.close()
    return compiled, out


@pytest.mark.parametrize("kind,precision", [("nscbc", 0), ("nscbc", 1), ("kernel", 0), ("kernel", 1)])
def test_compiled_mechanism_equals_table_loop(kind, precision):
    from sengapy import abi
    from sengapy import case as C
    spec = (C.batch_specs(1, seed=41) if kind == "nscbc" else C.kernel_specs(1, seed=41))[0]
    spec["n"] = (32, 20, 16)
    case = C.make_case(spec)
    print("mechanism", kind, precision, spec["n"], flush=True)
    fixed, a = _run(case, precision)
    cfg = abi.config_from_case(case, precision)
    cfg.rcsrif[abi.MAX_FALL - 1][7] = 1.0      # unused by every step
    table, b = _run(case, precision, cfg)
    assert fixed and not table
    for name, x, y in zip(("registers", "T", "p", "rates"), a, b):
        assert np.array_equal(x, y), f"{name} differ between the compiled and the table-driven loop"


def test_other_mechanisms_use_the_table_loop():
    from sengapy import case as C
    from sengapy.abi import Solver
    one = C.periodic_1step(n=(16, 16, 16), nsteps=1)
    s = Solver(one)
    assert not s.mechanism_compiled()
    s.close()
