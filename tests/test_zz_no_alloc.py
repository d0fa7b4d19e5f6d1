"""sg_derivatives_noalloc: the no-allocation variant of sg_derivatives (needs the GPU).

Called through the C ABI on two handles in a row -- the second with a smaller
grid, so that it runs on the staging array the first call left behind -- and
each result compared bit for bit with sg_derivatives on the same handle and
field; a grid above SG_NOALLOC_MAX_POINTS is refused without a launch."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SG_EINVAL = 1
MAX_POINTS = 1 << 19


def _handle(n, bounded):
    from sengapy import case as Cs
    from sengapy.abi import Solver
    if bounded:
        c = Cs.h2_nscbc(n=n, length=tuple(1e-4 * (m - 1) for m in n), mixavg=False, nsteps=1)
    else:
        c = Cs.periodic_1step(n=n, nsteps=1)
    return Solver(c)


def _noalloc(s, f):
    from sengapy.abi import _dp
    f = np.ascontiguousarray(f, dtype=np.float64)
    out = np.zeros((9,) + f.shape)
    rc = s.lib.sg_derivatives_noalloc(s.h, _dp(f), _dp(out))
    return rc, out


def test_noalloc_derivatives_equal_allocating_ones():
    rng = np.random.default_rng(11)
    for n, bounded in (((48, 40, 36), True), ((20, 16, 12), False)):
        print("noalloc", n, "bounded x" if bounded else "periodic", flush=True)
        s = _handle(n, bounded)
        f = rng.standard_normal((n[2], n[1], n[0]))
        ref = s.derivatives(f)
        rc, got = _noalloc(s, f)
        assert rc == 0
        assert np.array_equal(got, ref)
        s.close()


def test_noalloc_refuses_grids_above_its_staging():
    n = (96, 96, 64)       # 589,824 points > 2^19
    assert n[0] * n[1] * n[2] > MAX_POINTS
    print("noalloc refusal", n, flush=True)
    s = _handle(n, False)
    rc, _ = _noalloc(s, np.zeros((n[2], n[1], n[0])))
    assert rc == SG_EINVAL
    s.close()
