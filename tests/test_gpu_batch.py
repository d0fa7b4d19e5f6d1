"""Batched execution (sg_batch_*) against the same cases run alone (needs the GPU).

A batch launches every substage kernel once over all its members, from a
member table and block map, and replays one captured graph per RK step.  Each
member evaluates exactly the arithmetic it evaluates alone, so the bar is
bit-identity with the single-case run (FP64 and mixed), including members
that join or leave between steps; each member is also checked against the
oracle (FP64 1e-10, mixed 1e-5).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SG_EINVAL, SG_ESTATE = 1, 6


def _specs(kind, shapes, seed=5):
    from sengapy import case as C
    base = C.batch_specs(len(shapes), seed=seed) if kind == "nscbc" else C.kernel_specs(len(shapes), seed=seed)
    for s, n in zip(base, shapes):
        s["n"] = n
    return base


def _solver(spec, precision=0):
    from sengapy import case as C
    from sengapy.abi import Solver
    c = C.make_case(spec)
    s = Solver(c, precision=precision)
    s.set_timestep(c.dump.tstep)
    s.upload(c.dump.state, 0.0)
    return s


def _alone(spec, steps, precision=0):
    s = _solver(spec, precision)
    s.step(steps)
    run, T, P, _ = s.download()
    s.close()
    return run, T, P


SHAPES = [(32, 16, 16), (24, 20, 16), (40, 16, 24), (16, 16, 16)]


def _oracle(spec, steps):
    from pyoracle import Oracle
    from sengapy import case as C
    c = C.make_case(spec)
    o = Oracle(c)
    o.set_timestep(c.dump.tstep)
    o.set_state(c.dump.state, 0.0)
    o.step(steps)
    run, T, _, _ = o.get()
    return run, T


@pytest.mark.parametrize("kind,precision", [("nscbc", 0), ("nscbc", 1), ("kernel", 0), ("kernel", 1)])
def test_batch_matches_alone(kind, precision):
    """Each member of a batch of uneven cases equals its solo run bit for bit
    and the oracle within the precision's bar (FP64 1e-10, mixed 1e-5)."""
    from conftest import assert_state_close
    from sengapy.abi import Batch
    specs = _specs(kind, SHAPES)
    print("batch", kind, precision, [s["n"] for s in specs], flush=True)
    steps = 3
    ref = [_alone(s, steps, precision) for s in specs]
    b = Batch()
    sol = [_solver(s, precision) for s in specs]
    for s in sol:
        b.add(s)
    assert b.size() == len(sol)
    b.step(steps)
    b.sync()
    tol, tol_T = (1e-10, 1e-12) if precision == 0 else (1e-5, 1e-5)
    for q, s in enumerate(sol):
        run, T, P, _ = s.download()
        assert np.array_equal(run, ref[q][0]), f"member {q} registers differ"
        assert np.array_equal(T, ref[q][1]), f"member {q} temperature differs"
        assert np.array_equal(P, ref[q][2]), f"member {q} pressure differs"
        run_o, T_o = _oracle(specs[q], steps)
        assert_state_close(f"member {q} vs oracle", run, T, run_o, T_o, tol, tol_T)
    b.close()


def test_members_join_and_leave():
    """Members added and removed between steps keep their single-case results;
    the graph is recaptured for the new membership."""
    from sengapy.abi import Batch
    specs = _specs("nscbc", SHAPES, seed=9)
    print("join/leave", [s["n"] for s in specs], flush=True)
    r2 = [_alone(s, 2) for s in specs]
    r4 = [_alone(s, 4) for s in specs[:3]]
    b = Batch()
    sol = [_solver(s) for s in specs]
    for s in sol[:3]:
        b.add(s)
    b.step(2)
    b.remove(sol[1])            # leaves after 2 steps
    b.add(sol[3])               # joins for 2 steps
    b.step(2)
    b.sync()
    got = [s.download() for s in sol]
    assert np.array_equal(got[0][0], r4[0][0])
    assert np.array_equal(got[1][0], r2[1][0])
    assert np.array_equal(got[2][0], r4[2][0])
    assert np.array_equal(got[3][0], r2[3][0])
    b.close()


def test_batch_rejects_mismatched_members():
    from sengapy import case as C
    from sengapy.abi import Batch, SgError, Solver
    specs = _specs("nscbc", SHAPES[:2], seed=3)
    b = Batch()
    a = _solver(specs[0])
    b.add(a)
    # a different mechanism (2 species) and a different x boundary kind
    one = C.periodic_1step(n=(16, 16, 16), nsteps=1)
    s1 = Solver(one)
    s1.set_timestep(one.dump.tstep)
    s1.upload(one.dump.state, 0.0)
    with pytest.raises(SgError) as e:
        b.add(s1)
    assert e.value.code == SG_EINVAL
    kspec = _specs("kernel", SHAPES[:1], seed=3)[0]
    k = _solver(kspec)
    with pytest.raises(SgError) as e:
        b.add(k)
    assert e.value.code == SG_EINVAL
    # mixed precision with FP64 members
    m = _solver(specs[1], precision=1)
    with pytest.raises(SgError) as e:
        b.add(m)
    assert e.value.code == SG_EINVAL
    # a member cannot be stepped on its own or added twice
    with pytest.raises(SgError) as e:
        a.step(1)
    assert e.value.code == SG_ESTATE
    with pytest.raises(SgError) as e:
        b.add(a)
    assert e.value.code == SG_ESTATE
    # a destroyed member leaves the batch
    a.close()
    assert b.size() == 0
    b.step(1)
    b.sync()
    b.close()


@pytest.mark.parametrize("kind,precision", [("nscbc", 0), ("kernel", 1)])
def test_tiles_and_autotune_leave_results_unchanged(kind, precision):
    """Every pass-tile choice computes bit-identical results, and
    the autotuner restores the state it times on."""
    specs = _specs(kind, [(48, 20, 16)], seed=21)
    spec = specs[0]
    print("tiles", kind, precision, spec["n"], flush=True)
    ref = _alone(spec, 2, precision)
    for narrow in ((1, 1, 1), (1, 0, 1), (0, 1, 0)):
        s = _solver(spec, precision)
        s.set_tiles(narrow)
        assert s.tiles() == narrow
        s.step(2)
        run, T, P, _ = s.download()
        s.close()
        assert np.array_equal(run, ref[0]), f"tiles {narrow}"
        assert np.array_equal(T, ref[1])
    s = _solver(spec, precision)
    nw, ms = s.autotune(3)
    print("autotuned", nw, round(ms, 4), flush=True)
    assert ms > 0 and all(v in (0, 1) for v in nw)
    s.step(2)
    run, T, _, _ = s.download()
    s.close()
    assert np.array_equal(run, ref[0]) and np.array_equal(T, ref[1])


def test_batch_autotune_keeps_results():
    from sengapy.abi import Batch
    specs = _specs("nscbc", [(48, 16, 16), (32, 20, 16), (16, 16, 24)], seed=31)
    ref = [_alone(s, 2) for s in specs]
    b = Batch()
    sol = [_solver(s) for s in specs]
    for s in sol:
        b.add(s)
    nw, ms = b.autotune(3)
    print("batch autotuned", nw, round(ms, 4), flush=True)
    b.step(2)
    b.sync()
    for q, s in enumerate(sol):
        assert np.array_equal(s.download()[0], ref[q][0]), f"member {q}"
    b.set_tiles((1, 1, 1))
    assert b.tiles() == (1, 1, 1)
    b.close()


def test_batch_with_transport_options():
    """Cases with the four molecular-transport options and radiation (the
    passes with those terms, the thermal-diffusion and radiation kernels)
    batched: each member equals its solo run bit for bit; a member without
    the options is refused."""
    from sengapy import case as C
    from sengapy.abi import Batch, SgError, Solver
    cases = []
    for n, L in (((24, 16, 16), (3e-3, 2e-3, 2e-3)), ((32, 20, 16), (4e-3, 2.5e-3, 2e-3))):
        c = C.h2_nscbc(n=n, length=L, mixavg=True, nsteps=3)
        c.control.nfmixw = c.control.nfmixp = c.control.nfmsor = c.control.nfmduf = 1
        c.control.nfradn = 1
        cases.append(c)
    print("batch with transport options", [c.shape for c in cases], flush=True)

    def solver(c):
        s = Solver(c)
        s.set_timestep(c.dump.tstep)
        s.upload(c.dump.state, 0.0)
        return s
    ref = []
    for c in cases:
        s = solver(c)
        s.step(3)
        ref.append(s.download())
        s.close()
    b = Batch()
    sol = [solver(c) for c in cases]
    for s in sol:
        b.add(s)
    plain = C.h2_nscbc(n=(24, 16, 16), length=(3e-3, 2e-3, 2e-3), mixavg=True, nsteps=1)
    p = solver(plain)
    with pytest.raises(SgError) as e:
        b.add(p)
    assert e.value.code == SG_EINVAL
    b.step(3)
    b.sync()
    for q, s in enumerate(sol):
        run, T, P, _ = s.download()
        assert np.array_equal(run, ref[q][0]), f"member {q} registers differ"
        assert np.array_equal(T, ref[q][1]), f"member {q} temperature differs"
    b.close()
    p.close()
