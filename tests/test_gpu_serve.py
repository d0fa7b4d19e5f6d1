"""Continuous-batching serving loop (sengapy.serve) on the GPU.

Requests arrive while others are in flight, join the batch at the next chunk
boundary and leave it when their own step count is done.  Each result must be
the one the case gives when solved alone: the batch kernels evaluate every
member's own arithmetic, so the bar is bit-identity (FP64 and mixed); each
result is also checked against the oracle (FP64 1e-10, mixed 1e-5).
"""
import numpy as np
import pytest

from conftest import assert_state_close

pytestmark = pytest.mark.gpu


def _requests(precision_seed=0):
    from sengapy import case as C
    from sengapy.serve import Request
    specs = C.kernel_specs(5, seed=77 + precision_seed)
    shapes = [(16, 16, 16), (24, 16, 20), (20, 20, 16), (16, 24, 16), (32, 16, 16)]
    steps = [3, 7, 4, 5, 2]
    arrivals = [0.0, 0.0, 0.05, 0.1, 0.3]
    out = []
    for q, (sp, n) in enumerate(zip(specs, shapes)):
        sp["n"] = n
        out.append(Request(q, C.make_case(sp), steps[q], arrivals[q]))
    return out


def _alone(req, precision):
    from sengapy.abi import Solver
    s = Solver(req.case, precision=precision)
    s.set_timestep(req.case.dump.tstep)
    s.upload(req.case.dump.state, 0.0)
    s.step(req.nsteps)
    run, T, P, _ = s.download()
    s.close()
    return run, T, P


def _oracle(req):
    from pyoracle import Oracle
    o = Oracle(req.case)
    o.set_timestep(req.case.dump.tstep)
    o.set_state(req.case.dump.state, 0.0)
    o.step(req.nsteps)
    run, T, _, _ = o.get()
    return run, T


@pytest.mark.parametrize("precision,point_budget", [(0, 24e6), (1, 24e6), (0, 5000)])
def test_served_results_equal_alone(precision, point_budget):
    from sengapy.serve import Server, summary
    reqs = _requests(precision)
    print("serve", precision, point_budget, flush=True)
    srv = Server(chunk=2, precision=precision, point_budget=point_budget)
    res = srv.run(reqs)
    assert [r.rid for r in res] == list(range(len(reqs)))
    tol, tol_T = (1e-10, 1e-12) if precision == 0 else (1e-5, 1e-5)
    for r, q in zip(res, reqs):
        run, T, P = _alone(q, precision)
        assert np.array_equal(r.run, run), f"request {q.rid}: registers differ from the solo run"
        assert np.array_equal(r.T, T)
        assert np.array_equal(r.P, P)
        assert r.finished >= r.admitted >= r.arrival - 1e-9
        run_o, T_o = _oracle(q)
        assert_state_close(f"request {q.rid} vs oracle", r.run, r.T, run_o, T_o, tol, tol_T)
    if point_budget < 2 * 16 ** 3:
        # a budget below any two cases together: one case in flight at a time
        assert srv.max_points_in_flight == max(r.points for r in res)
    else:
        assert srv.max_points_in_flight > max(r.points for r in res)
    s = summary(res)
    print({k: round(v, 4) for k, v in s.items()}, flush=True)
    assert s["cases"] == len(reqs) and s["p99_s"] >= s["p50_s"] > 0


def test_sequential_baseline_equals_alone():
    from sengapy.serve import run_sequential
    reqs = _requests()[:3]
    for q in reqs:
        q.arrival = 0.0
    res = run_sequential(reqs)
    for r, q in zip(res, reqs):
        run, _, _ = _alone(q, 0)
        assert np.array_equal(r.run, run)
