"""Continuous-batching request loop for flame cases on one GPU.

Requests (an initial dump, its time step and a step count) arrive over time.
The loop admits waiting requests first-come first-served while the device
memory they need (sg_footprint) fits the memory budget and the grid points
already in flight stay below a point budget, keeps every admitted case's state
resident on the device, and advances all admitted cases together through one
sg_batch (one launch per substage kernel over all cases, one captured graph
per RK step).  Cases join the batch as soon as they are admitted and leave it
as soon as their steps are done, so requests that arrive while others are on
the GPU start at the next chunk boundary; their uploads and the downloads of
finished cases overlap the batch's kernels.

Every case computes in the batch exactly what it computes alone (the batch
kernels evaluate each member's own arithmetic), which the GPU tests check
bit for bit.

The reference has no serving layer; this is the host side of the batched
extension of the src/senga2.f:96-194 step loop.
"""
from __future__ import annotations

import time
from collections import deque
from dataclasses import dataclass, field
from typing import List, Optional

import numpy as np

from . import abi


@dataclass
class Request:
    rid: int
    case: object            # sengapy.case.Case: initial dump (state, tstep) and configuration
    nsteps: int
    arrival: float = 0.0    # seconds after the start of the run


@dataclass
class Result:
    rid: int
    arrival: float
    admitted: float
    finished: float
    nsteps: int
    points: int
    run: Optional[np.ndarray] = field(default=None, repr=False)
    T: Optional[np.ndarray] = field(default=None, repr=False)
    P: Optional[np.ndarray] = field(default=None, repr=False)

    @property
    def latency(self) -> float:
        """Time to solution: arrival to downloaded result."""
        return self.finished - self.arrival


def _points(case) -> int:
    nx, ny, nz = case.shape
    return nx * ny * nz


class _Live:
    __slots__ = ("req", "solver", "left", "admitted", "bytes", "points")

    def __init__(self, req, solver, admitted, nbytes):
        self.req, self.solver, self.left = req, solver, req.nsteps
        self.admitted, self.bytes, self.points = admitted, nbytes, _points(req.case)


class Server:
    """Continuous batching on the current device.

    hbm_budget: device bytes the admitted cases may hold together.
    point_budget: grid points in flight beyond which waiting requests stay
        queued (a few million points already fill the GPU; admitting more
        only stretches every admitted case's time to solution).  Measured on
        a Poisson stream of 48-160^3 cases: 3-6 M points keep the throughput
        of larger budgets at the sequential time to solution or better.  A
        request larger than the budget still runs when nothing else is in
        flight.
    chunk: RK steps between admission/retirement points.
    precision: 0 = FP64, 1 = mixed (every case of a server shares it).
    keep: download the final state of each case into its Result.
    """

    def __init__(self, hbm_budget=240e9, point_budget=6e6, chunk=5, precision=0, lib=None, keep=True):
        self.lib = lib or abi.load_library()
        self.hbm_budget, self.point_budget = float(hbm_budget), float(point_budget)
        self.chunk, self.precision, self.keep = int(chunk), int(precision), keep
        self.max_points_in_flight = 0
        self.max_bytes_in_flight = 0

    def _finish(self, live, t0, clock):
        res = Result(live.req.rid, live.req.arrival, live.admitted, 0.0, live.req.nsteps, live.points)
        if self.keep:
            res.run, res.T, res.P, _ = live.solver.download()
        else:
            live.solver.sync()
        res.finished = clock() - t0
        return res

    def run(self, requests: List[Request], clock=time.perf_counter) -> List[Result]:
        pending = deque(sorted(requests, key=lambda r: r.arrival))
        waiting: deque = deque()
        active: List[_Live] = []
        retired: List[_Live] = []
        results: List[Result] = []
        batch = abi.Batch(self.lib)
        mem = pts = 0
        t0 = clock()
        try:
            while pending or waiting or active or retired:
                now = clock() - t0
                while pending and pending[0].arrival <= now:
                    waiting.append(pending.popleft())
                # admission, first come first served
                while waiting:
                    r = waiting[0]
                    need = abi.footprint(r.case, self.precision, self.lib)
                    busy = bool(active)
                    if mem + need > self.hbm_budget and busy:
                        break
                    if busy and pts + _points(r.case) > self.point_budget:
                        break
                    waiting.popleft()
                    s = abi.Solver(r.case, self.lib, self.precision)
                    s.set_timestep(r.case.dump.tstep)
                    s.upload(r.case.dump.state, 0.0)
                    live = _Live(r, s, clock() - t0, need)
                    batch.add(s)
                    active.append(live)
                    mem += need
                    pts += live.points
                self.max_points_in_flight = max(self.max_points_in_flight, pts)
                self.max_bytes_in_flight = max(self.max_bytes_in_flight, mem)
                if not active:
                    for live in retired:
                        results.append(self._finish(live, t0, clock))
                        live.solver.close()
                        mem -= live.bytes
                    retired = []
                    if pending and not waiting:
                        time.sleep(max(0.0, pending[0].arrival - (clock() - t0)))
                    continue
                k = min(self.chunk, min(x.left for x in active))
                batch.step(k)
                # results of the cases that left at the last boundary come
                # back while the batch computes
                for live in retired:
                    results.append(self._finish(live, t0, clock))
                done_now = retired
                retired = []
                batch.sync()
                for live in done_now:
                    live.solver.close()
                    mem -= live.bytes
                for live in active:
                    live.left -= k
                for live in [x for x in active if x.left == 0]:
                    batch.remove(live.solver)
                    active.remove(live)
                    pts -= live.points
                    retired.append(live)
        finally:
            for live in active + retired:
                live.solver.close()
            batch.close()
        results.sort(key=lambda r: r.rid)
        return results


def run_sequential(requests: List[Request], precision=0, lib=None, keep=True, clock=time.perf_counter):
    """Baseline: each request solved alone, one after another in arrival
    order (create, upload, step, download), waiting for arrivals."""
    lib = lib or abi.load_library()
    out = []
    t0 = clock()
    for r in sorted(requests, key=lambda r: r.arrival):
        now = clock() - t0
        if r.arrival > now:
            time.sleep(r.arrival - now)
        adm = clock() - t0
        s = abi.Solver(r.case, lib, precision)
        s.set_timestep(r.case.dump.tstep)
        s.upload(r.case.dump.state, 0.0)
        s.step(r.nsteps)
        res = Result(r.rid, r.arrival, adm, 0.0, r.nsteps, _points(r.case))
        if keep:
            res.run, res.T, res.P, _ = s.download()
        else:
            s.sync()
        res.finished = clock() - t0
        s.close()
        out.append(res)
    out.sort(key=lambda r: r.rid)
    return out


def summary(results: List[Result]) -> dict:
    """cases/s over the run's makespan and time-to-solution percentiles."""
    lat = np.array([r.latency for r in results])
    span = max(r.finished for r in results) - min(r.arrival for r in results)
    return {"cases": len(results), "makespan_s": span, "cases_per_s": len(results) / span,
            "p50_s": float(np.percentile(lat, 50)), "p99_s": float(np.percentile(lat, 99)),
            "mean_s": float(lat.mean()),
            "mpoint_substages_per_s": sum(r.points * 5 * r.nsteps for r in results) / span / 1e6}
