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

    def _admit(self, waiting, busy, mem, pts, t0, clock):
        """Admit waiting requests first come first served while they fit the
        budgets: each gets its handle, its state uploaded and its initial
        temperature solved on its own stream.  busy: cases already in flight
        (a request above the budgets still runs when nothing else does)."""
        out = []
        while waiting:
            r = waiting[0]
            need = abi.footprint(r.case, self.precision, self.lib)
            busy_now = busy or bool(out)
            if busy_now and mem + need > self.hbm_budget:
                break
            if busy_now and pts + _points(r.case) > self.point_budget:
                break
            waiting.popleft()
            s = abi.Solver(r.case, self.lib, self.precision)
            s.set_timestep(r.case.dump.tstep)
            s.upload(r.case.dump.state, 0.0)
            live = _Live(r, s, clock() - t0, need)
            out.append(live)
            mem += need
            pts += live.points
        return out, mem, pts

    def run(self, requests: List[Request], clock=time.perf_counter) -> List[Result]:
        """Serve the requests.  Each chunk of k RK steps is queued for the
        batch; while it computes, the cases that left at the previous
        boundary download their results and newly arrived requests are
        admitted and uploaded on their own streams.  They join the batch at
        the next boundary."""
        pending = deque(sorted(requests, key=lambda r: r.arrival))
        waiting: deque = deque()
        active: List[_Live] = []
        staged: List[_Live] = []     # admitted and uploaded, joining at the next boundary
        retired: List[_Live] = []
        results: List[Result] = []
        batch = abi.Batch(self.lib)
        mem = pts = 0
        t0 = clock()

        def arrivals():
            now = clock() - t0
            while pending and pending[0].arrival <= now:
                waiting.append(pending.popleft())

        try:
            while pending or waiting or active or staged or retired:
                arrivals()
                if not active and not staged:
                    # idle GPU: results still held, then admit (blocking)
                    for live in retired:
                        results.append(self._finish(live, t0, clock))
                        live.solver.close()
                        mem -= live.bytes
                    retired = []
                    staged, mem, pts = self._admit(waiting, False, mem, pts, t0, clock)
                    if not staged:
                        if pending:
                            time.sleep(max(0.0, pending[0].arrival - (clock() - t0)))
                        continue
                for live in staged:
                    batch.add(live.solver)
                    active.append(live)
                staged = []
                self.max_points_in_flight = max(self.max_points_in_flight, pts)
                self.max_bytes_in_flight = max(self.max_bytes_in_flight, mem)
                k = min(self.chunk, min(x.left for x in active))
                batch.step(k)
                # while the batch computes: results of the cases that left at
                # the last boundary come back, and new arrivals upload
                for live in retired:
                    results.append(self._finish(live, t0, clock))
                done_now = retired
                retired = []
                arrivals()
                staged, mem, pts = self._admit(waiting, True, mem, pts, t0, clock)
                self.max_points_in_flight = max(self.max_points_in_flight, pts)
                self.max_bytes_in_flight = max(self.max_bytes_in_flight, mem)
                batch.sync()
                # freeing device memory synchronises the device: after the chunk
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
            for live in active + staged + retired:
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
