#!/usr/bin/env python3
"""Throughput of the SENGA+ RK substage on one MI355X.

Workload: the 128^3 H2-air case with NSCBC inflow/outflow in x, periodic y/z,
9 species / 21 steps with Lindemann falloff, NASA-7 thermodynamics and
mixture-averaged transport, FP64, seeded Passot-Pouquet turbulence
(sengapy.case.h2_nscbc).  A "step" is one full RK4(3)5 time step (five
right-hand-side evaluations and register updates); the metric counts
grid-point-substages per second with the state resident in HBM.

Beside the headline figure, "batched" advances 64 uneven cases together
(sg_batch: one launch per substage kernel over all cases, one HIP graph per
RK step) against the same cases stepped one after another, and "serving"
pushes a Poisson stream of flame-kernel requests through the continuous-
batching loop (sengapy.serve) against sequential FP64 solves, with p50/p99
time to solution.  "batched_small" and "serving_small" repeat both on cases
of 16-32^3, which are too small to fill the GPU alone.

Output: one JSON line.  roofline: the dominant kernel's algorithmic HBM bytes
per launch over its mean duration (HIP events on the library's stream);
cpu_baseline: the reference SENGA2 compiled from its own sources (one rank,
one core; the C oracle where that binary is absent) on a bounded 32^3 sample
of the same physics.
"""
import argparse
import ctypes as C
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "senga-versiona_amd"))
sys.path.insert(0, os.path.join(ROOT, "oracle"))

HBM_PEAK = 8.0e12   # MI355X HBM3E, bytes/s


def hip():
    lib = C.CDLL("/opt/rocm/lib/libamdhip64.so")
    lib.hipEventCreate.argtypes = [C.POINTER(C.c_void_p)]
    lib.hipEventRecord.argtypes = [C.c_void_p, C.c_void_p]
    lib.hipEventSynchronize.argtypes = [C.c_void_p]
    lib.hipEventElapsedTime.argtypes = [C.POINTER(C.c_float), C.c_void_p, C.c_void_p]
    lib.hipDeviceSynchronize.argtypes = []
    lib.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    lib.hipFree.argtypes = [C.c_void_p]
    lib.hipMemsetAsync.argtypes = [C.c_void_p, C.c_int, C.c_size_t, C.c_void_p]
    lib.hipMemcpyAsync.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int, C.c_void_p]
    return lib


def copy_calibration(h, nbytes=1 << 30):
    """One device-to-device copy of a known size, the reference point for the
    FETCH_SIZE / WRITE_SIZE counters (a plain copy reads and writes nbytes)."""
    a, b = C.c_void_p(), C.c_void_p()
    if h.hipMalloc(C.byref(a), nbytes) or h.hipMalloc(C.byref(b), nbytes):
        raise RuntimeError("hipMalloc failed")
    h.hipMemsetAsync(a, 1, nbytes, None)
    h.hipDeviceSynchronize()
    h.hipMemcpyAsync(b, a, nbytes, 3, None)   # hipMemcpyDeviceToDevice
    h.hipDeviceSynchronize()
    h.hipFree(a)
    h.hipFree(b)
    print(f"copy calibration: {nbytes} bytes device-to-device", file=sys.stderr, flush=True)


def kernel_bytes(name, ns, bounded, nx=128):
    """Algorithmic bytes per grid point of one launch: each distinct input
    field read once and each output written once (FP64 fields, 4-byte index).
    The face kernels touch the nx = 1 and nx = n planes only: their bytes are
    spread over the grid (divided by nx)."""
    f = 8
    nprim = 9 + 3 * ns
    if name in ("boundt", "bountt"):   # inflow values of 5+N fields at x-left
        return f * (5 + ns) / nx
    if name == "nscbc":                # face data (6+N) + stage register (5+N) + corrections (5+N), two faces
        return 2 * f * ((6 + ns) + 2 * (5 + ns)) / nx
    if name == "prim":          # reads stage (5+N), T, index; writes primitives, rate, index
        return f * (5 + ns + 1) + 4 + f * (nprim + ns) + 4
    if name == "pass_x":        # rho, 3 momenta, primitives, rate -> accumulators, d mu/dx
        return f * (4 + nprim + ns) + f * (5 + ns) + f
    if name in ("pass_y", "pass_z"):   # accumulators read-modify-write
        return f * (4 + nprim) + 2 * f * (5 + ns) + f
    if name == "final":         # accumulators, 3 viscosity gradients, u,v,w,mu, run (r/w), stage (w)
        return f * (5 + ns) + f * 3 + f * 4 + 2 * f * (5 + ns) + f * (5 + ns)
    return 0


FP64_PEAK = 78.6e12   # MI355X vector FP64, FLOP/s


def _profile_json(name, n, prec):
    """Per-kernel entries of a committed rocprofv3 summary (profiles/NAME.json)
    for the h2_nscbc_mix n^3 workload and precision tag, or {}."""
    path = os.path.join(ROOT, "profiles", name + ".json")
    if not os.path.exists(path):
        return {}
    d = json.load(open(path))
    if d.get("workload") != f"h2_nscbc_mix_{n}^3":
        return {}
    return d.get(prec, {})


def kernel_table(kern, ns, npts, n, prec):
    """Per kernel: mean launch time (events on the library's stream),
    algorithmic bytes per point and the HBM rate and fraction they give,
    counted FP64 + FP32 flops per point and bytes moved per point (committed
    rocprofv3 counters of the same workload), and the kernel's floor
    max(algorithmic bytes / 8 TB/s, counted flops / 78.6 TFLOP/s)."""
    valu = _profile_json("valu", n, prec)
    traf = _profile_json("traffic", n, prec)
    out = {}
    for k, (tot, cnt) in kern.items():
        ms = tot / cnt
        ab = kernel_bytes(k, ns, True, n)
        e = {"ms": ms, "algorithmic_bytes_per_point": ab,
             "hbm_achieved_gbs": ab * npts / (ms / 1e3) / 1e9,
             "hbm_frac": ab * npts / (ms / 1e3) / HBM_PEAK}
        v = valu.get(k)
        fl = None
        if v:
            fl = v["fp64_flops_per_point"] + v["fp32_flops_per_point"]
            e["flops_per_point"] = fl
            e["valu_issue_use"] = v["valu_issue_util"]
        if k in traf:
            e["traffic_bytes_per_point"] = traf[k]
        t_hbm = ab * npts / HBM_PEAK * 1e3
        t_alu = fl * npts / FP64_PEAK * 1e3 if fl is not None else 0.0
        e["floor_ms"] = max(t_hbm, t_alu)
        e["floor_bound"] = "fp64 valu" if t_alu > t_hbm else "hbm"
        e["floor_frac"] = e["floor_ms"] / ms
        out[k] = e
    return out


def roofline_of(e, name, npts):
    """The roofline object of one kernel, against the bound its floor names:
    HBM (algorithmic bytes per launch over the mean launch time, 8 TB/s) or
    the FP64 vector ALU (counted flops per launch over the same time,
    78.6 TFLOP/s; the HBM fraction is given beside it)."""
    t = e["ms"] / 1e3
    traffic = e.get("traffic_bytes_per_point")
    r = {"kernel": name, "mean_launch_ms": e["ms"], "algorithmic_bytes_per_point": e["algorithmic_bytes_per_point"],
         "traffic": traffic * npts if traffic is not None else None}
    if e["floor_bound"] == "fp64 valu":
        ach = e["flops_per_point"] * npts / t
        r.update({"bound": "fp64 valu", "achieved": ach / 1e12, "peak": FP64_PEAK / 1e12, "unit": "TFLOP/s",
                  "frac": ach / FP64_PEAK, "flops_per_point": e["flops_per_point"], "hbm_frac": e["hbm_frac"],
                  "valu_issue_use": e.get("valu_issue_use")})
    else:
        r.update({"bound": "hbm", "achieved": e["hbm_achieved_gbs"], "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                  "frac": e["hbm_frac"]})
    return r


def _baseline_case():
    from sengapy import case as Cs
    return Cs.h2_nscbc(n=(32, 32, 32), length=(1.25e-3, 1.25e-3, 1.25e-3), mixavg=True, nsteps=1)


def cpu_baseline_port(seconds_target=10.0):
    """Oracle (scalar C restatement, one thread) on a 32^3 sample of the same
    H2 NSCBC mixture-averaged physics; whole RK steps."""
    from pyoracle import Oracle
    c = _baseline_case()
    o = Oracle(c)
    threads = o.threads()
    o.threads(1)             # the scalar restatement on one core
    o.set_timestep(c.dump.tstep)
    o.set_state(c.dump.state, 0.0)
    npts = 32 ** 3
    t0 = time.perf_counter()
    nsub = 0
    while True:
        for r in range(1, 6):
            o.substage(r)
            nsub += 1
        o.lib.orc_end_step(o.h)
        if time.perf_counter() - t0 > seconds_target:
            break
    dt = time.perf_counter() - t0
    o.threads(threads)
    return {"value": npts * nsub / dt / 1e6, "unit": "Mpoint-RK-substages/s", "cores": 1, "kind": "port",
            "sample": f"C oracle, 32^3 H2-air NSCBC mixture-averaged FP64, {nsub} substages "
                      f"({nsub // 5} RK steps) in {dt:.1f} s on one host core"}


def cpu_baseline(nsteps=10):
    """The reference SENGA2 itself, compiled from its sources (oracle/_ref,
    single rank, one core) on a 32^3 sample of the same H2 NSCBC mixture-
    averaged physics: wall time of nsteps full RK steps minus that of a
    0-step run (input, TEMPIN and output).  Falls back to the C oracle where
    the reference binary was not built; the oracle's figure is reported too."""
    sys.path.insert(0, os.path.join(ROOT, "oracle", "ref"))
    port = cpu_baseline_port()
    try:
        import run_ref
        c = _baseline_case()
        run_ref.build_for(c)
        t = {}
        for n in (0, nsteps):
            t0 = time.perf_counter()
            run_ref.run(c, 2, nsteps=n)
            t[n] = time.perf_counter() - t0
        dt = t[nsteps] - t[0]
    except (FileNotFoundError, RuntimeError, ImportError) as e:
        print(f"reference binary unavailable ({e}); oracle baseline only", file=sys.stderr, flush=True)
        return port
    npts = 32 ** 3
    return {"value": npts * 5 * nsteps / dt / 1e6, "unit": "Mpoint-RK-substages/s", "cores": 1, "kind": "reference",
            "sample": f"reference SENGA2 built from its own sources (flang -O2, one rank, its mpi_stub), "
                      f"32^3 H2-air NSCBC mixture-averaged FP64, {nsteps} RK steps: {t[nsteps]:.1f} s minus "
                      f"{t[0]:.1f} s for a 0-step run, one host core",
            "port": port}


def mixed_run(abi, c, a, h, s64):
    """The mixed-precision mode on the same case and step counts: throughput,
    speed-up over FP64 and max relative error per field (normalised by the
    field's maximum) against the FP64 state after warmup + steps steps."""
    import numpy as np
    s = abi.Solver(c, precision=1)
    s.set_timestep(c.dump.tstep)
    s.upload(c.dump.state, 0.0)
    tuned = (s.tiles(), None) if a.no_autotune else s.autotune(5)
    stream = C.c_void_p(s.stream())
    ev0, ev1 = C.c_void_p(), C.c_void_p()
    h.hipEventCreate(C.byref(ev0))
    h.hipEventCreate(C.byref(ev1))
    s.step(a.warmup)
    s.sync()
    h.hipEventRecord(ev0, stream)
    s.step(a.steps)
    h.hipEventRecord(ev1, stream)
    h.hipEventSynchronize(ev1)
    s.sync()
    ms = C.c_float(0)
    h.hipEventElapsedTime(C.byref(ms), ev0, ev1)
    # the FP64 run's profiled steps, so both states are at the same step
    s.profile(True)
    s.step(a.profile_steps)
    s.sync()
    prof = s.profile_read()
    s.profile(False)
    r32, T32, _, _ = s.download()
    r64, T64, _, _ = s64.download()

    def rel(x, y):
        return float(np.max(np.abs(x - y)) / max(float(np.max(np.abs(y))), 1e-300))

    ns = c.nspec
    n3 = c.shape[0] * c.shape[1] * c.shape[2]
    val = n3 * 5 * a.steps / (ms.value / 1e3) / 1e6
    print(f"mixed: timed {a.steps} steps: {ms.value:.2f} ms", file=sys.stderr, flush=True)
    return {"value": round(val, 3), "unit": "Mpoint-RK-substages/s", "ms_per_step": round(ms.value / a.steps, 4),
            "dtype": "f32 compute / f64 registers",
            "max_rel_err_T": rel(T32, T64), "max_rel_err_Y": max(rel(r32[5 + k], r64[5 + k]) for k in range(ns)),
            "max_rel_err_conserved": max(rel(r32[v], r64[v]) for v in range(5 + ns)),
            "steps_compared": a.warmup + a.steps + a.profile_steps,
            "tiles": {"narrow_pass_tiles": list(tuned[0]), "autotuned": not a.no_autotune},
            "kernel_ms_per_launch": {k: round(v[0] / v[1], 4) for k, v in prof.items() if v[1] > 0},
            "kernels": kernel_table({k: v for k, v in prof.items() if v[1] > 0 and kernel_bytes(k, ns, True, c.shape[0]) > 0},
                                    ns, n3, c.shape[0], "mixed")}


def large_run(abi, h, c, steps, warmup=2):
    """FP64 and mixed throughput on a larger grid of the same physics (state
    resident, tiles autotuned), timed with events on the library's stream."""
    n3 = c.shape[0] * c.shape[1] * c.shape[2]
    out = {"grid": list(c.shape), "steps": steps}
    ev0, ev1 = C.c_void_p(), C.c_void_p()
    h.hipEventCreate(C.byref(ev0))
    h.hipEventCreate(C.byref(ev1))
    for tag, prec in (("fp64", 0), ("mixed", 1)):
        s = abi.Solver(c, precision=prec)
        s.set_timestep(c.dump.tstep)
        s.upload(c.dump.state, 0.0)
        tiles, _ = s.autotune(3)
        s.step(warmup)
        s.sync()
        stream = C.c_void_p(s.stream())
        h.hipEventRecord(ev0, stream)
        s.step(steps)
        h.hipEventRecord(ev1, stream)
        h.hipEventSynchronize(ev1)
        s.sync()
        ms = C.c_float(0)
        h.hipEventElapsedTime(C.byref(ms), ev0, ev1)
        out[tag] = {"value": round(n3 * 5 * steps / (ms.value / 1e3) / 1e6, 3), "unit": "Mpoint-RK-substages/s",
                    "ms_per_step": round(ms.value / steps, 4), "device_bytes": s.memory(),
                    "narrow_pass_tiles": list(tiles)}
        s.close()
        print(f"{c.shape} {tag}: {ms.value:.2f} ms for {steps} steps", file=sys.stderr, flush=True)
    out["speedup_mixed"] = out["mixed"]["value"] / out["fp64"]["value"]
    return out


def _make(spec):
    from sengapy import case as Cs
    return Cs.make_case(spec)


def make_cases(specs, workers=8):
    """Initial dumps of a case set, built by a pool of host processes (before
    any GPU use in this process)."""
    import multiprocessing as mp
    t0 = time.perf_counter()
    with mp.get_context("fork").Pool(workers) as pool:
        cases = pool.map(_make, specs, chunksize=1)
    print(f"{len(cases)} cases built in {time.perf_counter() - t0:.1f} s", file=sys.stderr, flush=True)
    return cases


def batched_run(abi, h, cases, steps):
    """Uneven cases advanced together (one launch per substage kernel over all
    cases, one graph per RK step) against the same cases advanced one after
    another on their own handles; states resident in HBM in both runs.  Rates
    are normalised to 100-step cases."""
    import numpy as np
    out = {}
    pts = sum(c.shape[0] * c.shape[1] * c.shape[2] for c in cases)
    ev0, ev1 = C.c_void_p(), C.c_void_p()
    h.hipEventCreate(C.byref(ev0))
    h.hipEventCreate(C.byref(ev1))
    for prec in (0, 1):
        sols = []
        for c in cases:
            s = abi.Solver(c, precision=prec)
            s.set_timestep(c.dump.tstep)
            s.upload(c.dump.state, 0.0)
            sols.append(s)
        # sequential: each case on its own (tiles autotuned per case), one after the other
        if prec == 0:
            for s in sols:
                s.autotune(3)
                s.step(1)
            h.hipDeviceSynchronize()
            t0 = time.perf_counter()
            for s in sols:
                s.step(steps)
            h.hipDeviceSynchronize()
            t_seq = time.perf_counter() - t0
        b = abi.Batch()
        for s in sols:
            b.add(s)
        tiles, _ = b.autotune(3)
        b.step(1)
        b.sync()
        t0 = time.perf_counter()
        b.step(steps)
        b.sync()
        t_b = time.perf_counter() - t0
        tag = "fp64" if prec == 0 else "mixed"
        out[tag] = {"cases_per_s": len(cases) / (t_b * 100.0 / steps), "s_per_step": t_b / steps,
                    "mpoint_substages_per_s": pts * 5 * steps / t_b / 1e6, "autotuned_narrow_tiles": list(tiles)}
        if prec == 0:
            out["sequential_fp64"] = {"cases_per_s": len(cases) / (t_seq * 100.0 / steps),
                                      "s_per_step": t_seq / steps,
                                      "mpoint_substages_per_s": pts * 5 * steps / t_seq / 1e6}
        b.close()
        for s in sols:
            s.close()
        print(f"batched {tag}: {t_b:.3f} s for {steps} steps of {len(cases)} cases", file=sys.stderr, flush=True)
    out["speedup_fp64_vs_sequential"] = out["fp64"]["cases_per_s"] / out["sequential_fp64"]["cases_per_s"]
    out["speedup_mixed_vs_sequential_fp64"] = out["mixed"]["cases_per_s"] / out["sequential_fp64"]["cases_per_s"]
    return out


def serve_run(cases, specs, steps, point_budget=6e6, chunk=5):
    """Poisson request stream through the continuous-batching loop, FP64 and
    mixed, against sequential FP64 solves of the same stream."""
    from sengapy import serve
    reqs = [serve.Request(q, c, steps, sp["arrival"]) for q, (c, sp) in enumerate(zip(cases, specs))]
    out = {}
    for tag, prec in (("fp64", 0), ("mixed", 1)):
        srv = serve.Server(precision=prec, keep=True, point_budget=point_budget, chunk=chunk)
        res = srv.run(reqs)
        out[tag] = serve.summary(res)
        out[tag]["max_points_in_flight"] = srv.max_points_in_flight
        print(f"served {tag}: {out[tag]}", file=sys.stderr, flush=True)
    res = serve.run_sequential(reqs, precision=0)
    out["sequential_fp64"] = serve.summary(res)
    print(f"sequential: {out['sequential_fp64']}", file=sys.stderr, flush=True)
    out["speedup_fp64_vs_sequential"] = out["fp64"]["cases_per_s"] / out["sequential_fp64"]["cases_per_s"]
    out["speedup_mixed_vs_sequential_fp64"] = out["mixed"]["cases_per_s"] / out["sequential_fp64"]["cases_per_s"]
    return out


def precision_study(abi, h, c, steps):
    """FP64 and mixed runs of the same case for `steps` RK steps: throughput of
    each and, per field, the max and L2 relative error of the mixed state
    against the FP64 state (normalised by the field's max and L2 norm)."""
    import numpy as np
    n3 = c.shape[0] * c.shape[1] * c.shape[2]
    res = {}
    states = {}
    for tag, prec in (("fp64", 0), ("mixed", 1), ("mixed_renorm", 2)):
        s = abi.Solver(c, precision=prec)
        s.set_timestep(c.dump.tstep)
        s.upload(c.dump.state, 0.0)
        s.sync()
        t0 = time.perf_counter()
        done = 0
        while done < steps:   # progress lines every 10 steps
            k = min(10, steps - done)
            s.step(k)
            s.sync()
            done += k
            print(f"{tag}: {done}/{steps} steps, {time.perf_counter() - t0:.1f} s", file=sys.stderr, flush=True)
        dt = time.perf_counter() - t0
        res[tag] = {"mpoint_substages_per_s": n3 * 5 * steps / dt / 1e6, "s_per_step": dt / steps,
                    "device_bytes": s.memory()}
        states[tag] = s.download()[:3]
        s.close()
    names = ["rho", "rho_u", "rho_v", "rho_w", "rho_E"] + [f"rho_Y_{k}" for k in range(c.nspec)]
    r64, T64, P64 = states["fp64"]

    def both(x, y):
        d = x - y
        return {"max_rel": float(np.max(np.abs(d)) / max(float(np.max(np.abs(y))), 1e-300)),
                "l2_rel": float(np.sqrt(np.sum(d * d)) / max(float(np.sqrt(np.sum(y * y))), 1e-300))}

    for tag in ("mixed", "mixed_renorm"):
        r32, T32, P32 = states[tag]
        err = {}
        for v, nm in enumerate(names):
            err[nm] = both(r32[v], r64[v])
        err["T"] = both(T32, T64)
        err["p"] = both(P32, P64)
        # mass fractions Y_k = rho Y_k / rho
        ymax = 0.0
        for k in range(c.nspec):
            e = both(r32[5 + k] / r32[0], r64[5 + k] / r64[0])
            err[f"Y_{k}"] = e
            ymax = max(ymax, e["max_rel"])
        sfx = "" if tag == "mixed" else "_renorm"
        res["speedup_mixed" + sfx] = res[tag]["mpoint_substages_per_s"] / res["fp64"]["mpoint_substages_per_s"]
        res["errors_mixed" + sfx + "_vs_fp64"] = err
        res["max_rel_T" + sfx] = err["T"]["max_rel"]
        res["max_rel_Y" + sfx] = ymax
        res["species_sum_drift" + sfx] = float(np.abs(r32[5:].sum(axis=0) / r32[0] - 1.0).max())
    res["species_sum_drift_fp64"] = float(np.abs(r64[5:].sum(axis=0) / r64[0] - 1.0).max())
    return res


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n", type=int, default=128)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--profile-steps", type=int, default=2)
    ap.add_argument("--copy-calibration", action="store_true",
                    help="also run one 1 GiB device-to-device copy (counter calibration)")
    ap.add_argument("--no-mixed", action="store_true",
                    help="skip the mixed-precision run reported beside the FP64 figure")
    ap.add_argument("--batch-cases", type=int, default=64,
                    help="uneven H2-air cases of the batched run (0: skip)")
    ap.add_argument("--batch-steps", type=int, default=10)
    ap.add_argument("--batch-small-cases", type=int, default=64,
                    help="cases of the second batched run, on small grids {16,24,32}^3 (0: skip)")
    ap.add_argument("--batch-small-steps", type=int, default=20)
    ap.add_argument("--serve-requests", type=int, default=16,
                    help="flame-kernel requests of the served run (0: skip)")
    ap.add_argument("--serve-steps", type=int, default=40)
    ap.add_argument("--serve-rate", type=float, default=4.0)
    ap.add_argument("--serve-point-budget", type=float, default=6e6,
                    help="grid points in flight beyond which the serving loop queues requests")
    ap.add_argument("--serve-chunk", type=int, default=5, help="RK steps between admission points")
    ap.add_argument("--serve-small-requests", type=int, default=48,
                    help="requests of the second served run, on small grids {16,24,32}^3 (0: skip)")
    ap.add_argument("--serve-small-rate", type=float, default=200.0)
    ap.add_argument("--serve-small-steps", type=int, default=100)
    ap.add_argument("--large-n", type=int, default=256,
                    help="grid of the second FP64/mixed throughput figure (0: skip)")
    ap.add_argument("--large-steps", type=int, default=5)
    ap.add_argument("--no-autotune", action="store_true",
                    help="keep the default tiles (profiling runs: no tuning launches in the trace)")
    ap.add_argument("--study-steps", type=int, default=0,
                    help="instead of the throughput run: FP64 vs mixed on the --n case for this many "
                         "steps, per-field errors")
    a = ap.parse_args()
    if a.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {a.gpus} is not supported (multi-GPU sharding is "
              "design-only in this build)", file=sys.stderr)
        sys.exit(2)
    from sengapy import abi
    from sengapy import case as Cs
    if a.study_steps > 0:
        n = a.n
        t0 = time.perf_counter()
        c = Cs.h2_nscbc(n=(n, n, n), length=(5.0e-3 * (n - 1) / 127,) * 3, mixavg=True, nsteps=a.study_steps)
        print(f"case built in {time.perf_counter() - t0:.1f} s", file=sys.stderr, flush=True)
        out = {"study": "precision", "workload": f"H2-air 9-species NSCBC mixture-averaged {n}^3, grid spacing of "
                                                 f"the 128^3 / 5 mm case, {a.study_steps} RK steps",
               "grid": [n, n, n], "steps": a.study_steps, "tstep": c.dump.tstep}
        out.update(precision_study(abi, hip(), c, a.study_steps))
        print(json.dumps(out), flush=True)
        return
    bspecs = Cs.batch_specs(a.batch_cases) if a.batch_cases > 0 else []
    bsspecs = (Cs.batch_specs(a.batch_small_cases, seed=77, sizes=(16, 24, 32))
               if a.batch_small_cases > 0 else [])
    ssspecs = (Cs.kernel_specs(a.serve_small_requests, seed=99, rate=a.serve_small_rate, sizes=(16, 24, 32),
                               nsteps=a.serve_small_steps) if a.serve_small_requests > 0 else [])
    sspecs = Cs.kernel_specs(a.serve_requests, rate=a.serve_rate, nsteps=a.serve_steps) if a.serve_requests > 0 else []
    lspecs = ([dict(kind="nscbc", n=(a.large_n,) * 3, phi=0.7, uprime_over_sl=5.0, Tu=300.0, seed=7, nsteps=100)]
              if a.large_n > 0 else [])
    allspecs = lspecs + bspecs + bsspecs + sspecs + ssspecs
    built = make_cases(allspecs) if allspecs else []
    lcases, built = built[:len(lspecs)], built[len(lspecs):]
    bcases, built = built[:len(bspecs)], built[len(bspecs):]
    bscases, built = built[:len(bsspecs)], built[len(bsspecs):]
    scases, sscases = built[:len(sspecs)], built[len(sspecs):]
    n = a.n
    t0 = time.perf_counter()
    c = Cs.h2_nscbc(n=(n, n, n), length=(5.0e-3, 5.0e-3, 5.0e-3), mixavg=True, nsteps=a.steps)
    print(f"case built in {time.perf_counter() - t0:.1f} s, dt = {c.dump.tstep:.3e}", file=sys.stderr, flush=True)
    s = abi.Solver(c)
    s.set_timestep(c.dump.tstep)
    s.upload(c.dump.state, 0.0)
    tuned = (s.tiles(), None) if a.no_autotune else s.autotune(5)
    print(f"tiles {tuned}", file=sys.stderr, flush=True)
    h = hip()
    if a.copy_calibration:
        copy_calibration(h)
    stream = C.c_void_p(s.stream())
    ev0, ev1 = C.c_void_p(), C.c_void_p()
    h.hipEventCreate(C.byref(ev0))
    h.hipEventCreate(C.byref(ev1))
    s.step(a.warmup)
    s.sync()
    h.hipDeviceSynchronize()
    h.hipEventRecord(ev0, stream)
    s.step(a.steps)
    h.hipEventRecord(ev1, stream)
    h.hipEventSynchronize(ev1)
    s.sync()
    ms = C.c_float(0)
    h.hipEventElapsedTime(C.byref(ms), ev0, ev1)
    total_s = ms.value / 1e3
    npts = n ** 3
    value = npts * 5 * a.steps / total_s / 1e6
    print(f"timed {a.steps} steps: {ms.value:.2f} ms", file=sys.stderr, flush=True)
    if a.profile_steps < 1:
        print(json.dumps({"metric": "Mpoint-RK-substages/s", "value": round(npts * 5 * a.steps / total_s / 1e6, 3),
                          "steps": a.steps, "note": "no profiled steps"}), flush=True)
        return
    # per-kernel durations with events around each launch (separate, profiled steps)
    s.profile(True)
    s.step(a.profile_steps)
    s.sync()
    prof = s.profile_read()
    s.profile(False)
    ns = c.nspec
    kern = {k: v for k, v in prof.items() if v[1] > 0 and kernel_bytes(k, ns, True, n) > 0}
    table = kernel_table(kern, ns, npts, n, "fp64")
    dom = max(table, key=lambda k: table[k]["ms"])
    roof = roofline_of(table[dom], dom, npts)
    substage_bytes = 32 * (5 + ns) + 16      # compulsory state traffic of one fused substage
    sub_ms = ms.value / a.steps / 5
    floor_ms = sum(v["floor_ms"] for v in table.values())
    meas_ms = sum(v["ms"] for v in table.values())
    traffic = [v.get("traffic_bytes_per_point") for v in table.values()]
    substage = {
        "measured_ms": sub_ms,
        "kernel_ms_sum": meas_ms,
        "floor_ms": floor_ms,
        "floor_frac": floor_ms / sub_ms,
        "compulsory_bytes_per_point": substage_bytes,
        "compulsory_hbm_frac": substage_bytes * value * 1e6 / HBM_PEAK,
        "moved_bytes_per_point": sum(traffic) if all(t is not None for t in traffic) else None,
        "note": "floor_ms: sum over kernels of max(algorithmic bytes / 8 TB/s, counted flops / 78.6 TFLOP/s)",
    }
    if substage["moved_bytes_per_point"] is not None:
        substage["moved_hbm_frac"] = substage["moved_bytes_per_point"] * npts / (sub_ms / 1e3) / HBM_PEAK
    out = {
        "metric": "Mpoint-RK-substages/s",
        "value": round(value, 3),
        "unit": "Mpoint-RK-substages/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(ms.value / a.steps, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded turbulence + planar erf flame initial dump)",
        "config": {"workload": f"H2-air 9-species 21-step, {n}^3, NSCBC x inflow/outflow, periodic y/z, "
                               "mixture-averaged transport, FP64, RK4(3)5",
                   "grid": [n, n, n], "nspec": ns, "nstep_chem": c.tables.nstep, "tstep": c.dump.tstep},
        "roofline": roof,
        "substage": substage,
        "kernels": table,
        "kernel_ms_per_launch": {k: round(v["ms"], 4) for k, v in table.items()},
        "tiles": {"narrow_pass_tiles": list(tuned[0]), "autotuned": not a.no_autotune},
    }
    if not a.no_mixed:
        out["mixed"] = mixed_run(abi, c, a, h, s)
    s.close()
    if lcases:
        out["large"] = large_run(abi, h, lcases[0], a.large_steps)
        out["large"]["workload"] = (f"the same H2-air NSCBC physics on {a.large_n}^3 (grid spacing of the 128^3 "
                                    f"/ 5 mm case), {a.large_steps} RK steps timed after 2, tiles autotuned")
    if bcases:
        out["batched"] = batched_run(abi, h, bcases, a.batch_steps)
        out["batched"]["workload"] = (f"{len(bcases)} H2-air NSCBC cases, n^3 with n from {{48,64,96,128}}, "
                                      "phi~U(0.5,1.2), u'/S_L~U(1,10), T_u~U(300,700) K; "
                                      f"{a.batch_steps} RK steps timed, rates per 100-step case")
    if bscases:
        # where batching pays: cases too small to fill the GPU alone, whose
        # solo steps are bound by launch latency
        out["batched_small"] = batched_run(abi, h, bscases, a.batch_small_steps)
        out["batched_small"]["workload"] = (f"{len(bscases)} H2-air NSCBC cases, n^3 with n from {{16,24,32}}, "
                                            "phi~U(0.5,1.2), u'/S_L~U(1,10), T_u~U(300,700) K; "
                                            f"{a.batch_small_steps} RK steps timed, rates per 100-step case")
    if scases:
        out["serving"] = serve_run(scases, sspecs, a.serve_steps, a.serve_point_budget, a.serve_chunk)
        out["serving"]["workload"] = (f"{len(scases)} H2-air flame-kernel requests, Poisson {a.serve_rate}/s, "
                                      f"n^3 with n in 48..160, {a.serve_steps} RK steps each, "
                                      f"240 GB HBM budget, {a.serve_point_budget:.3g} points in flight, "
                                      f"{a.serve_chunk}-step chunks, time to solution = arrival to downloaded state")
    if sscases:
        out["serving_small"] = serve_run(sscases, ssspecs, a.serve_small_steps, a.serve_point_budget, a.serve_chunk)
        out["serving_small"]["workload"] = (f"{len(sscases)} H2-air flame-kernel requests, Poisson "
                                            f"{a.serve_small_rate}/s, n^3 with n in {{16,24,32}}, "
                                            f"{a.serve_small_steps} RK steps each, {a.serve_point_budget:.3g} "
                                            f"points in flight, {a.serve_chunk}-step chunks")
    if not a.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
