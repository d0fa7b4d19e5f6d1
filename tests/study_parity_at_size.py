#!/usr/bin/env python3
"""Parity at the measured workload's own size and step count (needs the GPU).

The 128^3 H2-air NSCBC mixture-averaged case (the throughput workload,
sengapy.case.h2_nscbc) is advanced for STEPS RK steps by the FP64 GPU path,
by the mixed-precision GPU path (without and with the FP64 species
renormalisation, precision 1 and 2) and by the oracle, all from the same
initial dump and time step.  For every conserved field, T and p the script reports the
max-norm relative error (max|a-b| / max|b|), the L2 relative error and the
pointwise relative error with a floor of 1e-6 of the field's maximum, of both
GPU states against the oracle.  The oracle runs on the host's cores (OpenMP)
and prints a line per step.

Not collected by pytest (a study of minutes, not a unit test); the per-step
bars it checks are the tests' own: FP64 <= 1e-10, mixed <= 1e-5 max-norm.

usage: study_parity_at_size.py [--n 128] [--steps 100] [--json OUT]
"""
import argparse
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
for p in (os.path.join(ROOT, "senga-versiona_amd"), os.path.join(ROOT, "oracle"), HERE):
    if p not in sys.path:
        sys.path.insert(0, p)

from conftest import pw_err, rel_err  # noqa: E402


def l2_err(a, b):
    d = a - b
    return float(np.sqrt(np.sum(d * d)) / max(float(np.sqrt(np.sum(b * b))), 1e-300))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--n", type=int, default=128)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--json", default=None)
    a = ap.parse_args()
    from pyoracle import Oracle
    from sengapy import case as C
    from sengapy.abi import Solver
    n = a.n
    t0 = time.perf_counter()
    c = C.h2_nscbc(n=(n, n, n), length=(5.0e-3 * (n - 1) / 127,) * 3, mixavg=True, nsteps=a.steps)
    print(f"case {c.name} built in {time.perf_counter() - t0:.1f} s, dt {c.dump.tstep:.4e}", flush=True)
    states = {}
    for tag, prec in (("fp64", 0), ("mixed", 1), ("mixed_renorm", 2)):
        s = Solver(c, precision=prec)
        s.set_timestep(c.dump.tstep)
        s.upload(c.dump.state, 0.0)
        s.step(a.steps)
        run, T, P, _ = s.download()
        s.close()
        states[tag] = (run, T, P)
        print(f"gpu {tag}: {a.steps} steps done", flush=True)
    o = Oracle(c)
    print(f"oracle threads {o.threads()}", flush=True)
    o.set_timestep(c.dump.tstep)
    o.set_state(c.dump.state, 0.0)
    t0 = time.perf_counter()
    for k in range(a.steps):
        o.step(1)
        print(f"oracle step {k + 1}/{a.steps}, {time.perf_counter() - t0:.1f} s", flush=True)
    run_o, T_o, P_o, _ = o.get()
    names = ["rho", "rho_u", "rho_v", "rho_w", "rho_E"] + [f"rho_Y_{sp}" for sp in c.mech.species]
    out = {"study": "parity at size", "workload": f"H2-air 9-species 21-step NSCBC mixture-averaged {n}^3, "
                                                  f"{a.steps} RK steps, dt {c.dump.tstep:.6e} s",
           "grid": [n, n, n], "steps": a.steps, "oracle_threads": o.threads(),
           "oracle_s": time.perf_counter() - t0, "pointwise_floor": 1e-6, "errors": {}}
    for tag in ("fp64", "mixed", "mixed_renorm"):
        run, T, P = states[tag]
        e = {}
        for v, nm in enumerate(names):
            e[nm] = {"max": rel_err(run[v], run_o[v]), "l2": l2_err(run[v], run_o[v]),
                     "pointwise": pw_err(run[v], run_o[v])}
        e["T"] = {"max": rel_err(T, T_o), "l2": l2_err(T, T_o), "pointwise": pw_err(T, T_o)}
        e["p"] = {"max": rel_err(P, P_o), "l2": l2_err(P, P_o), "pointwise": pw_err(P, P_o)}
        out["errors"][tag] = e
        worst = max(v["max"] for v in e.values())
        print(f"{tag} vs oracle after {a.steps} steps: worst max-norm {worst:.2e}", flush=True)
        for nm, v in e.items():
            print(f"  {tag:5s} {nm:10s} max {v['max']:.2e}  l2 {v['l2']:.2e}  pointwise {v['pointwise']:.2e}",
                  flush=True)
    ok64 = max(v["max"] for v in out["errors"]["fp64"].values()) <= 1e-10
    okmx = max(max(v["max"] for v in out["errors"][t].values()) for t in ("mixed", "mixed_renorm")) <= 1e-5
    out["fp64_within_1e-10"] = ok64
    out["mixed_within_1e-5"] = okmx
    out["species_sum_drift"] = {t: float(np.abs(states[t][0][5:].sum(axis=0) / states[t][0][0] - 1.0).max())
                                for t in states}
    out["species_sum_drift"]["oracle"] = float(np.abs(run_o[5:].sum(axis=0) / run_o[0] - 1.0).max())
    print("max |sum rho Y_k / rho - 1|:", out["species_sum_drift"], flush=True)
    if a.json:
        with open(a.json, "w") as f:
            json.dump(out, f, indent=1)
    print(json.dumps({k: out[k] for k in ("workload", "fp64_within_1e-10", "mixed_within_1e-5", "oracle_s")}),
          flush=True)
    sys.exit(0 if (ok64 and okmx) else 1)


if __name__ == "__main__":
    main()
