"""Run the reference SENGA2 build (oracle/_ref) on a case and parse its outputs.

Test infrastructure only: writes the case's input/ and output/ directories in
a scratch directory, runs oracle/_ref/<build>/refrun (driver oracle/ref/refrun.f)
and returns numpy arrays.  Arrays come back as (field, z, y, x) float64.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, HERE)
sys.path.insert(0, os.path.join(ROOT, "senga-versiona_amd"))

import build_ref  # noqa: E402


def build_name(case):
    nx, ny, nz = case.shape
    return f"n{nx}x{ny}x{nz}_s{case.nspec}_r{case.tables.nstep}"


def build_for(case):
    """The reference binary for this case's compile-time sizes: built from
    the reference sources where they are present, else the one built earlier
    (the build directory travels with the repository)."""
    nx, ny, nz = case.shape
    name = build_name(case)
    exe = os.path.join(ROOT, "oracle", "_ref", name, "refrun")
    if not os.path.isdir(build_ref.REF):
        if os.path.exists(exe):
            return exe
        raise FileNotFoundError(f"{exe} is not built and the reference sources are absent")
    return build_ref.build(name, nx, ny, nz, case.nspec, case.tables.nstep)


def run(case, mode, nsteps=None, opin=None, keep=None):
    """Mode 5 (steps with OUTPUT) also returns the run's output files:
    "rept" and "stat" (text of output/rept.res and output/stat.res) and
    "outres" ({file name: raw bytes} of output/out*.res)."""
    exe = build_for(case)
    d = tempfile.mkdtemp(prefix="refrun_")
    try:
        if nsteps is not None:
            case.control.ntime = nsteps
        case.write_inputs(d)
        open(os.path.join(d, "refrun.ctl"), "w").write(f"{mode}\nout.bin\n")
        if opin is not None:
            np.asarray(opin, "<f8").tofile(os.path.join(d, "opin.bin"))
        r = subprocess.run([exe], cwd=d, capture_output=True, text=True)
        if r.returncode != 0 or not os.path.exists(os.path.join(d, "out.bin")):
            raise RuntimeError(f"refrun failed ({r.returncode}):\n{r.stdout[-2000:]}\n{r.stderr[-2000:]}")
        raw = open(os.path.join(d, "out.bin"), "rb").read()
        files = {}
        if mode == 5:
            od = os.path.join(d, "output")
            files["rept"] = open(os.path.join(od, "rept.res")).read()
            files["stat"] = open(os.path.join(od, "stat.res")).read()
            files["outres"] = {f: open(os.path.join(od, f), "rb").read()
                               for f in sorted(os.listdir(od)) if f.startswith("out")}
        if keep:
            shutil.copy(os.path.join(d, "out.bin"), keep)
    finally:
        shutil.rmtree(d, ignore_errors=True)
    out = parse(raw)
    out.update(files)
    return out


def parse(raw):
    hdr = np.frombuffer(raw[:20], "<i4")
    nx, ny, nz, ns, mode = (int(v) for v in hdr)
    n = nx * ny * nz
    pos = 20
    out = {"shape": (nx, ny, nz), "nspec": ns, "mode": mode}

    def take(cnt, dt="<f8"):
        nonlocal pos
        sz = np.dtype(dt).itemsize * cnt
        a = np.frombuffer(raw[pos:pos + sz], dt)
        pos += sz
        return a

    out["T0"] = take(n).reshape(nz, ny, nx)
    out["itndex0"] = take(n, "<i4").reshape(nz, ny, nx)
    if mode in (1, 4):
        out["rhs"] = take(n * (5 + ns)).reshape(5 + ns, nz, ny, nx)
        out["T"] = take(n).reshape(nz, ny, nx)
        out["P"] = take(n).reshape(nz, ny, nx)
        out["rate"] = take(n * ns).reshape(ns, nz, ny, nx)
        if mode == 4:
            out["run"] = take(n * (5 + ns)).reshape(5 + ns, nz, ny, nx)
            out["stage"] = take(n * (5 + ns)).reshape(5 + ns, nz, ny, nx)
    elif mode in (2, 5):
        out["run"] = take(n * (5 + ns)).reshape(5 + ns, nz, ny, nx)
        out["T"] = take(n).reshape(nz, ny, nx)
        out["P"] = take(n).reshape(nz, ny, nx)
        out["rate"] = take(n * ns).reshape(ns, nz, ny, nx)
        out["etime"], out["tstep"] = take(2)
    elif mode == 3:
        out["ops"] = take(9 * n).reshape(9, nz, ny, nx)
    return out
