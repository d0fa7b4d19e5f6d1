"""Build the MI355X library (lib/libsenga_gpu.so) and, for tests, the oracle.

hipcc cross-compiles for gfx950 without a GPU.  Every supported species count
is one object of csrc/sg_kernels.hip (-DSG_NS=N), compiled in parallel.
"""
from __future__ import annotations

import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
OBJ = os.path.join(PKG, "_build")
LIB = os.path.join(PKG, "lib", "libsenga_gpu.so")
HIPCC = "/opt/rocm/bin/hipcc"
SPECIES = list(range(2, 19))
# Packed FP32 instructions are disabled for all device code: some of their forms have
# given wrong results on this hardware, and unpacked FP32 keeps the arithmetic exact.
HIPFLAGS = ["--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC",
            "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
            "-I", os.path.join(ROOT, "include"), "-I", CSRC]


def _run(cmd, cwd=None):
    r = subprocess.run(cmd, capture_output=True, text=True, cwd=cwd)
    if r.returncode != 0:
        raise RuntimeError(f"command failed: {' '.join(cmd)}\n{r.stdout[-4000:]}\n{r.stderr[-4000:]}")
    return r


def _stale(obj, deps):
    if not os.path.exists(obj):
        return True
    t = os.path.getmtime(obj)
    return any(os.path.getmtime(d) > t for d in deps)


def build_library(jobs=8, force=False, species=None):
    # the compiled mechanism header follows data/h2_9sp/chem.dat
    sys.path.insert(0, PKG)
    import mechgen
    if mechgen.main(["--check"]) != 0:
        mechgen.main([])
    os.makedirs(OBJ, exist_ok=True)
    os.makedirs(os.path.dirname(LIB), exist_ok=True)
    headers = [os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith(".h")]
    headers.append(os.path.join(ROOT, "include", "senga_gpu.h"))
    kern = os.path.join(CSRC, "sg_kernels.hip")
    units = []
    for n in (species or SPECIES):
        units.append((kern, os.path.join(OBJ, f"sg_kernels_ns{n}.o"), [f"-DSG_NS={n}"]))
    units.append((kern, os.path.join(OBJ, "sg_kernels_derivs.o"), ["-DSG_DERIVS"]))
    units.append((os.path.join(CSRC, "sg_api.hip"), os.path.join(OBJ, "sg_api.o"), []))
    units.append((os.path.join(CSRC, "sg_dispatch.cpp"), os.path.join(OBJ, "sg_dispatch.o"), ["-x", "hip"]))
    cmds = []
    for src, obj, extra in units:
        if force or _stale(obj, [src] + headers):
            cmds.append([HIPCC, *HIPFLAGS, *extra, "-c", src, "-o", obj])
    with ThreadPoolExecutor(jobs) as ex:
        list(ex.map(_run, cmds))
    objs = [u[1] for u in units]
    _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", *objs, "-o", LIB])
    return LIB


def build_oracle():
    """oracle/lib/liboracle.so: the C restatement (test infrastructure)."""
    src = os.path.join(ROOT, "oracle", "senga_oracle.c")
    out = os.path.join(ROOT, "oracle", "lib", "liboracle.so")
    os.makedirs(os.path.dirname(out), exist_ok=True)
    _run(["gcc", "-O2", "-ffp-contract=off", "-fopenmp", "-fPIC", "-shared", "-I", os.path.join(ROOT, "include"),
          src, "-o", out, "-lm"])
    return out



def build_fortran_driver():
    """The Fortran driver (ISO_C_BINDING) when flang and its sources exist."""
    fdir = os.path.join(PKG, "fortran")
    flang = "/opt/rocm/lib/llvm/bin/flang"
    srcs = [os.path.join(fdir, f) for f in ("senga_gpu_mod.f90", "senga_io.f90", "senga_gpu_driver.f90")]
    if not os.path.exists(flang) or not all(os.path.exists(s) for s in srcs):
        return None
    os.makedirs(OBJ, exist_ok=True)
    exe = os.path.join(PKG, "lib", "senga_gpu")
    _run([flang, "-O2", "-J", OBJ, *srcs, "-o", exe, "-L", os.path.dirname(LIB), "-lsenga_gpu",
          f"-Wl,-rpath,$ORIGIN", "-Wl,-rpath," + os.path.dirname(LIB)])
    return exe


if __name__ == "__main__":
    force = "--force" in sys.argv
    print(build_oracle())
    print(build_library(force=force))
    print(build_fortran_driver())
