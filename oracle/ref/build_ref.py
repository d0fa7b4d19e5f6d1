#!/usr/bin/env python3
"""Build the reference SENGA2 Fortran sources into a per-case oracle binary.

This is test infrastructure: the binary it produces is used only to write the
golden fixtures under tests/golden/ and to cross-check the C restatement in
oracle/.  Nothing in the product links or runs it.

The reference keeps its array sizes as compile-time PARAMETERs in
src/com_senga2.h (src/com_senga2.h:12-34,74); its users set them per case (the
casefiles ship one header each, e.g. src/casefiles/1step/com_senga2.h).  This
recipe does the same: it writes a header with the case's grid, species and step
counts (single rank: NXPROC=NYPROC=NZPROC=1) into oracle/_ref/<case>/src/ and
compiles the reference sources *in place* through symlinks placed beside that
header, so Fortran INCLUDE resolves to the case header while the source text is
read from /root/reference.  Outputs go only to oracle/_ref/<case>/.

src/flamin.f is replaced by the reference's own source-swap option
src/flamin_options/flamin_stub.f (cases here always restart from a dump).
The single-rank message-passing layer is the reference's own src/mpi_stub.
src/senga2.f (main), src/inflow.f and src/hdf5io.f (direct MPI/HDF5 calls, not on
the right-hand-side path) and src/finish.f are left out; the oracle's own driver
oracle/ref/refrun.f replaces the main program.

Compiler: ROCm flang, -O2 -cpp -ffixed-form -mcmodel=medium.  On x86-64 without
-march flags it emits no fused multiply-add, so the oracle binary evaluates the
Fortran expressions in plain IEEE double.

usage: build_ref.py NAME NX NY NZ NSPEC NSTEP
"""
import os
import re
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

REF = "/root/reference/src"
HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
FLANG = "/opt/rocm/lib/llvm/bin/flang"
FFLAGS = ["-O2", "-cpp", "-ffixed-form", "-mcmodel=medium"]

EXCLUDE = {"senga2.f", "inflow.f", "hdf5io.f", "finish.f"}


def reference_sources():
    """Source list of src/Makefile.codegen (SRC = ...) minus EXCLUDE."""
    text = open(os.path.join(REF, "Makefile.codegen")).read()
    m = re.search(r"^SRC\s*=(.*?)(?:\n\s*\n)", text, re.S | re.M)
    names = m.group(1).replace("\\", " ").split()
    return [n for n in names if n.endswith(".f") and n not in EXCLUDE]


def case_header(nx, ny, nz, nspec, nstep):
    """Reference header with the case's compile-time sizes substituted."""
    h = open(os.path.join(REF, "com_senga2.h")).read()
    nmax = max(nx, ny, nz)
    subs = [
        (r"PARAMETER\(NXGLBL=\d+, NYGLBL=\d+, NZGLBL=\d+\)",
         f"PARAMETER(NXGLBL={nx}, NYGLBL={ny}, NZGLBL={nz})"),
        (r"PARAMETER\(NGZMAX=NXGLBL\)", f"PARAMETER(NGZMAX={nmax})"),
        (r"PARAMETER\(NXPROC=\d+, NYPROC=\d+, NZPROC=\d+\)",
         "PARAMETER(NXPROC=1, NYPROC=1, NZPROC=1)"),
        (r"PARAMETER\(NXSIZE=\d+, NYSIZE=\d+, NZSIZE=\d+\)",
         f"PARAMETER(NXSIZE={nx}, NYSIZE={ny}, NZSIZE={nz})"),
        (r"PARAMETER\(NSZMAX=NXSIZE\)", f"PARAMETER(NSZMAX={nmax})"),
        (r"PARAMETER\(NSPCMX=\d+, NSTPMX=\d+\)",
         f"PARAMETER(NSPCMX={nspec}, NSTPMX={max(nstep, 1)})"),
    ]
    for pat, rep in subs:
        h, n = re.subn(pat, rep, h)
        if n != 1:
            raise RuntimeError(f"header substitution failed for {pat}")
    return h


def link(src, dst):
    if os.path.islink(dst) or os.path.exists(dst):
        os.remove(dst)
    os.symlink(src, dst)


def build(name, nx, ny, nz, nspec, nstep, jobs=8):
    out = os.path.join(ROOT, "oracle", "_ref", name)
    srcd = os.path.join(out, "src")
    objd = os.path.join(out, "obj")
    os.makedirs(srcd, exist_ok=True)
    os.makedirs(objd, exist_ok=True)
    hdr = os.path.join(srcd, "com_senga2.h")
    text = case_header(nx, ny, nz, nspec, nstep)
    if not os.path.exists(hdr) or open(hdr).read() != text:   # unchanged: objects stay valid
        with open(hdr, "w") as f:
            f.write(text)
    link(os.path.join(REF, "com_espect.h"), os.path.join(srcd, "com_espect.h"))
    link(os.path.join(REF, "mpi_stub", "mpif.h"), os.path.join(srcd, "mpif.h"))
    units = []
    for n in reference_sources():
        if n == "flamin.f":
            # the cold-start flame generator is hard-wired to the 9-species
            # H2 case; cases here always restart from a dump, so the
            # reference's own stub option is used (src/flamin_options/)
            units.append((os.path.join(REF, "flamin_options", "flamin_stub.f"), n))
        else:
            units.append((os.path.join(REF, n), n))
    for sub in ("fftlib", "rndlib", "mpi_stub"):
        d = os.path.join(REF, sub)
        for n in sorted(os.listdir(d)):
            if n.endswith(".f"):
                units.append((os.path.join(d, n), f"{sub}_{n}"))
    units.append((os.path.join(HERE, "refrun.f"), "refrun.f"))
    objs = []
    cmds = []
    for path, lname in units:
        lpath = os.path.join(srcd, lname)
        link(path, lpath)
        obj = os.path.join(objd, lname[:-2] + ".o")
        objs.append(obj)
        if (not os.path.exists(obj) or os.path.getmtime(obj) < os.path.getmtime(path)
                or os.path.getmtime(obj) < os.path.getmtime(os.path.join(srcd, "com_senga2.h"))):
            cmds.append([FLANG, *FFLAGS, "-c", lpath, "-o", obj])

    def run(cmd):
        r = subprocess.run(cmd, capture_output=True, text=True, cwd=objd)
        if r.returncode != 0:
            raise RuntimeError(f"compile failed: {' '.join(cmd)}\n{r.stderr[-3000:]}")

    with ThreadPoolExecutor(jobs) as ex:
        list(ex.map(run, cmds))
    exe = os.path.join(out, "refrun")
    r = subprocess.run([FLANG, *FFLAGS, *objs, "-o", exe], capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed\n{r.stderr[-3000:]}")
    return exe


if __name__ == "__main__":
    if len(sys.argv) != 7:
        sys.exit(__doc__)
    a = sys.argv[1:]
    print(build(a[0], *map(int, a[1:])))
