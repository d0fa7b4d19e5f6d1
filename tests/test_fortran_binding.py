"""The Fortran ISO_C_BINDING module (senga-versiona_amd/fortran/senga_gpu_mod.f90)
declares sg_config_t with the same size and member offsets as the C struct
sg_config of include/senga_gpu.h, and every bind(C) name it declares is
exported by the library header."""
import os
import re
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
FLANG = "/opt/rocm/lib/llvm/bin/flang"
MOD = os.path.join(ROOT, "senga-versiona_amd", "fortran", "senga_gpu_mod.f90")


def _parse(out):
    return dict((ln.split()[0], int(ln.split()[1])) for ln in out.strip().splitlines())


@pytest.mark.skipif(not os.path.exists(FLANG), reason="flang not present")
def test_fortran_config_layout_matches_c(tmp_path):
    fexe, cexe = tmp_path / "flc", tmp_path / "clc"
    r = subprocess.run([FLANG, "-O1", "-J", str(tmp_path), MOD, os.path.join(HERE, "native", "fortran_layout_check.f90"),
                        "-o", str(fexe)], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-3000:]
    r = subprocess.run(["gcc", "-O1", "-I", os.path.join(ROOT, "include"), os.path.join(HERE, "native", "c_layout_check.c"),
                        "-o", str(cexe)], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-3000:]
    f = _parse(subprocess.run([str(fexe)], capture_output=True, text=True, check=True).stdout)
    c = _parse(subprocess.run([str(cexe)], capture_output=True, text=True, check=True).stdout)
    assert f == c


def test_fortran_bind_names_are_exported():
    names = set(re.findall(r'bind\(C, name="(\w+)"\)', open(MOD).read()))
    header = open(os.path.join(ROOT, "include", "senga_gpu.h")).read()
    declared = set(re.findall(r"^\w[\w\s\*]*?\b(sg_\w+)\(", header, re.M))
    assert names and names <= declared, names - declared
