"""Python handle on the C restatement in oracle/ (test infrastructure only).

Used by tests/, the smoke check and bench.py's cpu_baseline leg as the parity
checker; the product never imports it.
"""
from __future__ import annotations

import ctypes as C
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
sys.path.insert(0, os.path.join(ROOT, "senga-versiona_amd"))

from sengapy.abi import Config, config_from_case, _dp  # noqa: E402

LIBPATH = os.path.join(HERE, "lib", "liboracle.so")


class Oracle:
    """The C restatement (oracle/lib/liboracle.so) -- test infrastructure."""

    def __init__(self, case, libpath=None):
        libpath = libpath or LIBPATH
        lib = C.CDLL(libpath)
        lib.orc_create.restype = C.c_void_p
        lib.orc_create.argtypes = [C.POINTER(Config), C.POINTER(C.c_int)]
        for n in ("orc_set_state",):
            getattr(lib, n).argtypes = [C.c_void_p, C.POINTER(C.c_double), C.c_double]
        lib.orc_set_timestep.argtypes = [C.c_void_p, C.c_double]
        lib.orc_rhs.argtypes = [C.c_void_p, C.c_int, C.POINTER(C.c_double)]
        lib.orc_substage.argtypes = [C.c_void_p, C.c_int]
        lib.orc_end_step.argtypes = [C.c_void_p]
        lib.orc_step.argtypes = [C.c_void_p, C.c_int]
        lib.orc_get.argtypes = [C.c_void_p] + [C.POINTER(C.c_double)] * 4
        lib.orc_get_stage.argtypes = [C.c_void_p, C.POINTER(C.c_double)]
        lib.orc_get_itndex.argtypes = [C.c_void_p, C.POINTER(C.c_int)]
        lib.orc_derivatives.argtypes = [C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_double)]
        lib.orc_destroy.argtypes = [C.c_void_p]
        lib.orc_set_adaptive.argtypes = [C.c_void_p, C.c_int, C.c_double, C.c_double]
        lib.orc_get_timestep.argtypes = [C.c_void_p] + [C.POINTER(C.c_double)] * 3
        lib.orc_threads.argtypes = [C.c_int]
        lib.orc_threads.restype = C.c_int
        self.lib = lib
        self.case = case
        self.cfg = config_from_case(case)
        err = C.c_int(0)
        self.h = lib.orc_create(C.byref(self.cfg), C.byref(err))
        if not self.h:
            raise RuntimeError(f"orc_create failed: {err.value}")
        self.nx, self.ny, self.nz = case.shape
        self.ns = case.nspec

    def __del__(self):
        if getattr(self, "h", None):
            self.lib.orc_destroy(self.h)
            self.h = None

    def _chk(self, rc, what):
        if rc != 0:
            raise RuntimeError(f"{what} failed: {rc}")

    def threads(self, n=0):
        """OpenMP threads of the oracle's point loops: n > 0 sets the count
        (process-wide); returns the count in use."""
        return int(self.lib.orc_threads(int(n)))

    def set_state(self, run, etime=0.0):
        run = np.ascontiguousarray(run, dtype=np.float64)
        self._chk(self.lib.orc_set_state(self.h, _dp(run), etime), "orc_set_state")

    def set_timestep(self, dt):
        self.lib.orc_set_timestep(self.h, dt)

    def set_adaptive(self, on, errold=1.0e-4, errldr=1.0e-4):
        self.lib.orc_set_adaptive(self.h, int(on), errold, errldr)

    def timestep(self):
        v = [C.c_double() for _ in range(3)]
        self.lib.orc_get_timestep(self.h, *[C.byref(x) for x in v])
        return tuple(x.value for x in v)

    def rhs(self, irk=1):
        out = np.zeros((5 + self.ns, self.nz, self.ny, self.nx))
        self._chk(self.lib.orc_rhs(self.h, irk, _dp(out)), "orc_rhs")
        return out

    def substage(self, irk):
        self._chk(self.lib.orc_substage(self.h, irk), "orc_substage")

    def step(self, n=1):
        self._chk(self.lib.orc_step(self.h, n), "orc_step")

    def get(self):
        run = np.zeros((5 + self.ns, self.nz, self.ny, self.nx))
        T = np.zeros((self.nz, self.ny, self.nx)); P = np.zeros_like(T)
        rate = np.zeros((self.ns, self.nz, self.ny, self.nx))
        self.lib.orc_get(self.h, _dp(run), _dp(T), _dp(P), _dp(rate))
        return run, T, P, rate

    def stage(self):
        st = np.zeros((5 + self.ns, self.nz, self.ny, self.nx))
        self.lib.orc_get_stage(self.h, _dp(st))
        return st

    def itndex(self):
        a = np.zeros((self.nz, self.ny, self.nx), dtype=np.int32)
        self.lib.orc_get_itndex(self.h, a.ctypes.data_as(C.POINTER(C.c_int)))
        return a

    def derivatives(self, f):
        f = np.ascontiguousarray(f, dtype=np.float64)
        out = np.zeros((9,) + f.shape)
        self.lib.orc_derivatives(self.h, _dp(f), _dp(out))
        return out
