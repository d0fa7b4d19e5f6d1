"""ctypes binding of the C ABI in include/senga_gpu.h.

``Config`` mirrors ``sg_config`` field by field; ``config_from_case`` fills it
from a ``sengapy.case.Case`` (cont.dat control data plus the chem.dat/diff.dat
derived tables).  ``Library`` loads libsenga_gpu.so and wraps the entry points;
it raises if the library is missing (there is no CPU fallback).
"""
from __future__ import annotations

import ctypes as C
import math
import os

import numpy as np

MAX_SPEC = 18
MAX_STEP = 128
MAX_SSL = 10
MAX_BODY = 16
MAX_FALL = 32
NKPRAD = 6
NCOF = 7
NTINT = 2
NTRC = 4

PKG = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIBDIR = os.path.join(PKG, "lib")
LIBPATH = os.path.join(LIBDIR, "libsenga_gpu.so")

d = C.c_double
i32 = C.c_int32


class Config(C.Structure):
    _fields_ = [
        ("nx", i32), ("ny", i32), ("nz", i32), ("nspec", i32), ("nstep", i32),
        ("mixavg", i32), ("precision", i32), ("physics", i32),
        ("xgdlen", d), ("ygdlen", d), ("zgdlen", d),
        ("bc_type", i32 * 6), ("bc_ipar", (i32 * 4) * 6), ("bc_rpar", (d * 4) * 6),
        ("prin", d), ("trin", d), ("urin", d), ("vrin", d), ("wrin", d),
        ("yrin", d * MAX_SPEC),
        ("alamda", d), ("rlamda", d), ("prantl", d), ("olewis", d * MAX_SPEC),
        ("wmolar", d * MAX_SPEC), ("ovwmol", d * MAX_SPEC), ("rgspec", d * MAX_SPEC),
        ("ntint", i32 * MAX_SPEC),
        ("ncpoly", (i32 * NTINT) * MAX_SPEC), ("ncenth", (i32 * NTINT) * MAX_SPEC),
        ("ncenpy", (i32 * NTINT) * MAX_SPEC), ("tinthi", (d * NTINT) * MAX_SPEC),
        ("amascp", ((d * NCOF) * NTINT) * MAX_SPEC), ("amasct", ((d * NCOF) * NTINT) * MAX_SPEC),
        ("amasch", ((d * NCOF) * NTINT) * MAX_SPEC), ("amolgb", ((d * NCOF) * NTINT) * MAX_SPEC),
        ("rparam", (d * 3) * MAX_STEP),
        ("nsslen", i32 * MAX_STEP), ("nsspec", (i32 * MAX_SSL) * MAX_STEP),
        ("diffmu", (d * MAX_SSL) * MAX_STEP), ("diffmw", (d * MAX_SSL) * MAX_STEP),
        ("nrslen", i32 * MAX_STEP), ("nrspec", (i32 * MAX_SSL) * MAX_STEP),
        ("npslen", i32 * MAX_STEP), ("npspec", (i32 * MAX_SSL) * MAX_STEP),
        ("nrclen", i32 * MAX_STEP), ("nrcpec", (i32 * MAX_SSL) * MAX_STEP),
        ("crspec", (d * MAX_SSL) * MAX_STEP),
        ("npclen", i32 * MAX_STEP), ("npcpec", (i32 * MAX_SSL) * MAX_STEP),
        ("cpspec", (d * MAX_SSL) * MAX_STEP),
        ("mblist", i32 * MAX_STEP), ("mglist", i32 * MAX_STEP), ("mllist", i32 * MAX_STEP),
        ("mtlist", i32 * MAX_STEP), ("mslist", i32 * MAX_STEP),
        ("effy3b", (d * MAX_SPEC) * MAX_BODY),
        ("rclind", (d * 4) * MAX_FALL), ("rctroe", (d * 12) * MAX_FALL), ("rcsrif", (d * 8) * MAX_FALL),
        ("clnten", d),
        ("pdifgb", d), ("tdifgb", d),
        ("ncovis", i32), ("ncocon", i32), ("ncodif", i32), ("reserved1", i32),
        ("viscco", (d * NTRC) * MAX_SPEC), ("condco", (d * NTRC) * MAX_SPEC),
        ("diffco", ((d * NTRC) * MAX_SPEC) * MAX_SPEC),
        ("wilko1", (d * MAX_SPEC) * MAX_SPEC), ("wilko2", (d * MAX_SPEC) * MAX_SPEC),
        ("ncotdr", i32), ("reserved2", i32), ("wmltdr", d),
        ("tdrcco", ((d * NTRC) * MAX_SPEC) * MAX_SPEC),
        ("nsprad", i32), ("reserved3", i32), ("nsprid", i32 * MAX_SPEC), ("nkprad", i32 * MAX_SPEC),
        ("trefrn", d), ("akprad", (d * NKPRAD) * MAX_SPEC),
    ]


def fill_chemistry(c: Config, t) -> None:
    """The chemistry members of sg_config (step lists, rate parameters, third
    bodies, falloff constants) from derived tables ``t`` (sengapy.chem.derive)."""
    m = t.mech
    ns, nst = t.nspec, t.nstep
    for st in range(nst):
        c.rparam[st][0], c.rparam[st][1], c.rparam[st][2] = t.rparam[:, st]
        c.nsslen[st] = len(m.nsspec[st])
        for q, sp in enumerate(m.nsspec[st]):
            c.nsspec[st][q] = sp - 1
            c.diffmu[st][q] = t.diffmu[st][q]
            c.diffmw[st][q] = t.diffmw[st][q]
        c.nrslen[st] = len(m.nrspec[st])
        for q, sp in enumerate(m.nrspec[st]):
            c.nrspec[st][q] = sp - 1
        c.npslen[st] = len(m.npspec[st])
        for q, sp in enumerate(m.npspec[st]):
            c.npspec[st][q] = sp - 1
        c.nrclen[st] = len(m.nrcpec[st])
        for q, (sp, cf) in enumerate(m.nrcpec[st]):
            c.nrcpec[st][q] = sp - 1
            c.crspec[st][q] = cf
        c.npclen[st] = len(m.npcpec[st])
        for q, (sp, cf) in enumerate(m.npcpec[st]):
            c.npcpec[st][q] = sp - 1
            c.cpspec[st][q] = cf
        c.mblist[st] = m.mblist[st] if m.bodies else 0
        c.mglist[st] = m.mglist[st]
        c.mllist[st] = m.mllist[st]
        c.mtlist[st] = m.mtlist[st]
        c.mslist[st] = m.mslist[st]
    for b in range(len(m.bodies)):
        for s in range(ns):
            c.effy3b[b][s] = m.effy3b[s, b]
    for j in range(t.rclind.shape[1]):
        for q in range(4):
            c.rclind[j][q] = t.rclind[q, j]
    for j in range(t.rctroe.shape[1]):
        for q in range(12):
            c.rctroe[j][q] = t.rctroe[q, j]
    for j in range(t.rcsrif.shape[1]):
        for q in range(8):
            c.rcsrif[j][q] = t.rcsrif[q, j]
    c.clnten = math.log(1.0e1)          # src/indata.f:1071


# sg_config.physics bits (include/senga_gpu.h SG_PHYS_*)
PHYS_MIXW, PHYS_MIXP, PHYS_SORET, PHYS_DUFOUR, PHYS_RADN = 1, 2, 4, 8, 16


def physics_bits(ctl) -> int:
    """SG_PHYS_* bits of the optional physics switches of cont.dat that are on
    (src/contin.f:194-209): the molecular-transport options count only with
    mixture-averaged transport on, as src/diffin.f:66,384-407 applies them."""
    bits = 0
    if ctl.nfmavt == 1:
        bits |= PHYS_MIXW if ctl.nfmixw == 1 else 0
        bits |= PHYS_MIXP if ctl.nfmixp == 1 else 0
        bits |= PHYS_SORET if ctl.nfmsor == 1 else 0
        bits |= PHYS_DUFOUR if ctl.nfmduf == 1 else 0
    bits |= PHYS_RADN if ctl.nfradn == 1 else 0
    return bits


def config_from_case(case, precision=0) -> Config:
    """Fill sg_config from a Case (control + derived tables); precision 0 is
    FP64, 1 the mixed FP32/FP64 mode, 2 the mixed mode with the FP64 species
    renormalisation at the end of every step."""
    c = Config()
    ctl, t, m = case.control, case.tables, case.tables.mech
    ns, nst = t.nspec, t.nstep
    if ns > MAX_SPEC or nst > MAX_STEP:
        raise ValueError("mechanism too large for sg_config")
    c.nx, c.ny, c.nz = ctl.nx, ctl.ny, ctl.nz
    c.nspec, c.nstep = ns, nst
    c.mixavg = 1 if ctl.nfmavt == 1 else 0
    c.physics = physics_bits(ctl)
    c.precision = precision
    c.xgdlen, c.ygdlen, c.zgdlen = ctl.xgdlen, ctl.ygdlen, ctl.zgdlen
    for f, (typ, ip, rp) in enumerate(ctl.bc):
        c.bc_type[f] = typ
        for q in range(4):
            c.bc_ipar[f][q] = ip[q]
            c.bc_rpar[f][q] = rp[q]
    c.prin, c.trin, c.urin, c.vrin, c.wrin = ctl.prin, ctl.trin, ctl.urin, ctl.vrin, ctl.wrin
    for s in range(ns):
        c.yrin[s] = ctl.yrin[s] if s < len(ctl.yrin) else 0.0
    from .chem import RLAMDA, PRANTL
    c.alamda, c.rlamda, c.prantl = t.alamda, RLAMDA, PRANTL
    for s in range(ns):
        c.olewis[s] = t.olewis[s]
        c.wmolar[s], c.ovwmol[s], c.rgspec[s] = t.wmolar[s], t.ovwmol[s], t.rgspec[s]
        c.ntint[s] = int(t.ntint[s])
        for it in range(int(t.ntint[s])):
            c.ncpoly[s][it] = int(t.ncpoly[it, s])
            c.ncenth[s][it] = int(t.ncenth[it, s])
            c.ncenpy[s][it] = int(t.ncenpy[it, s])
            c.tinthi[s][it] = t.tinthi[it, s]
            for q in range(NCOF):
                c.amascp[s][it][q] = t.amascp[q, it, s]
                c.amasct[s][it][q] = t.amasct[q, it, s]
                c.amasch[s][it][q] = t.amasch[q, it, s]
                c.amolgb[s][it][q] = t.amolgb[q, it, s]
    fill_chemistry(c, t)
    dt = case.diff
    if dt is not None:
        c.pdifgb, c.tdifgb = dt.pdifgb, dt.tdifgb
        c.ncovis, c.ncocon, c.ncodif = dt.ncovis, dt.ncocon, dt.ncodif
        for s in range(ns):
            for q in range(dt.ncovis):
                c.viscco[s][q] = dt.viscco[q, s]
            for q in range(dt.ncocon):
                c.condco[s][q] = dt.condco[q, s]
            for j in range(ns):
                for q in range(dt.ncodif):
                    c.diffco[s][j][q] = dt.diffco[q, j, s]
                c.wilko1[s][j] = dt.wilko1[j, s]
                c.wilko2[s][j] = dt.wilko2[j, s]
                for q in range(dt.ncotdr):
                    c.tdrcco[s][j][q] = dt.tdrcco[q, j, s]
        c.ncotdr = dt.ncotdr
    c.wmltdr = ctl.wmltdr
    c.trefrn = ctl.trefrn
    if ctl.nfradn == 1:
        rd = case.radn
        if rd is None:
            raise ValueError("NFRADN = 1 needs the case's radn.dat")
        c.nsprad = len(rd.nsprid)
        for r, (k, nk) in enumerate(zip(rd.nsprid, rd.nkprad)):
            c.nsprid[r], c.nkprad[r] = k, nk
            for q in range(nk):
                c.akprad[r][q] = rd.akprad[q, r]
    return c


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


ABI_FUNCS = ["sg_abi_version", "sg_config_size", "sg_create", "sg_destroy", "sg_upload_state",
             "sg_download_state", "sg_set_timestep", "sg_substage", "sg_end_step", "sg_step",
             "sg_rhs", "sg_derivatives", "sg_sync", "sg_last_error", "sg_get_stream",
             "sg_profile", "sg_profile_read", "sg_memory", "sg_set_adaptive", "sg_get_timestep",
             "sg_batch_create", "sg_batch_destroy", "sg_batch_add", "sg_batch_remove", "sg_batch_size",
             "sg_batch_member", "sg_batch_step", "sg_batch_sync", "sg_footprint", "sg_set_tiles",
             "sg_get_tiles", "sg_autotune", "sg_batch_set_tiles", "sg_batch_get_tiles", "sg_batch_autotune",
             "sg_mechanism_compiled", "sg_derivatives_noalloc"]


def load_library(path=None):
    """Load libsenga_gpu.so; raises OSError if it is not built (no fallback)."""
    path = path or LIBPATH
    if not os.path.exists(path):
        raise OSError(f"{path} is not built; run senga-versiona_amd/build.py")
    lib = C.CDLL(path)
    P = C.POINTER
    dp = P(C.c_double)
    vp = C.c_void_p
    lib.sg_abi_version.restype = C.c_int
    lib.sg_config_size.restype = C.c_size_t
    lib.sg_create.argtypes = [P(Config), P(vp)]
    lib.sg_destroy.argtypes = [vp]
    lib.sg_upload_state.argtypes = [vp, dp, C.c_double]
    lib.sg_download_state.argtypes = [vp, dp, dp, dp, dp]
    lib.sg_set_timestep.argtypes = [vp, C.c_double]
    lib.sg_substage.argtypes = [vp, C.c_int]
    lib.sg_end_step.argtypes = [vp]
    lib.sg_step.argtypes = [vp, C.c_int]
    lib.sg_rhs.argtypes = [vp, C.c_int, dp]
    lib.sg_derivatives.argtypes = [vp, dp, dp]
    lib.sg_derivatives_noalloc.argtypes = [vp, dp, dp]
    lib.sg_sync.argtypes = [vp]
    lib.sg_last_error.argtypes = [vp, P(C.c_int), P(C.c_int), C.c_char_p, C.c_size_t]
    lib.sg_get_stream.argtypes = [vp, P(vp)]
    lib.sg_profile.argtypes = [vp, C.c_int]
    lib.sg_profile_read.argtypes = [vp, C.c_int, C.c_char_p, C.c_size_t, dp, P(C.c_int64)]
    lib.sg_memory.argtypes = [vp, P(C.c_size_t)]
    lib.sg_set_adaptive.argtypes = [vp, C.c_int, C.c_double, C.c_double]
    lib.sg_get_timestep.argtypes = [vp, dp, dp, dp]
    lib.sg_footprint.argtypes = [P(Config), P(C.c_size_t)]
    i3 = C.c_int * 3
    lib.sg_set_tiles.argtypes = [vp, i3]
    lib.sg_get_tiles.argtypes = [vp, i3]
    lib.sg_autotune.argtypes = [vp, C.c_int, dp]
    lib.sg_mechanism_compiled.argtypes = [vp, P(C.c_int)]
    lib.sg_batch_set_tiles.argtypes = [vp, i3]
    lib.sg_batch_get_tiles.argtypes = [vp, i3]
    lib.sg_batch_autotune.argtypes = [vp, C.c_int, dp]
    lib.sg_batch_create.argtypes = [P(vp)]
    lib.sg_batch_destroy.argtypes = [vp]
    lib.sg_batch_add.argtypes = [vp, vp]
    lib.sg_batch_remove.argtypes = [vp, vp]
    lib.sg_batch_size.argtypes = [vp, P(C.c_int)]
    lib.sg_batch_member.argtypes = [vp, C.c_int, P(vp)]
    lib.sg_batch_step.argtypes = [vp, C.c_int]
    lib.sg_batch_sync.argtypes = [vp, P(C.c_int)]
    for f in ABI_FUNCS:
        if f not in ("sg_config_size",):
            getattr(lib, f).restype = C.c_int
    return lib


def footprint(case, precision=0, lib=None):
    """Device bytes a Solver for this case allocates (sg_footprint; no device
    access)."""
    lib = lib or load_library()
    cfg = config_from_case(case, precision)
    b = C.c_size_t(0)
    rc = lib.sg_footprint(C.byref(cfg), C.byref(b))
    if rc:
        raise SgError(rc, "sg_footprint: configuration rejected")
    return b.value


class SgError(RuntimeError):
    def __init__(self, code, msg, ijk=(0, 0, 0)):
        super().__init__(f"{msg} (code {code}, point {tuple(ijk)})")
        self.code, self.ijk = code, tuple(ijk)


class Solver:
    """One case on the current GPU through the C ABI."""

    def __init__(self, case, lib=None, precision=0, cfg=None):
        """cfg: a Config to create from instead of config_from_case(case)."""
        self.lib = lib or load_library()
        if self.lib.sg_config_size() != C.sizeof(Config):
            raise RuntimeError("sg_config layout mismatch between library and binding")
        self.case = case
        self.cfg = cfg if cfg is not None else config_from_case(case, precision)
        h = C.c_void_p()
        rc = self.lib.sg_create(C.byref(self.cfg), C.byref(h))
        if rc != 0:
            raise SgError(rc, "sg_create failed")
        self.h = h
        self.nx, self.ny, self.nz = case.shape
        self.ns = case.nspec

    def close(self):
        if getattr(self, "h", None):
            self.lib.sg_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def _chk(self, rc):
        if rc != 0:
            code = C.c_int(0)
            ijk = (C.c_int * 3)()
            msg = C.create_string_buffer(256)
            self.lib.sg_last_error(self.h, C.byref(code), ijk, msg, 256)
            raise SgError(rc, msg.value.decode(), list(ijk))

    def upload(self, run, etime=0.0):
        run = np.ascontiguousarray(run, dtype=np.float64)
        self._chk(self.lib.sg_upload_state(self.h, _dp(run), etime))

    def set_timestep(self, dt):
        self._chk(self.lib.sg_set_timestep(self.h, dt))

    def set_adaptive(self, on, errold=1.0e-4, errldr=1.0e-4):
        """Adaptive time step (ADAPTT after each step); errold/errldr are the
        controller history from the dump."""
        self._chk(self.lib.sg_set_adaptive(self.h, int(on), errold, errldr))

    def timestep(self):
        v = [C.c_double() for _ in range(3)]
        self._chk(self.lib.sg_get_timestep(self.h, *[C.byref(x) for x in v]))
        return tuple(x.value for x in v)

    def substage(self, irk):
        self._chk(self.lib.sg_substage(self.h, irk))

    def end_step(self):
        self._chk(self.lib.sg_end_step(self.h))

    def step(self, n=1):
        self._chk(self.lib.sg_step(self.h, n))

    def sync(self):
        self._chk(self.lib.sg_sync(self.h))

    def rhs(self, irk=1):
        out = np.zeros((5 + self.ns, self.nz, self.ny, self.nx))
        self._chk(self.lib.sg_rhs(self.h, irk, _dp(out)))
        return out

    def download(self):
        run = np.zeros((5 + self.ns, self.nz, self.ny, self.nx))
        T = np.zeros((self.nz, self.ny, self.nx)); Pr = np.zeros_like(T)
        rate = np.zeros((self.ns, self.nz, self.ny, self.nx))
        self._chk(self.lib.sg_download_state(self.h, _dp(run), _dp(T), _dp(Pr), _dp(rate)))
        return run, T, Pr, rate

    def derivatives(self, f):
        f = np.ascontiguousarray(f, dtype=np.float64)
        out = np.zeros((9,) + f.shape)
        self._chk(self.lib.sg_derivatives(self.h, _dp(f), _dp(out)))
        return out

    def stream(self):
        s = C.c_void_p()
        self._chk(self.lib.sg_get_stream(self.h, C.byref(s)))
        return s.value

    def profile(self, on):
        self._chk(self.lib.sg_profile(self.h, 1 if on else 0))

    def profile_read(self):
        out = {}
        name = C.create_string_buffer(64)
        ms = C.c_double(0)
        cnt = C.c_int64(0)
        i = 0
        while self.lib.sg_profile_read(self.h, i, name, 64, C.byref(ms), C.byref(cnt)) == 0:
            out[name.value.decode()] = (ms.value, cnt.value)
            i += 1
        return out

    def memory(self):
        b = C.c_size_t(0)
        self._chk(self.lib.sg_memory(self.h, C.byref(b)))
        return b.value

    def mechanism_compiled(self) -> bool:
        """True when the chemistry is the mechanism compiled into the library
        (the point kernels then run its unrolled step loop)."""
        v = C.c_int(0)
        self._chk(self.lib.sg_mechanism_compiled(self.h, C.byref(v)))
        return bool(v.value)

    def set_tiles(self, narrow=(0, 0, 0)):
        """Pass tiles per direction (0: rows of 32, 1: rows of 16); results do
        not change."""
        self._chk(self.lib.sg_set_tiles(self.h, (C.c_int * 3)(*narrow)))

    def tiles(self):
        nw = (C.c_int * 3)()
        self._chk(self.lib.sg_get_tiles(self.h, nw))
        return tuple(nw)

    def autotune(self, reps=5):
        """Time the tile choices on the current state and keep the fastest;
        returns (narrow, ms per substage of the tuned passes)."""
        ms = C.c_double(0)
        self._chk(self.lib.sg_autotune(self.h, int(reps), C.byref(ms)))
        return self.tiles(), ms.value


class BatchError(SgError):
    def __init__(self, code, msg, member=-1, ijk=(0, 0, 0)):
        super().__init__(code, msg, ijk)
        self.member = member


class Batch:
    """Cases advanced together (sg_batch_*): one launch per substage kernel
    over all members, one captured graph per RK step."""

    def __init__(self, lib=None):
        self.lib = lib or load_library()
        b = C.c_void_p()
        rc = self.lib.sg_batch_create(C.byref(b))
        if rc != 0:
            raise SgError(rc, "sg_batch_create failed")
        self.b = b
        self.members = []

    def close(self):
        if getattr(self, "b", None):
            self.lib.sg_batch_destroy(self.b)
            self.b = None
            self.members = []

    def __del__(self):
        self.close()

    def add(self, solver):
        solver._chk(self.lib.sg_batch_add(self.b, solver.h))
        self.members.append(solver)

    def remove(self, solver):
        solver._chk(self.lib.sg_batch_remove(self.b, solver.h))
        self.members.remove(solver)

    def size(self):
        n = C.c_int(0)
        rc = self.lib.sg_batch_size(self.b, C.byref(n))
        if rc:
            raise SgError(rc, "sg_batch_size failed")
        return n.value

    def step(self, n=1):
        rc = self.lib.sg_batch_step(self.b, n)
        if rc:
            raise SgError(rc, "sg_batch_step failed")

    def set_tiles(self, narrow=(0, 0, 0)):
        rc = self.lib.sg_batch_set_tiles(self.b, (C.c_int * 3)(*narrow))
        if rc:
            raise SgError(rc, "sg_batch_set_tiles failed")

    def tiles(self):
        nw = (C.c_int * 3)()
        rc = self.lib.sg_batch_get_tiles(self.b, nw)
        if rc:
            raise SgError(rc, "sg_batch_get_tiles failed")
        return tuple(nw)

    def autotune(self, reps=5):
        ms = C.c_double(0)
        rc = self.lib.sg_batch_autotune(self.b, int(reps), C.byref(ms))
        if rc:
            raise SgError(rc, "sg_batch_autotune failed")
        return self.tiles(), ms.value

    def sync(self):
        failed = C.c_int(-1)
        rc = self.lib.sg_batch_sync(self.b, C.byref(failed))
        if rc:
            m = failed.value
            if 0 <= m < len(self.members):
                try:
                    self.members[m]._chk(rc)
                except SgError as e:
                    raise BatchError(rc, str(e), m, e.ijk) from None
            raise BatchError(rc, "sg_batch_sync failed", m)

