"""chem.dat reader/writer and the derived thermochemical tables.

The file layout is the one the reference reads in src/chemin.f:155-344 (fixed
Fortran edit descriptors, listed beside each field below).  The derived tables
follow src/chemin.f:822-983 operation by operation, in float64, so that the
values handed to the GPU library are bit-identical to what the reference
computes from the same file.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

RGUNIV = 8.3142e3          # src/com_senga2.h:118-119
ALAMDC, RLAMDA, TLAMDA = 2.58e-5, 0.7, 298.0   # src/com_senga2.h:112-113
PRANTL = 0.7               # src/com_senga2.h:115


def _ffloat(s: str, scale: int = 0) -> float:
    """Fortran E/F input of one fixed-width field (scale factor kP applies only
    when the field carries no exponent)."""
    t = s.strip().replace("D", "E").replace("d", "e")
    if t == "":
        return 0.0
    v = float(t)
    if scale and ("E" not in t.upper()):
        v *= 10.0 ** (-scale)
    return v


def _fields(line: str, widths):
    out, pos = [], 0
    for w in widths:
        out.append(line[pos:pos + w])
        pos += w
    return out


@dataclass
class Mechanism:
    species: list
    prefgb: float
    wmolar: np.ndarray          # (N,)
    clewis: np.ndarray          # (N,)
    ntint: np.ndarray           # (N,)
    tintlo: np.ndarray          # (NT, N)
    tinthi: np.ndarray          # (NT, N)
    ncofcp: np.ndarray          # (NT, N)
    amolcp: np.ndarray          # (7, NT, N) raw NASA coefficients (a1..a7)
    rparam: np.ndarray          # (3, S) A, n, E as read
    nsslen: list = field(default_factory=list)
    nsspec: list = field(default_factory=list)   # per step list of species (1-based)
    nrspec: list = field(default_factory=list)
    npspec: list = field(default_factory=list)
    nrcpec: list = field(default_factory=list)   # (species, coef) pairs
    npcpec: list = field(default_factory=list)
    diffmu: list = field(default_factory=list)   # per step list of delta-mu
    bodies: list = field(default_factory=list)
    mblist: list = field(default_factory=list)
    effy3b: np.ndarray = None   # (N, NB)
    mglist: list = field(default_factory=list)
    mllist: list = field(default_factory=list)
    rclind: np.ndarray = None   # (4, NL)
    mtlist: list = field(default_factory=list)
    rctroe: np.ndarray = None   # (12, NT)
    mslist: list = field(default_factory=list)
    rcsrif: np.ndarray = None   # (8, NS)

    @property
    def nspec(self):
        return len(self.species)

    @property
    def nstep(self):
        return self.rparam.shape[1]


def read_chem(path: str) -> Mechanism:
    """Parse chem.dat as src/chemin.f:155-344 does."""
    lines = open(path).read().split("\n")
    it = iter(lines)
    nxt = lambda: next(it)
    for _ in range(5):
        nxt()
    nxt()                                            # 'Species list:'
    nspec = int(nxt()[:5])                           # (I5)
    species = [""] * nspec
    for _ in range(nspec):
        ln = nxt()
        species[int(ln[:5]) - 1] = ln[8:].strip()    # (I5,3X,A)
    nxt()                                            # 'Species data:'
    prefgb = _ffloat(nxt())                          # list-directed
    nt = 2
    wm = np.zeros(nspec); cl = np.zeros(nspec); ntint = np.zeros(nspec, int)
    tlo = np.zeros((nt, nspec)); thi = np.zeros((nt, nspec)); ncof = np.zeros((nt, nspec), int)
    amolcp = np.zeros((7, nt, nspec))
    for _ in range(nspec):
        a, b = _fields(nxt(), (5, 12))               # (I5,1PE12.4)
        isp = int(a) - 1
        wm[isp] = _ffloat(b, 1)
        cl[isp] = _ffloat(nxt()[:12], 1)             # (1PE12.4)
        ntint[isp] = int(nxt()[:5])                  # (I5)
        for itn in range(ntint[isp]):
            a, b, c = _fields(nxt(), (12, 12, 5))    # (2(1PE12.4),I5)
            tlo[itn, isp] = _ffloat(a, 1); thi[itn, isp] = _ffloat(b, 1); ncof[itn, isp] = int(c)
            for icp in range(ncof[itn, isp]):
                amolcp[icp, itn, isp] = _ffloat(nxt()[:15], 1)   # (1PE15.7)
    nxt()                                            # 'Step rate data:'
    nstep = int(nxt()[:5])
    rparam = np.zeros((3, nstep))
    for _ in range(nstep):
        f = _fields(nxt(), (5, 12, 12, 12))          # (I5,3(1PE12.4))
        ist = int(f[0]) - 1
        rparam[:, ist] = [_ffloat(x, 1) for x in f[1:]]

    def steplist():
        nxt()
        lens, lists = [0] * nstep, [[] for _ in range(nstep)]
        for _ in range(nstep):
            f = _fields(nxt(), (5, 5))               # (2I5)
            ist = int(f[0]) - 1
            lens[ist] = int(f[1])
            tmp = [0] * lens[ist]
            for _ in range(lens[ist]):
                g = _fields(nxt(), (5, 5, 8))        # (2I5,I8)
                tmp[int(g[1]) - 1] = int(g[2])
            lists[ist] = tmp
        return lists

    nsspec = steplist()
    nrspec = steplist()
    npspec = steplist()

    def coeflist():
        nxt()
        lists = [[] for _ in range(nstep)]
        for _ in range(nstep):
            f = _fields(nxt(), (5, 5))
            ist = int(f[0]) - 1
            n = int(f[1])
            tmp = [None] * n
            for _ in range(n):
                g = _fields(nxt(), (5, 5, 8, 12))    # (2I5,I8,1PE12.4)
                tmp[int(g[1]) - 1] = (int(g[2]), _ffloat(g[3], 1))
            lists[ist] = tmp
        return lists

    nrcpec = coeflist()
    npcpec = coeflist()
    nxt()                                            # 'Species delta-list:'
    diffmu = [[0.0] * len(nsspec[s]) for s in range(nstep)]
    for s in range(nstep):
        for _ in range(len(nsspec[s])):
            g = _fields(nxt(), (5, 5, 5))            # (2I5,F5.1)
            diffmu[int(g[0]) - 1][int(g[1]) - 1] = _ffloat(g[2])
    nxt()                                            # 'Third-body list:'
    nbody = int(nxt()[:5])
    bodies = [""] * nbody
    for _ in range(nbody):
        ln = nxt()
        bodies[int(ln[:5]) - 1] = ln[8:].strip()
    nxt()                                            # 'Third-body step list:'
    mblist = [0] * nstep
    if nbody > 0:
        for _ in range(nstep):
            g = _fields(nxt(), (5, 8))               # (I5,I8)
            mblist[int(g[0]) - 1] = int(g[1])
    nxt()                                            # 'Third-body efficiencies:'
    effy = np.zeros((nspec, max(nbody, 1)))
    for _ in range(nbody):
        for _ in range(nspec):
            g = _fields(nxt(), (5, 5, 12))           # (2I5,1PE12.4)
            effy[int(g[1]) - 1, int(g[0]) - 1] = _ffloat(g[2], 1)

    def flaglist():
        nxt()
        n = int(nxt()[:5])
        lst = [0] * nstep
        if n > 0:
            for _ in range(nstep):
                g = _fields(nxt(), (5, 8))
                lst[int(g[0]) - 1] = int(g[1])
        return n, lst

    ngibb, mglist = flaglist()
    nlind, mllist = flaglist()
    nxt()                                            # 'Lindemann step rate data:'
    rclind = np.zeros((4, max(nlind, 1)))
    for _ in range(nlind):
        f = _fields(nxt(), (5, 12, 12, 12, 12))      # (I5,4(1PE12.4))
        rclind[:, int(f[0]) - 1] = [_ffloat(x, 1) for x in f[1:]]
    ntroe, mtlist = flaglist()
    nxt()
    rctroe = np.zeros((12, max(ntroe, 1)))
    for _ in range(ntroe):
        f = _fields(nxt(), (5,) + (12,) * 6)         # (I5,6(1PE12.4)) x2
        rctroe[0:6, int(f[0]) - 1] = [_ffloat(x, 1) for x in f[1:]]
        f = _fields(nxt(), (5,) + (12,) * 6)
        rctroe[6:12, int(f[0]) - 1] = [_ffloat(x, 1) for x in f[1:]]
    nsrif, mslist = flaglist()
    nxt()
    rcsrif = np.zeros((8, max(nsrif, 1)))
    for _ in range(nsrif):
        f = _fields(nxt(), (5, 12, 12, 12))          # (I5,3(1PE12.4))
        rcsrif[0:3, int(f[0]) - 1] = [_ffloat(x, 1) for x in f[1:]]
        f = _fields(nxt(), (5,) + (12,) * 5)         # (I5,5(1PE12.4))
        rcsrif[3:8, int(f[0]) - 1] = [_ffloat(x, 1) for x in f[1:]]
    return Mechanism(species, prefgb, wm, cl, ntint, tlo, thi, ncof, amolcp, rparam,
                     [len(x) for x in nsspec], nsspec, nrspec, npspec, nrcpec, npcpec,
                     diffmu, bodies, mblist, effy, mglist, mllist, rclind[:, :nlind],
                     mtlist, rctroe[:, :ntroe], mslist, rcsrif[:, :nsrif])


def _e(x: float, w: int = 12, d: int = 4) -> str:
    """Fortran 1PEw.d output."""
    s = f"{x:.{d}E}"
    mant, ex = s.split("E")
    return f"{mant}E{int(ex):+03d}".rjust(w)


def write_chem(m: Mechanism, path: str, title: str = "chem data file") -> None:
    """Write chem.dat in the layout src/chemin.f:155-344 reads."""
    L = [" " + "*" * 29, " *" + " " * 27 + "*", " *  " + title[:24].ljust(24) + " *",
         " *" + " " * 27 + "*", " " + "*" * 29, " Species list:", f"{m.nspec:5d}"]
    for i, s in enumerate(m.species):
        L.append(f"{i + 1:5d}   {s:<10s}")
    L.append(" Species data:")
    L.append(_e(m.prefgb))
    for i in range(m.nspec):
        L.append(f"{i + 1:5d}{_e(m.wmolar[i])}")
        L.append(_e(m.clewis[i]))
        L.append(f"{m.ntint[i]:5d}")
        for t in range(m.ntint[i]):
            L.append(f"{_e(m.tintlo[t, i])}{_e(m.tinthi[t, i])}{m.ncofcp[t, i]:5d}")
            for c in range(m.ncofcp[t, i]):
                L.append(_e(m.amolcp[c, t, i], 15, 7))
    L.append(" Step rate data:")
    L.append(f"{m.nstep:5d}")
    for s in range(m.nstep):
        L.append(f"{s + 1:5d}" + "".join(_e(v) for v in m.rparam[:, s]))
    for name, lists in (("Step species-list:", m.nsspec), ("Step reactant-list:", m.nrspec),
                        ("Step product-list:", m.npspec)):
        L.append(" " + name)
        for s in range(m.nstep):
            L.append(f"{s + 1:5d}{len(lists[s]):5d}")
            for j, sp in enumerate(lists[s]):
                L.append(f"{s + 1:5d}{j + 1:5d}{sp:8d}")
    for name, lists in (("Step reactant coefficient-list:", m.nrcpec),
                        ("Step product coefficient-list:", m.npcpec)):
        L.append(" " + name)
        for s in range(m.nstep):
            L.append(f"{s + 1:5d}{len(lists[s]):5d}")
            for j, (sp, c) in enumerate(lists[s]):
                L.append(f"{s + 1:5d}{j + 1:5d}{sp:8d}{_e(c)}")
    L.append(" Species delta-list:")
    for s in range(m.nstep):
        for j, v in enumerate(m.diffmu[s]):
            L.append(f"{s + 1:5d}{j + 1:5d}{v:5.1f}")
    L.append(" Third-body list:")
    L.append(f"{len(m.bodies):5d}")
    for i, b in enumerate(m.bodies):
        L.append(f"{i + 1:5d}   {b:<10s}")
    L.append(" Third-body step list:")
    if m.bodies:
        for s in range(m.nstep):
            L.append(f"{s + 1:5d}{m.mblist[s]:8d}")
    L.append(" Third-body efficiencies:")
    for b in range(len(m.bodies)):
        for i in range(m.nspec):
            L.append(f"{b + 1:5d}{i + 1:5d}{_e(m.effy3b[i, b])}")
    for name, n, lst in (("Gibbs step-list:", sum(1 for x in m.mglist if x), m.mglist),
                         ("Lindemann step-list:", m.rclind.shape[1], m.mllist)):
        L.append(" " + name)
        L.append(f"{n:5d}")
        if n:
            for s in range(m.nstep):
                L.append(f"{s + 1:5d}{lst[s]:8d}")
        if name.startswith("Lindemann"):
            L.append(" Lindemann step rate data:")
            for j in range(m.rclind.shape[1]):
                L.append(f"{j + 1:5d}" + "".join(_e(v) for v in m.rclind[:, j]))
    L.append(" Troe step-list:")
    L.append(f"{m.rctroe.shape[1]:5d}")
    if m.rctroe.shape[1]:
        for s in range(m.nstep):
            L.append(f"{s + 1:5d}{m.mtlist[s]:8d}")
    L.append(" Troe step rate data:")
    for j in range(m.rctroe.shape[1]):
        L.append(f"{j + 1:5d}" + "".join(_e(v) for v in m.rctroe[0:6, j]))
        L.append(f"{j + 1:5d}" + "".join(_e(v) for v in m.rctroe[6:12, j]))
    L.append(" SRI step-list:")
    L.append(f"{m.rcsrif.shape[1]:5d}")
    if m.rcsrif.shape[1]:
        for s in range(m.nstep):
            L.append(f"{s + 1:5d}{m.mslist[s]:8d}")
    L.append(" SRI step rate data:")
    for j in range(m.rcsrif.shape[1]):
        L.append(f"{j + 1:5d}" + "".join(_e(v) for v in m.rcsrif[0:3, j]))
        L.append(f"{j + 1:5d}" + "".join(_e(v) for v in m.rcsrif[3:8, j]))
    L.append(" End of file")
    open(path, "w").write("\n".join(L) + "\n")


@dataclass
class ChemTables:
    """Derived tables, src/chemin.f:822-983 (arrays indexed [coef, interval, species])."""
    nspec: int
    nstep: int
    wmolar: np.ndarray
    ovwmol: np.ndarray
    rgspec: np.ndarray
    olewis: np.ndarray
    ntint: np.ndarray
    tinthi: np.ndarray
    ncpoly: np.ndarray
    ncenth: np.ndarray
    ncenpy: np.ndarray
    amascp: np.ndarray
    amasct: np.ndarray
    amasch: np.ndarray
    amolgb: np.ndarray
    rparam: np.ndarray
    rclind: np.ndarray
    rctroe: np.ndarray
    rcsrif: np.ndarray
    diffmu: list
    diffmw: list
    alamda: float
    mech: Mechanism


def derive(m: Mechanism) -> ChemTables:
    """src/chemin.f:822-983, in the reference's operation order."""
    ns, nst = m.nspec, m.nstep
    rparam = m.rparam.copy()
    for s in range(nst):
        rparam[0, s] = math.log(rparam[0, s])
        rparam[2, s] = rparam[2, s] / RGUNIV
    rclind = m.rclind.copy()
    for j in range(rclind.shape[1]):
        rclind[0, j] = math.log(rclind[0, j]); rclind[2, j] = rclind[2, j] / RGUNIV
    rctroe = m.rctroe.copy()
    for j in range(rctroe.shape[1]):
        rctroe[0, j] = math.log(rctroe[0, j]); rctroe[2, j] = rctroe[2, j] / RGUNIV
        rctroe[4, j] = -1.0 / rctroe[4, j]; rctroe[6, j] = -1.0 / rctroe[6, j]
    rcsrif = m.rcsrif.copy()
    for j in range(rcsrif.shape[1]):
        rcsrif[0, j] = math.log(rcsrif[0, j]); rcsrif[2, j] = rcsrif[2, j] / RGUNIV
        rcsrif[4, j] = -rcsrif[4, j]; rcsrif[5, j] = -1.0 / rcsrif[5, j]
    diffmw = [[m.diffmu[s][j] * m.wmolar[m.nsspec[s][j] - 1] for j in range(len(m.nsspec[s]))]
              for s in range(nst)]
    ovwmol = np.array([1.0 / w for w in m.wmolar])
    rgspec = np.array([RGUNIV * o for o in ovwmol])
    nt = m.tintlo.shape[0]
    ncpoly = np.zeros((nt, ns), int); ncenth = np.zeros((nt, ns), int); ncenpy = np.zeros((nt, ns), int)
    amascp = np.zeros((7, nt, ns)); amasct = np.zeros((7, nt, ns))
    amasch = np.zeros((7, nt, ns)); amolgb = np.zeros((7, nt, ns))
    giblet = math.log(m.prefgb / RGUNIV)
    for i in range(ns):
        for t in range(m.ntint[i]):
            nc = int(m.ncofcp[t, i])
            ncpoly[t, i] = nc - 2; ncenth[t, i] = nc - 1; ncenpy[t, i] = nc
            for c in range(nc):
                amascp[c, t, i] = m.amolcp[c, t, i] * RGUNIV * ovwmol[i]
            amasct[0, t, i] = amascp[0, t, i] - rgspec[i]
            for c in range(2, nc - 1):
                amasct[c - 1, t, i] = amascp[c - 1, t, i] / float(c)
            amasct[nc - 2, t, i] = 0.0; amasct[nc - 1, t, i] = 0.0
            amasch[0, t, i] = amascp[0, t, i]
            for c in range(2, nc - 1):
                amasch[c - 1, t, i] = amascp[c - 1, t, i] / float(c)
            amasch[nc - 2, t, i] = amascp[nc - 2, t, i]; amasch[nc - 1, t, i] = 0.0
            amolgb[0, t, i] = m.amolcp[nc - 1, t, i] - m.amolcp[0, t, i] + giblet
            for c in range(2, nc - 1):
                amolgb[c - 1, t, i] = m.amolcp[c - 1, t, i] / float(c * (c - 1))
            amolgb[nc - 2, t, i] = m.amolcp[nc - 2, t, i]
            amolgb[nc - 1, t, i] = m.amolcp[0, t, i] - 1.0
    olewis = np.array([1.0 / c for c in m.clewis])
    alamda = ALAMDC * math.exp(-RLAMDA * math.log(TLAMDA))
    return ChemTables(ns, nst, m.wmolar.copy(), ovwmol, rgspec, olewis, m.ntint.copy(),
                      m.tinthi.copy(), ncpoly, ncenth, ncenpy, amascp, amasct, amasch, amolgb,
                      rparam, rclind, rctroe, rcsrif, m.diffmu, diffmw, alamda, m)
