"""cont.dat (run control) reader and writer.

Line layout follows src/contin.f:83-214 (list-directed reads after fixed
header/label lines).  Boundary types are the reference codes of
src/com_senga2.h (NSPERI=1, NSBCI3=13, NSBCO1=21, ...).
"""
from __future__ import annotations

from dataclasses import dataclass, field


def _d(x: float) -> str:
    s = f"{x:.16E}"
    m, e = s.split("E")
    return f"{m}D{int(e):+03d}"


@dataclass
class Control:
    xgdlen: float
    ygdlen: float
    zgdlen: float
    nx: int
    ny: int
    nz: int
    tstep: float
    ntime1: int = 1
    ntime: int = 1
    nstpsw: int = 0
    ntdump: int = 1000000
    ntrept: int = 1000000
    ntstat: int = 1000000
    ndofmt: int = 0
    ncdmpi: int = 1
    ndifmt: int = 0
    inturb: int = 0
    inseed: int = 1
    sparam: tuple = (0.0, 0.0, 0.0, 0.0)
    inflam: int = 0
    prin: float = 1.0e5
    trin: float = 300.0
    urin: float = 0.0
    vrin: float = 0.0
    wrin: float = 0.0
    yrin: list = field(default_factory=list)
    # per face: (type, [4 ints], [4 reals]) in order xl, xr, yl, yr, zl, zr
    bc: list = field(default_factory=lambda: [(1, [0, 0, 0, 0], [0.0] * 4)] * 6)
    nfmavt: int = 0
    nfmixw: int = 0
    nfmixp: int = 0
    nfmsor: int = 0
    nfmduf: int = 0
    wmltdr: float = 5.0
    nfradn: int = 0
    trefrn: float = 300.0

    def write(self, path: str) -> None:
        L = ["*" * 49, "**" + " " * 45 + "**", "**       SENGA2: Run control data file         **",
             "**" + " " * 45 + "**", "*" * 49, "",
             "Global domain size (x,y,z) in metres",
             f" {_d(self.xgdlen)} {_d(self.ygdlen)} {_d(self.zgdlen)}", "",
             "Global domain size (nx,ny,nz)", f" {self.nx} {self.ny} {self.nz}", "",
             "No. of processors (x,y,z)", " 1 1 1", "",
             "Time step; start step, no of steps, step switch (0=fixed, 1=adaptive))",
             f" {_d(self.tstep)} {self.ntime1} {self.ntime} {self.nstpsw}", "",
             "Intervals between dumps, reports, statistics, dump output (0/1=un/formatted)",
             f" {self.ntdump} {self.ntrept} {self.ntstat} {self.ndofmt}", "",
             "Cold start switch (0=cold start, 1=restart), dump input (0/1=un/formatted)",
             f" {self.ncdmpi} {self.ndifmt}", "",
             "Initial turbulence generator",
             "switch (0=off, 1=on); random seed; spectrum parameters",
             f" {self.inturb} {self.inseed} " + " ".join(_d(v) for v in self.sparam), "",
             "Flame generator switch (0=off, 1=on)", f" {self.inflam}", "",
             "Default initial conditions",
             "pressure, temperature, velocity components u,v,w",
             f" {_d(self.prin)} {_d(self.trin)} {_d(self.urin)} {_d(self.vrin)} {_d(self.wrin)}",
             "mass fractions", f" {len(self.yrin)}"]
        for i, y in enumerate(self.yrin):
            L.append(f" {i + 1} {_d(y)}")
        L += ["", "Global boundary condition types",
              "one per line: x-left; x-right; y-left; y-right; z-left; z-right",
              "(1=periodic; 1a=inlet; 2b=outlet; 3c=wall; a,b,c denotes BC subtype)",
              "(four integer and four real parameters allowed for each)"]
        for t, ip, rp in self.bc:
            L.append(f"{t} " + " ".join(str(i) for i in ip) + " " + " ".join(_d(r) for r in rp))
        L += ["", "Molecular transport control data",
              "switches: mix.av, mol.mass, pressure, Soret, Dufour; mol.mass limit",
              f"{self.nfmavt} {self.nfmixw} {self.nfmixp} {self.nfmsor} {self.nfmduf} {_d(self.wmltdr)}",
              "", "Radiation control data", "switch, reference temperature",
              f"{self.nfradn} {_d(self.trefrn)}", "", "End of file"]
        open(path, "w").write("\n".join(L) + "\n")


def _num(t: str):
    t = t.replace("D", "E").replace("d", "e")
    try:
        return int(t)
    except ValueError:
        return float(t)


def read_cont(path: str) -> Control:
    """Parse cont.dat as src/contin.f:83-214 does."""
    lines = open(path).read().split("\n")
    p = [6]

    def skip(n=1):
        p[0] += n

    def vals():
        v = [_num(t) for t in lines[p[0]].split()]
        p[0] += 1
        return v

    skip(); xg, yg, zg = vals(); skip()
    skip(); nx, ny, nz = vals()[:3]; skip()
    skip(); vals(); skip()
    skip(); tstep, ntime1, ntime, nstpsw = vals()[:4]; skip()
    skip(); ntdump, ntrept, ntstat, ndofmt = vals()[:4]; skip()
    skip(); ncdmpi, ndifmt = vals()[:2]; skip()
    skip(); skip(); t = vals(); skip()
    skip(); inflam = vals()[0]; skip()
    skip(); skip(); prin, trin, urin, vrin, wrin = vals()[:5]; skip()
    nsp = vals()[0]
    yrin = [0.0] * nsp
    for _ in range(nsp):
        v = vals(); yrin[int(v[0]) - 1] = float(v[1])
    skip()
    skip(4)
    bc = []
    for _ in range(6):
        v = vals()
        bc.append((int(v[0]), [int(x) for x in v[1:5]], [float(x) for x in v[5:9]]))
    skip()
    c = Control(float(xg), float(yg), float(zg), int(nx), int(ny), int(nz), float(tstep),
                int(ntime1), int(ntime), int(nstpsw), int(ntdump), int(ntrept), int(ntstat),
                int(ndofmt), int(ncdmpi), int(ndifmt), int(t[0]), int(t[1]),
                tuple(float(x) for x in t[2:6]), int(inflam), float(prin), float(trin),
                float(urin), float(vrin), float(wrin), yrin, bc)
    if lines[p[0]].startswith("Molecular transport control data"):
        skip(2)
        v = vals()
        c.nfmavt, c.nfmixw, c.nfmixp, c.nfmsor, c.nfmduf = (int(x) for x in v[:5])
        c.wmltdr = float(v[5])
        skip(3)
        v = vals()
        c.nfradn, c.trefrn = int(v[0]), float(v[1])
    return c
