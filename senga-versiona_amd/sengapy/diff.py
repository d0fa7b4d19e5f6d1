"""diff.dat reader and the derived mixture-averaged transport tables.

Layout as read by src/diffin.f:95-212 (list-directed); derived quantities as
src/diffin.f:280-405 (symmetric binary-diffusion matrix, antisymmetric
thermal-diffusion-ratio matrix, Wilke factors).
"""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np


def _toks(line):
    return [t.replace("D", "E").replace("d", "e") for t in line.split()]


@dataclass
class DiffTables:
    pdifgb: float
    tdifgb: float
    ncovis: int
    ncocon: int
    ncodif: int
    viscco: np.ndarray     # (NCOVIS, N)
    condco: np.ndarray     # (NCOCON, N)
    diffco: np.ndarray     # (NCODIF, N, N)  [coef, j, i], symmetric in (j, i)
    wilko1: np.ndarray     # (N, N) [j, i]
    wilko2: np.ndarray     # (N, N) [j, i]
    ncotdr: int = 0
    tdrcco: np.ndarray = None   # (NCOTDR, N, N) [coef, j, i], antisymmetric in (j, i)


def read_diff(path: str, wmolar) -> DiffTables:
    lines = open(path).read().split("\n")
    pos = 5
    pos += 1                                         # 'Species list:'
    nspec = int(lines[pos]); pos += 1
    pos += nspec
    pos += 1                                         # 'Molecular transport reference data:'
    t = _toks(lines[pos]); pos += 1
    pdifgb, tdifgb = float(t[0]), float(t[1])
    pos += 1                                         # 'Viscosity'
    viscco = None
    for _ in range(nspec):
        j, ncovis = map(int, lines[pos].split()); pos += 1
        if viscco is None:
            viscco = np.zeros((ncovis, nspec))
        viscco[:, j - 1] = [float(x) for x in _toks(lines[pos])[:ncovis]]; pos += 1
    pos += 1                                         # 'Thermal conductivity'
    condco = None
    for _ in range(nspec):
        j, ncocon = map(int, lines[pos].split()); pos += 1
        if condco is None:
            condco = np.zeros((ncocon, nspec))
        condco[:, j - 1] = [float(x) for x in _toks(lines[pos])[:ncocon]]; pos += 1
    pos += 1                                         # 'Binary diffusion coefficients'
    diffco = None
    for _ in range(nspec):
        i, ncodif = map(int, lines[pos].split()); pos += 1
        if diffco is None:
            diffco = np.zeros((ncodif, nspec, nspec))
        for _ in range(i):
            t = _toks(lines[pos]); pos += 1
            diffco[:, int(t[0]) - 1, i - 1] = [float(x) for x in t[1:1 + ncodif]]
    # symmetric fill, src/diffin.f:296-305
    for i in range(nspec):
        for j in range(i + 1, nspec):
            diffco[:, j, i] = diffco[:, i, j]
    # thermal diffusion ratios (src/diffin.f:180-195): species i lists j < i
    pos += 1                                         # 'Thermal diffusion ratio'
    tdrcco = None
    for _ in range(nspec):
        i, ncotdr = map(int, lines[pos].split()); pos += 1
        if tdrcco is None:
            tdrcco = np.zeros((ncotdr, nspec, nspec))
        for _ in range(i - 1):
            t = _toks(lines[pos]); pos += 1
            tdrcco[:, int(t[0]) - 1, i - 1] = [float(x) for x in t[1:1 + ncotdr]]
    # antisymmetric fill in the reference's order (src/diffin.f:345-368)
    for i in range(nspec):
        tdrcco[:, i, i] = 0.0
        for j in range(i):
            tdrcco[:, j, i] = -tdrcco[:, j, i]
        for j in range(i + 1, nspec):
            tdrcco[:, j, i] = tdrcco[:, i, j]
    w1 = np.zeros((nspec, nspec)); w2 = np.zeros((nspec, nspec))
    for i in range(nspec):
        for j in range(nspec):
            f = wmolar[i] / wmolar[j]
            w2[j, i] = 1.0 / math.sqrt(math.sqrt(f))
            f = 8.0 * (1.0 + f)
            w1[j, i] = 1.0 / math.sqrt(f)
    return DiffTables(pdifgb, tdifgb, viscco.shape[0], condco.shape[0], diffco.shape[0],
                      viscco, condco, diffco, w1, w2, tdrcco.shape[0], tdrcco)
