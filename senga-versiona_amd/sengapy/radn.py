"""radn.dat reader: the optically thin radiation data (src/radcin.f:76-126).

Layout: a five-line header, the species list (count, then index and symbol),
a label line, the number of radiating species NSPRAD and per radiating species
its index and coefficient count NKPRAD followed by the NKPRAD coefficients of
its Planck mean absorption coefficient, a polynomial in T (list-directed: the
coefficients may span lines).
"""
from __future__ import annotations

from dataclasses import dataclass
from typing import List

import numpy as np

NCFRMX = 6   # coefficients per radiating species (src/com_senga2.h:224-225)


@dataclass
class RadnData:
    nsprid: List[int]      # 0-based species index of each radiating species
    nkprad: List[int]      # coefficient count of each
    akprad: np.ndarray     # (NCFRMX, NSPRAD), AKPRAD(c, r)


def read_radn(path: str) -> RadnData:
    lines = open(path).read().split("\n")
    pos = 5
    pos += 1                                          # 'Species list:'
    nspec = int(lines[pos]); pos += 1
    pos += nspec
    pos += 1                                          # 'Planck mean absorption coefficient data'
    toks = " ".join(lines[pos:]).replace("D", "E").replace("d", "e").split()
    q = 0
    nsprad = int(toks[q]); q += 1
    ids, nks = [], []
    ak = np.zeros((NCFRMX, nsprad))
    for r in range(nsprad):
        j, nk = int(toks[q]), int(toks[q + 1]); q += 2
        if nk > NCFRMX:
            raise ValueError(f"{path}: {nk} coefficients for species {j}, at most {NCFRMX}")
        ak[:nk, r] = [float(x) for x in toks[q:q + nk]]; q += nk
        ids.append(j - 1)
        nks.append(nk)
    return RadnData(ids, nks, ak)
