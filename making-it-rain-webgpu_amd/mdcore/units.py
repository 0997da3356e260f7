"""Units and Ewald/mesh parameters, OpenMM's everywhere (``src/engine/units.ts``)."""
from __future__ import annotations

import math

BOLTZMANN_KJ_PER_MOL_K = 0.008314462618153242
ONE_4PI_EPS0 = 138.93545764498457
BAR_TO_KJ_PER_MOL_NM3 = 0.06022140762081242
AVOGADRO = 6.02214076e23
FS_PER_PS = 1000.0
PS_PER_NS = 1000.0


def kT(temperatureK: float) -> float:
    return BOLTZMANN_KJ_PER_MOL_K * temperatureK


def ewaldAlpha(cutoffNm: float, tolerance: float) -> float:
    """``(1/cutoff) sqrt(-log(2 tol))`` (``units.ts:53-54``)."""
    return (1.0 / cutoffNm) * math.sqrt(-math.log(2.0 * tolerance))


def pmeMinimumGridSize(alphaPerNm: float, boxLengthNm: float, tolerance: float) -> int:
    """``ceil(2 alpha L / (3 tol^0.2))``, never below 6 (``units.ts:65-70``)."""
    return max(6, math.ceil((2 * alphaPerNm * boxLengthNm) / (3 * tolerance ** 0.2)))


def smoothFftSize(minimum: float) -> int:
    """Smallest 2/3/5/7-smooth integer >= minimum (``units.ts:78-84``)."""
    n = max(1, math.ceil(minimum))
    while True:
        r = n
        for p in (2, 3, 5, 7):
            while r % p == 0:
                r //= p
        if r == 1:
            return n
        n += 1
