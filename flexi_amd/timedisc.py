"""Low-storage Runge-Kutta coefficients and CFL/DFL scaling of FLEXI's explicit time integration.

The tables are FLEXI's (src/timedisc/timedisc_vars.f90:93-214); they define the scheme and the
admissible time step, so they are reproduced value for value.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np


@dataclass(frozen=True)
class LSERK:
    name: str
    RKA: np.ndarray  # RKA(2:nStages)
    RKb: np.ndarray  # RKb(1:nStages)
    RKc: np.ndarray  # RKc(2:nStages)
    cfl_alpha_gauss: tuple
    cfl_alpha_lobatto: tuple
    relative_dfl: float

    @property
    def nStages(self) -> int:
        return len(self.RKb)


CARPENTER_RK4_5 = LSERK(
    name="CarpenterRK4-5",
    RKA=np.array([567301805773.0 / 1357537059087.0, 2404267990393.0 / 2016746695238.0,
                  3550918686646.0 / 2091501179385.0, 1275806237668.0 / 842570457699.0]),
    RKb=np.array([1432997174477.0 / 9575080441755.0, 5161836677717.0 / 13612068292357.0,
                  1720146321549.0 / 2090206949498.0, 3134564353537.0 / 4481467310338.0,
                  2277821191437.0 / 14882151754819.0]),
    RKc=np.array([1432997174477.0 / 9575080441755.0, 2526269341429.0 / 6820363962896.0,
                  2006345519317.0 / 3224310063776.0, 2802321613138.0 / 2924317926251.0]),
    cfl_alpha_gauss=(2.0351, 1.7595, 1.5401, 1.3702, 1.2375, 1.1318, 1.0440, 0.9709, 0.9079, 0.8539, 0.8066,
                     0.7650, 0.7290, 0.6952, 0.6660),
    cfl_alpha_lobatto=(4.7497, 3.4144, 2.8451, 2.4739, 2.2027, 1.9912, 1.8225, 1.6830, 1.5682, 1.4692, 1.3849,
                       1.3106, 1.2454, 1.1880, 1.1362),
    relative_dfl=1.853,
)

STANDARD_RK3_3 = LSERK(
    name="StandardRK3-3",
    RKA=np.array([5.0 / 9.0, 153.0 / 128.0]),
    RKb=np.array([1.0 / 3.0, 15.0 / 16.0, 8.0 / 15.0]),
    RKc=np.array([1.0 / 3.0, 0.75]),
    cfl_alpha_gauss=(1.2285, 1.0485, 0.9101, 0.8066, 0.7268, 0.6626, 0.6109, 0.5670, 0.5299, 0.4973, 0.4703,
                     0.4455, 0.4230, 0.4039, 0.3859),
    cfl_alpha_lobatto=(3.1871, 2.2444, 1.7797, 1.5075, 1.3230, 1.1857, 1.0800, 0.9945, 0.9247, 0.8651, 0.8134,
                       0.7695, 0.7301, 0.6952, 0.6649),
    relative_dfl=1.0,
)

SCHEMES = {"carpenterrk4-5": CARPENTER_RK4_5, "standardrk3-3": STANDARD_RK3_3}

# DFLScaleAlpha (timedisc_vars.f90:93-98) for the diagonal mass matrix
_DFL_ALPHA_GAUSS = (1.12, 0.76, 0.55, 0.41, 0.32, 0.25, 0.20, 0.17, 0.14, 0.12)
_DFL_ALPHA_LOBATTO = (4.50, 1.95, 1.18, 0.79, 0.56, 0.42, 0.32, 0.25, 0.20, 0.17)


def fill_cfl_dfl(N: int, node_type: int, CFLScale: float, DFLScale: float, scheme: LSERK = CARPENTER_RK4_5):
    """FillCFL_DFL (src/timedisc/timedisc_func.f90:463-527): scaled CFL and DFL numbers."""
    table = scheme.cfl_alpha_gauss if node_type == 1 else scheme.cfl_alpha_lobatto
    cfl = CFLScale * table[min(15, N) - 1]
    cfl = cfl / (2.0 * N + 1.0)
    alpha = (_DFL_ALPHA_GAUSS if node_type == 1 else _DFL_ALPHA_LOBATTO)[min(10, N) - 1] * scheme.relative_dfl
    dfl = DFLScale * alpha
    dfl = dfl / (2.0 * N + 1.0) ** 2
    return cfl, dfl
