"""Reference states of FLEXI's Navier-Stokes equation module, on the host (numpy).

``ref_state_prim`` builds RefStatePrim(6,nRefState) from the parameter file's RefState entries
(rho, u, v, w, p) as InitEquation does (src/equations/navierstokes/equation.f90:126-141):
T = p*(1/rho)/R. ``prim_to_cons`` is PrimToCons (idealgas/eos.f90:305-328), the conservative
state of IniExactFunc = 1 (exactfunc.f90:258-259: RefStateCons of IniRefState);
``cons_to_prim`` is ConsToPrim (idealgas/eos.f90:203-230).
"""
from __future__ import annotations

import numpy as np


def ref_state_prim(ref_states, R: float) -> np.ndarray:
    """(6, nRefState) Fortran-ordered: rho, u, v, w, p, T for each RefState (rho, u, v, w, p)."""
    rs = np.asarray(ref_states, dtype=np.float64).reshape(-1, 5)
    out = np.zeros((6, rs.shape[0]), order="F")
    out[:5] = rs.T
    out[5] = out[4] * (1.0 / out[0]) / R
    return out


def prim_to_cons(prim: np.ndarray, kappa: float) -> np.ndarray:
    """PrimToCons over the leading axis: prim (>=5, ...) -> cons (5, ...)."""
    prim = np.asarray(prim, dtype=np.float64)
    c = np.empty((5,) + prim.shape[1:], order="F")
    c[0] = prim[0]
    c[1] = prim[1] * prim[0]
    c[2] = prim[2] * prim[0]
    c[3] = prim[3] * prim[0]
    c[4] = (1.0 / (kappa - 1.0)) * prim[4] + 0.5 * (c[1] * prim[1] + c[2] * prim[2] + c[3] * prim[3])
    return c


def cons_to_prim(cons: np.ndarray, kappa: float, R: float) -> np.ndarray:
    """ConsToPrim over the leading axis: cons (5, ...) -> prim (6, ...) = rho, u, v, w, p, T."""
    c = np.asarray(cons, dtype=np.float64)
    p = np.empty((6,) + c.shape[1:], order="F")
    s_rho = 1.0 / c[0]
    p[0] = c[0]
    p[1] = c[1] * s_rho
    p[2] = c[2] * s_rho
    p[3] = c[3] * s_rho
    p[4] = (kappa - 1.0) * (c[4] - 0.5 * (c[1] * p[1] + c[2] * p[2] + c[3] * p[3]))
    p[5] = p[4] * s_rho / R
    return p
