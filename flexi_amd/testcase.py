"""Initial states of the Taylor-Green vortex test case and the other synthetic states of the
benchmark configurations.

TGV: ExactFuncTestcase/InitTestcase (src/testcase/taylorgreenvortex/testcase.f90:71-219) with
rho0 = U0 = 1, constant initial density, p0 = (U0/Ma)^2 rho0/kappa.
"""
from __future__ import annotations

import numpy as np


def prim_to_cons(prim: np.ndarray, kappa: float) -> np.ndarray:
    """PrimToCons (idealgas/eos.f90:305-328); prim(6,...) = (rho,u,v,w,p,T) -> cons(5,...)."""
    rho, u, v, w, p = prim[0], prim[1], prim[2], prim[3], prim[4]
    cons = np.empty((5,) + prim.shape[1:], order="F")
    cons[0] = rho
    cons[1] = u * rho
    cons[2] = v * rho
    cons[3] = w * rho
    cons[4] = (1.0 / (kappa - 1.0)) * p + 0.5 * (cons[1] * u + cons[2] * v + cons[3] * w)
    return cons


def tgv(xyz: np.ndarray, kappa: float = 1.4, R: float = 71.42857, Ma: float = 0.1) -> np.ndarray:
    """Conservative TGV state U(5,...) at points xyz(3,...)."""
    rho0, U0 = 1.0, 1.0
    p0 = (U0 / Ma) ** 2 / kappa * rho0
    x, y, z = xyz[0], xyz[1], xyz[2]
    prim = np.empty((6,) + xyz.shape[1:], order="F")
    prim[1] = U0 * np.sin(x) * np.cos(y) * np.cos(z)
    prim[2] = -U0 * np.cos(x) * np.sin(y) * np.cos(z)
    prim[3] = 0.0
    prim[4] = p0 + (rho0 * U0 ** 2) / 16.0 * (np.cos(2 * x) * np.cos(2.0 * z) + 2.0 * np.cos(2.0 * y)
                                              + 2.0 * np.cos(2.0 * x) + np.cos(2 * y) * np.cos(2.0 * z))
    prim[0] = rho0
    prim[5] = prim[4] / (prim[0] * R)
    return np.asfortranarray(prim_to_cons(prim, kappa))


def splitmix64_uniform(index: np.ndarray, seed: int = 1) -> np.ndarray:
    """Counter-based U[-1,1) draws: splitmix64 output for the state seed + (index+1)*golden, top 53 bits."""
    with np.errstate(over="ignore"):
        golden = np.uint64(0x9E3779B97F4A7C15)
        z = np.uint64(seed) + (np.asarray(index, dtype=np.uint64) + np.uint64(1)) * golden
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    u = (z >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)
    return 2.0 * u - 1.0


def tgv_density_perturbed(xyz: np.ndarray, amplitude: float = 1e-3, seed: int = 1, kappa: float = 1.4,
                          R: float = 71.42857, Ma: float = 0.1) -> np.ndarray:
    """TGV with rho += amplitude*xi, xi ~ U[-1,1) per node keyed by the global node index; p kept."""
    U = tgv(xyz, kappa, R, Ma)
    idx = np.arange(np.prod(xyz.shape[1:]), dtype=np.uint64).reshape(xyz.shape[1:], order="F")
    rho = 1.0 + amplitude * splitmix64_uniform(idx, seed)
    u, v, w = U[1] / U[0], U[2] / U[0], U[3] / U[0]
    p = (kappa - 1.0) * (U[4] - 0.5 * (U[1] * u + U[2] * v + U[3] * w))
    prim = np.stack([rho, u, v, w, p, p / (rho * R)])
    return np.asfortranarray(prim_to_cons(prim, kappa))


def freestream(shape, kappa: float = 1.4, rho=1.0, vel=(0.3, 0.2, 0.1), p=None) -> np.ndarray:
    """Uniform state U(5,shape...) with p = 1/kappa by default."""
    p = 1.0 / kappa if p is None else p
    prim = np.empty((6,) + tuple(shape), order="F")
    prim[0] = rho
    prim[1], prim[2], prim[3] = vel
    prim[4] = p
    prim[5] = 0.0
    return np.asfortranarray(prim_to_cons(prim, kappa))


def gaussian_pulse(xyz: np.ndarray, amplitude: float = 0.1, width: float = 0.5, kappa: float = 1.4,
                   vel=(0.3, 0.2, 0.1)) -> np.ndarray:
    """Freestream with rho = 1 + a*exp(-|x|^2/width^2) centred at the origin, p = 1/kappa."""
    r2 = xyz[0] ** 2 + xyz[1] ** 2 + xyz[2] ** 2
    prim = np.empty((6,) + xyz.shape[1:], order="F")
    prim[0] = 1.0 + amplitude * np.exp(-r2 / width ** 2)
    prim[1], prim[2], prim[3] = vel
    prim[4] = 1.0 / kappa
    prim[5] = 0.0
    return np.asfortranarray(prim_to_cons(prim, kappa))
