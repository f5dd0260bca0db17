"""The Taylor-Green vortex testcase's analysis on the host, and FLEXI's CSV output of it.

``tgv_analyze`` is AnalyzeTestcase (src/testcase/taylorgreenvortex/testcase.f90:241-479): the
state, the lifted velocity gradients and the Jacobian are interpolated to NAnalyze+1
Gauss-Lobatto points per direction (ChangeBasis3D), where the kinetic energies, enstrophy,
dissipation rates, maximum vorticity, mean temperature and entropy are integrated and normalised
by the domain volume. ``TGVAnalysisFile`` writes them as FLEXI's ``<ProjectName>_TGVAnalysis.csv``
(InitOutputToFile / OutputToFile, src/output/output.f90:647-869): a header line, then one line per
analysis in the Fortran format E23.14E5 separated by ", ".
"""
from __future__ import annotations

import math
import os

import numpy as np

from .basis import GAUSS_LOBATTO, nodes_weights

TGV_VARNAMES = ["Dissipation Rate Incompressible", "Dissipation Rate Compressible", "Ekin incomp", "Ekin comp",
                "Enstrophy comp", "DR_u", "DR_S", "DR_Sd", "DR_p", "Maximum Vorticity", "Mean Temperature", "uprime",
                "Mean Entropy", "ED_S", "ED_D"]


def lagrange_matrix(x_from: np.ndarray, x_to: np.ndarray) -> np.ndarray:
    """V(len(x_to), len(x_from)): the Lagrange polynomials of the nodes x_from at the points x_to
    (barycentric form, LagrangeInterpolationPolys, src/interpolation/basis.f90:679-716)."""
    n = len(x_from)
    wb = np.array([1.0 / np.prod([x_from[j] - x_from[k] for k in range(n) if k != j]) for j in range(n)])
    V = np.zeros((len(x_to), n))
    for a, x in enumerate(x_to):
        hit = np.nonzero(np.isclose(x, x_from, rtol=0.0, atol=1e-14))[0]
        if hit.size:
            V[a, hit[0]] = 1.0
            continue
        t = wb / (x - x_from)
        V[a] = t / t.sum()
    return V


def tgv_analyze(U: np.ndarray, gx: np.ndarray, gy: np.ndarray, gz: np.ndarray, sJ: np.ndarray, N: int,
                node_type: int, kappa: float, R: float, mu0: float, NAnalyze: int, rho0: float = 1.0,
                chunk: int = 2048) -> np.ndarray:
    """The 15 values of a TGVAnalysis line (its column order) from U(5,n,n,n,K), the lifted
    gradients g{x,y,z}(4,n,n,n,K) of (u,v,w,T) and sJ(n,n,n,K), all FLEXI layout."""
    x, w = nodes_weights(N, node_type)
    xa, wa = nodes_weights(NAnalyze, GAUSS_LOBATTO)
    V = lagrange_matrix(x, xa)
    W = w[:, None, None] * w[None, :, None] * w[None, None, :]
    Wa = wa[:, None, None] * wa[None, :, None] * wa[None, None, :]
    K = U.shape[-1]
    sKappaM1 = 1.0 / (kappa - 1.0)
    vol = float(np.sum(W[..., None] / sJ))
    s = dict(Ekin=0.0, Ekin_comp=0.0, T=0.0, Entropy=0.0, Enstr=0.0, DR_u=0.0, DR_S=0.0, DR_Sd=0.0, DR_p=0.0,
             ED_S=0.0, ED_D=0.0)
    vmax = 0.0

    def interp(A):  # (..., n, n, n, k) -> (..., na, na, na, k), i then j then k as ChangeBasis3D
        A = np.einsum("ai,...ijke->...ajke", V, A)
        A = np.einsum("bj,...ajke->...abke", V, A)
        return np.einsum("ck,...abke->...abce", V, A)

    for e0 in range(0, K, chunk):
        sl = slice(e0, min(K, e0 + chunk))
        sJa = interp(sJ[..., sl])
        Ua = interp(U[..., sl])
        G = np.stack([interp(g[0:3, ..., sl]) for g in (gx, gy, gz)], axis=1)  # (3 vel, 3 dir, a, b, c, k)
        rho = Ua[0]
        vel = Ua[1:4] / rho
        p = (kappa - 1.0) * (Ua[4] - 0.5 * np.sum(Ua[1:4] * vel, axis=0))
        T = p / rho / R
        divU = G[0, 0] + G[1, 1] + G[2, 2]
        S = 0.5 * (G + np.swapaxes(G, 0, 1))
        Sd = S.copy()
        for r in range(3):
            Sd[r, r] = Sd[r, r] - 1.0 / 3.0 * divU
        vort = np.stack([G[2, 1] - G[1, 2], G[0, 2] - G[2, 0], G[1, 0] - G[0, 1]])
        w2 = np.sum(vort * vort, axis=0)
        vmax = max(vmax, float(np.sqrt(w2).max()))
        IF = Wa[..., None] / sJa
        vv = np.sum(vel * vel, axis=0)
        s["T"] += np.sum(IF * T)
        s["Entropy"] += np.sum(IF * (-sKappaM1) * rho * (np.log(p) - kappa * np.log(rho)))
        s["Ekin"] += np.sum(IF * 0.5 * vv)
        s["Ekin_comp"] += np.sum(IF * 0.5 * rho * vv)
        s["Enstr"] += np.sum(IF * 0.5 * rho * w2)
        s["DR_u"] += np.sum(IF * np.sum(G * G, axis=(0, 1)))
        s["DR_S"] += np.sum(IF * np.sum(S * S, axis=(0, 1)))
        s["DR_Sd"] += np.sum(IF * np.sum(Sd * Sd, axis=(0, 1)))
        s["DR_p"] += np.sum(IF * p * divU)
        s["ED_S"] += np.sum(IF * mu0 * w2)
        s["ED_D"] += np.sum(IF * mu0 * divU * divU)
    Ekin = s["Ekin"] / vol
    Ekin_comp = s["Ekin_comp"] / vol / rho0
    DR_S = s["DR_S"] * 2.0 * mu0 / (rho0 * vol)
    DR_Sd = s["DR_Sd"] * 2.0 * mu0 / (rho0 * vol)
    DR_p = -s["DR_p"] / (rho0 * vol)
    return np.array([DR_S, DR_Sd + DR_p, Ekin, Ekin_comp, s["Enstr"] / (rho0 * vol), s["DR_u"] * mu0 / vol, DR_S,
                     DR_Sd, DR_p, vmax, s["T"] / vol, math.sqrt(2.0 / 3.0 * Ekin), s["Entropy"] / vol,
                     s["ED_S"] / vol, s["ED_D"] * 4.0 / 3.0 / vol])


def fortran_e(x: float, d: int = 14, e: int = 5) -> str:
    """Fortran's E(d+9).dEe edit of x (E23.14E5 by default): 0.ddd...E+eeeee, the leading zero
    dropped for negative numbers to fit the width."""
    x = float(x)
    if x == 0.0:
        mant, ex = "0" * d, 0
    else:
        m = f"{abs(x):.{d - 1}e}"  # d significant digits: D.DDDDe+XX
        digits, exp10 = m.split("e")
        mant = digits.replace(".", "")
        ex = int(exp10) + 1
    es = f"{'-' if ex < 0 else '+'}{abs(ex):0{e}d}"
    return (f"-.{mant}E{es}" if x < 0 else f"0.{mant}E{es}")


class TGVAnalysisFile:
    """``<ProjectName>_TGVAnalysis.csv`` with FLEXI's header and number format; ``append`` adds one
    line (time, 15 values). A new file is written unless ``keep`` asks to append to an existing
    one (a restart)."""

    def __init__(self, path: str, keep: bool = False):
        self.path = path
        if not (keep and os.path.exists(path)):
            with open(path, "w") as f:
                f.write(",".join(["Time"] + TGV_VARNAMES) + "\n")

    def append(self, t: float, values) -> None:
        with open(self.path, "a") as f:
            f.write(", ".join(fortran_e(v) for v in [t, *values]) + "\n")
