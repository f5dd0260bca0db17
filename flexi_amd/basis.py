"""Quadrature nodes and weights of the solution nodes (Gauss, Gauss-Lobatto) on [-1, 1].

Used on the host for volume integrals of device states (e.g. the kinetic energy functional);
the kernels build their operators in C++ (csrc/fxg_basis.h). Same nodes and weights as FLEXI's
LegendreGaussNodesAndWeights / LegGaussLobNodesAndWeights (src/interpolation/basis.f90:370-586).
"""
from __future__ import annotations

import numpy as np
from numpy.polynomial import legendre as L

GAUSS, GAUSS_LOBATTO = 1, 2


def nodes_weights(N: int, node_type: int):
    """(x, w) of the N+1 Gauss (node_type 1) or Gauss-Lobatto (2) nodes, ascending."""
    if node_type == GAUSS:
        x, w = L.leggauss(N + 1)
        return x, w
    if node_type == GAUSS_LOBATTO:
        if N == 0:
            raise ValueError("Gauss-Lobatto needs N >= 1")
        cN = np.zeros(N + 1)
        cN[N] = 1.0
        inner = L.legroots(L.legder(cN)) if N > 1 else np.zeros(0)
        x = np.concatenate([[-1.0], np.sort(inner), [1.0]])
        PN = L.legval(x, cN)
        w = 2.0 / (N * (N + 1) * PN * PN)
        return x, w
    raise ValueError("node_type must be 1 (Gauss) or 2 (Gauss-Lobatto)")
