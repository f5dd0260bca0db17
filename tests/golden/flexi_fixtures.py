"""Readers for the golden data files of FLEXI's own unit tests and regression checks.

The ``.bin`` files under ``flexi_unittests/`` are unmodified copies of the data files held by the
reference's ``unitTests/`` directory (Fortran unformatted sequential files: each record is framed
by 4-byte little-endian length markers). The producing programs and record layouts are:

* ``NodesAndWeights.bin``  (unitTests/NodesAndWeights.f90:24-44): xi,w,wBary(0:10,1:10,1:4)
* ``DerivativeMatrix.bin`` (unitTests/DerivativeMatrix.f90:15-48): D(0:10,0:10,1:40)
* ``Vandermonde.bin``      (unitTests/Vandermonde.f90:25-72): Vdm_Out_In(0:5,0:6,512), Vdm_In_Out(0:6,0:5,512)
* ``UnittestElementData3D.bin`` (unitTests/unittest.f90:29-53): one curved element at NRef=9
* ``ProlongToFace*.bin``, ``SurfInt*.bin`` (unitTests/ProlongToFace.f90, SurfInt.f90)

``tgv/TGV_Re1600_Split_TGVAnalysis_Reference.csv`` is the reference trajectory of the
``regressioncheck/checks/tgv/split`` case (N=7 Gauss-Lobatto, split PI, RoeEntropyFix, BR1, 8^3).
"""
from __future__ import annotations

import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
UT = os.path.join(HERE, "flexi_unittests")


def records(path: str) -> list[bytes]:
    b = open(path, "rb").read()
    out, pos = [], 0
    while pos < len(b):
        n = int(np.frombuffer(b[pos:pos + 4], "<i4")[0])
        out.append(b[pos + 4:pos + 4 + n])
        assert int(np.frombuffer(b[pos + 4 + n:pos + 8 + n], "<i4")[0]) == n
        pos += 8 + n
    return out


def _f8(buf: bytes, shape, offset=0):
    n = int(np.prod(shape))
    return np.frombuffer(buf, "<f8", count=n, offset=offset).reshape(shape, order="F"), offset + 8 * n


def _i4(buf: bytes, shape, offset=0):
    n = int(np.prod(shape))
    return np.frombuffer(buf, "<i4", count=n, offset=offset).reshape(shape, order="F"), offset + 4 * n


def nodes_and_weights():
    r = records(os.path.join(UT, "NodesAndWeights.bin"))[0]
    xi, o = _f8(r, (11, 10, 4))
    w, o = _f8(r, (11, 10, 4), o)
    wb, o = _f8(r, (11, 10, 4), o)
    return xi, w, wb


def derivative_matrix():
    D, _ = _f8(records(os.path.join(UT, "DerivativeMatrix.bin"))[0], (11, 11, 40))
    return D


def vandermonde():
    r = records(os.path.join(UT, "Vandermonde.bin"))[0]
    oi, o = _f8(r, (6, 7, 512))
    io, o = _f8(r, (7, 6, 512), o)
    return oi, io


NREF = 9


def element_data():
    """SideToElem, S2V2, S2V, L_Minus, L_Plus, L_HatPlus, L_HatMinus, sJ of the reference element."""
    r = records(os.path.join(UT, "UnittestElementData3D.bin"))[0]
    n = NREF + 1
    d = {}
    v, o = _i4(r, (1,))
    d["nElems"] = int(v[0])
    d["SideToElem"], o = _i4(r, (5, 6), o)
    v, o = _i4(r, (3,), o)
    d["firstMPISide_YOUR"], d["lastMPISide_MINE"], d["nSides"] = (int(x) for x in v)
    d["S2V2"], o = _i4(r, (2, n, n, 5, 6), o)
    d["S2V"], o = _i4(r, (3, n, n, n, 5, 6), o)
    for k in ("L_Minus", "L_Plus", "L_HatPlus", "L_HatMinus"):
        d[k], o = _f8(r, (n,), o)
    d["sJ"], o = _f8(r, (n, n, n, 2, 1), o)  # (0:N,0:N,0:N,0:FV_SIZE,nElems), FV_SIZE=1
    assert o == len(r)
    return d


def prolong_uvol():
    n = NREF + 1
    v, _ = _f8(records(os.path.join(UT, "ProlongToFaceUvol.bin"))[0], (n, n, n, 1))
    return v


def prolong_ref(kind: str):
    """kind 'G' or 'GL': the DG record Uface_master(1,:,:,1:6)."""
    n = NREF + 1
    v, _ = _f8(records(os.path.join(UT, f"ProlongToFace_{kind}3D.bin"))[0], (n, n, 6))
    return v


def surfint_flux():
    n = NREF + 1
    v, _ = _f8(records(os.path.join(UT, "SurfIntFlux.bin"))[0], (n, n, 6))
    return v


def surfint_ref(kind: str):
    n = NREF + 1
    v, _ = _f8(records(os.path.join(UT, f"SurfInt_{kind}3D.bin"))[0], (n, n, n, 1))
    return v


def tgv_split_csv():
    path = os.path.join(HERE, "tgv", "TGV_Re1600_Split_TGVAnalysis_Reference.csv")
    with open(path) as f:
        header = f.readline().strip().split(",")
    data = np.loadtxt(path, delimiter=",", skiprows=1)
    return header, data


def almost_equal_abs_or_rel(x, y, tol):
    """``ALMOSTEQUALABSORREL`` of src/flexi.h:57-60."""
    x, y = np.asarray(x), np.asarray(y)
    return (np.abs(x - y) <= tol) | (np.abs(x - y) <= np.maximum(np.abs(x), np.abs(y)) * tol)
