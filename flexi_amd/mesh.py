"""Hexahedral box meshes in FLEXI's connectivity format (periodic, or with boundary sides).

This is host-side set-up: it produces exactly the arrays FLEXI's mesh reader and
``setLocalSideIDs``/``fillMeshInfo`` hand to the DG operator (src/mesh/prepare_mesh.f90:50-737,
src/mesh/mesh_vars.f90:85-94):

* ``NodeCoords(3,0:NGeo,0:NGeo,0:NGeo,nElems)`` on equidistant nodes (HOPR convention,
  NodeTypeVISU, src/mesh/metrics.f90:75-156),
* ``ElemToSide(2,6,nElems)`` = (SideID, flip) per local side,
* ``SideToElem(5,nSides)`` = (ElemID, nbElemID, locSide, nbLocSide, flip),

all 1-based like the Fortran arrays. A periodic box has no boundary sides, so every side is an
inner side (``firstInnerSide = 1``, ``lastInnerSide = nSides``; src/mesh/mesh.f90:257-268);
``walled_box`` makes the faces of non-periodic directions boundary sides, numbered first.

Local side ids follow ``flexi.h``: 1 zeta-, 2 eta-, 3 xi+, 4 eta+, 5 xi-, 6 zeta+.
The master of a side is the element on its minus side (flip 0); the slave's flip is found
geometrically by matching the face corners through the side-to-volume mapping of
src/mesh/mappings.f90:188-446.
"""
from __future__ import annotations

from dataclasses import dataclass
from typing import Callable, Optional

import numpy as np

XI_MINUS, XI_PLUS, ETA_MINUS, ETA_PLUS, ZETA_MINUS, ZETA_PLUS = 5, 3, 2, 4, 1, 6

# Direction (0,1,2) and sign (-1,+1) of each local side.
_SIDE_DIR = {ZETA_MINUS: (2, -1), ETA_MINUS: (1, -1), XI_PLUS: (0, +1),
             ETA_PLUS: (1, +1), XI_MINUS: (0, -1), ZETA_PLUS: (2, +1)}


def flip_s2m(N: int, p: int, q: int, flip: int) -> tuple[int, int]:
    """``Flip_S2M`` (src/mesh/mappings.f90:188-221), 3D branch."""
    if flip == 0:
        return p, q
    if flip == 1:
        return q, p
    if flip == 2:
        return N - p, q
    if flip == 3:
        return N - q, N - p
    if flip == 4:
        return p, N - q
    raise ValueError(flip)


def cgns_side_to_vol(N: int, l: int, p: int, q: int, s: int) -> tuple[int, int, int]:
    """``CGNS_SideToVol`` (src/mesh/mappings.f90:274-309), 3D branch."""
    return {XI_MINUS: (l, q, p), XI_PLUS: (N - l, p, q), ETA_MINUS: (p, l, q),
            ETA_PLUS: (N - p, N - l, q), ZETA_MINUS: (q, p, l), ZETA_PLUS: (p, q, N - l)}[s]


def cgns_side_to_vol2(N: int, p: int, q: int, s: int) -> tuple[int, int]:
    """``CGNS_SideToVol2`` (src/mesh/mappings.f90:315-346), 3D branch."""
    return {XI_MINUS: (q, p), XI_PLUS: (p, q), ETA_MINUS: (p, q),
            ETA_PLUS: (N - p, q), ZETA_MINUS: (q, p), ZETA_PLUS: (p, q)}[s]


def side_to_vol(N: int, l: int, p: int, q: int, flip: int, s: int) -> tuple[int, int, int]:
    """``SideToVol`` (src/mesh/mappings.f90:421-431)."""
    a, b = flip_s2m(N, p, q, flip)
    return cgns_side_to_vol(N, l, a, b, s)


def side_to_vol2(N: int, p: int, q: int, flip: int, s: int) -> tuple[int, int]:
    """``SideToVol2`` (src/mesh/mappings.f90:436-446)."""
    a, b = flip_s2m(N, p, q, flip)
    return cgns_side_to_vol2(N, a, b, s)


def build_s2v2(N: int) -> np.ndarray:
    """``S2V2(2,0:N,0:N,0:4,1:6)`` as a Fortran-ordered int32 array."""
    out = np.zeros((2, N + 1, N + 1, 5, 6), dtype=np.int32, order="F")
    for s in range(1, 7):
        for f in range(5):
            for q in range(N + 1):
                for p in range(N + 1):
                    out[:, p, q, f, s - 1] = side_to_vol2(N, p, q, f, s)
    return out


@dataclass
class Mesh:
    """A conforming hexahedral mesh in FLEXI's array layout (all Fortran-ordered)."""

    nElems: int
    nSides: int
    NGeo: int
    NodeCoords: np.ndarray   # (3,NGeo+1,NGeo+1,NGeo+1,nElems) float64
    ElemToSide: np.ndarray   # (2,6,nElems) int32, 1-based SideID
    SideToElem: np.ndarray   # (5,nSides) int32, 1-based ElemIDs
    shape: Optional[tuple[int, int, int]] = None       # elements per direction of a generated box
    box_lo: Optional[tuple[float, float, float]] = None  # bounding box of the nodes
    box_len: Optional[tuple[float, float, float]] = None
    # boundary sides 1..nBCSides (numbered first, prepare_mesh.f90:196-222)
    nBCSides: int = 0
    BC: Optional[np.ndarray] = None            # (nBCSides) int32, 1-based boundary index of each side
    BoundaryType: Optional[np.ndarray] = None  # (nBCs,3) int32: type, state, alpha
    BCNames: Optional[list] = None

    @property
    def firstInnerSide(self) -> int:
        return self.nBCSides + 1

    @property
    def lastInnerSide(self) -> int:
        return self.nSides


def _box_nodes(dims, ijk, lo, length, NGeo, deform) -> np.ndarray:
    """Equidistant geometry nodes of every element of a box, (3,ng,ng,ng,K) in Fortran order."""
    h = np.array([length[d] / dims[d] for d in range(3)])
    xi = np.linspace(0.0, 1.0, NGeo + 1)
    a = xi[:, None, None, None]
    b = xi[None, :, None, None]
    c = xi[None, None, :, None]
    pts = np.stack(np.broadcast_arrays(lo[0] + (ijk[0] + a) * h[0], lo[1] + (ijk[1] + b) * h[1],
                                       lo[2] + (ijk[2] + c) * h[2]), axis=-1)  # (ng,ng,ng,K,3)
    if deform is not None:
        pts = deform(pts)
    return np.asfortranarray(np.moveaxis(pts, -1, 0))


def periodic_box(nx: int, ny: int, nz: int,
                 lo=(-np.pi, -np.pi, -np.pi), length=(2 * np.pi, 2 * np.pi, 2 * np.pi),
                 NGeo: int = 1,
                 deform: Optional[Callable[[np.ndarray], np.ndarray]] = None) -> Mesh:
    """Fully periodic box of ``nx*ny*nz`` hexahedra (element id = ix + nx*(iy + ny*iz) + 1).

    Side ``3*(e-1) + d + 1`` is the face between element ``e`` and its neighbour in the positive
    direction ``d`` (0: xi, 1: eta, 2: zeta); ``e`` is its master through its plus side, the
    neighbour its slave through its minus side. ``deform`` maps points ``(..., 3)`` to displaced
    points; it must be periodic on the box so that periodic faces still match. Each direction
    needs at least two elements so that the corner matching that fixes the flips is unambiguous.
    """
    if min(nx, ny, nz) < 2:
        raise ValueError("periodic_box needs at least 2 elements per direction")
    K = nx * ny * nz
    dims = (nx, ny, nz)
    e = np.arange(K)
    ijk = np.stack([e % nx, (e // nx) % ny, e // (nx * ny)])  # (3,K)

    NodeCoords = _box_nodes(dims, ijk, lo, length, NGeo, deform)

    # Flip of the slave (minus side) relative to the master (plus side), one per direction:
    # match the four face corners on the integer lattice of a 2x2x2 element block.
    plus = {0: XI_PLUS, 1: ETA_PLUS, 2: ZETA_PLUS}
    minus = {0: XI_MINUS, 1: ETA_MINUS, 2: ZETA_MINUS}
    flips = {}
    for d in range(3):
        off = np.zeros(3, dtype=int)
        off[d] = 1
        master = [tuple(np.array(side_to_vol(1, 0, p, q, 0, plus[d]))) for p in (0, 1) for q in (0, 1)]
        for f in range(1, 5):
            slave = [tuple(np.array(side_to_vol(1, 0, p, q, f, minus[d])) + off) for p in (0, 1) for q in (0, 1)]
            if slave == master:
                flips[d] = f
                break
        if d not in flips:
            raise RuntimeError("no consistent slave orientation found")

    nb = []
    for d in range(3):
        o = ijk.copy()
        o[d] = (o[d] + 1) % dims[d]
        nb.append(o[0] + nx * (o[1] + ny * o[2]))
    S = 3 * K
    SideToElem = np.zeros((5, S), dtype=np.int32, order="F")
    ElemToSide = np.zeros((2, 6, K), dtype=np.int32, order="F")
    for d in range(3):
        sid = 3 * e + d  # 0-based
        SideToElem[0, sid] = e + 1
        SideToElem[1, sid] = nb[d] + 1
        SideToElem[2, sid] = plus[d]
        SideToElem[3, sid] = minus[d]
        SideToElem[4, sid] = flips[d]
        ElemToSide[0, plus[d] - 1, e] = sid + 1
        ElemToSide[1, plus[d] - 1, e] = 0
        ElemToSide[0, minus[d] - 1, nb[d]] = sid + 1
        ElemToSide[1, minus[d] - 1, nb[d]] = flips[d]
    return Mesh(nElems=K, nSides=S, NGeo=NGeo, NodeCoords=NodeCoords, ElemToSide=ElemToSide,
                SideToElem=SideToElem, shape=dims, box_lo=tuple(lo), box_len=tuple(length))


def sine_deform(amplitude: float = 0.1) -> Callable[[np.ndarray], np.ndarray]:
    """``x_i <- x_i + a*sin(x)sin(y)sin(z)`` applied to every coordinate (2*pi-periodic)."""
    def f(pts: np.ndarray) -> np.ndarray:
        s = amplitude * np.sin(pts[..., 0]) * np.sin(pts[..., 1]) * np.sin(pts[..., 2])
        return pts + s[..., None]
    return f


# Boundary indices of a generated box, in HOPR's BCIndex order of a Cartesian box
# (z-, y-, x+, y+, x-, z+); (direction, sign) -> 1-based index.
BOX_BC_NAMES = ["BC_zminus", "BC_yminus", "BC_xplus", "BC_yplus", "BC_xminus", "BC_zplus"]
_BOX_BC_INDEX = {(2, -1): 1, (1, -1): 2, (0, +1): 3, (1, +1): 4, (0, -1): 5, (2, +1): 6}


def walled_box(nx: int, ny: int, nz: int, lo=(0.0, 0.0, 0.0), length=(1.0, 1.0, 1.0),
               periodic=(False, False, False), boundary_types=None, NGeo: int = 1) -> Mesh:
    """Box of ``nx*ny*nz`` hexahedra whose faces in each non-periodic direction are boundary
    sides. Sides are numbered as setLocalSideIDs does (src/mesh/prepare_mesh.f90:196-243):
    boundary sides first, in element / local-side order, then the inner sides; a boundary side's
    element is its master (flip 0) and it has no neighbour (SideToElem(2,:) = -1). ``BC`` holds the
    boundary index of each boundary side (``BOX_BC_NAMES`` order) and ``BoundaryType(6,3)`` the
    (type, state, alpha) of each boundary, from ``boundary_types`` ({name: (type, state)}, default
    (4, 1): isothermal walls at RefState 1)."""
    for d, n in enumerate((nx, ny, nz)):
        if n < (2 if periodic[d] else 1):
            raise ValueError("walled_box needs at least one element per direction (two where periodic)")
    dims = (nx, ny, nz)
    K = nx * ny * nz
    e = np.arange(K)
    ijk = np.stack([e % nx, (e // nx) % ny, e // (nx * ny)])
    NodeCoords = _box_nodes(dims, ijk, lo, length, NGeo, None)
    plus = {0: XI_PLUS, 1: ETA_PLUS, 2: ZETA_PLUS}
    minus = {0: XI_MINUS, 1: ETA_MINUS, 2: ZETA_MINUS}
    # the slave flip of an inner side per direction, as periodic_box finds it (sides 1..3 of a
    # 2x2x2 box are the xi, eta and zeta faces of element 1)
    flips = {d: int(f) for d, f in enumerate(periodic_box(2, 2, 2).SideToElem[4, :3])}
    # every local side: inner (master = element on the minus side) or boundary
    sides = {}  # (e, ls) -> ('bc', index) | ('in', master e, master ls, slave e, slave ls, flip)
    for el in range(K):
        for d in range(3):
            i = ijk[d, el]
            o = ijk[:, el].copy()
            if i + 1 < dims[d] or periodic[d]:
                o[d] = (i + 1) % dims[d]
                nb = o[0] + nx * (o[1] + ny * o[2])
                sides[(el, plus[d])] = ("in", el, plus[d], nb, minus[d], flips[d])
                sides[(nb, minus[d])] = sides[(el, plus[d])]
            else:
                sides[(el, plus[d])] = ("bc", _BOX_BC_INDEX[(d, +1)])
            if i == 0 and not periodic[d]:
                sides[(el, minus[d])] = ("bc", _BOX_BC_INDEX[(d, -1)])
    nBC = sum(1 for v in sides.values() if v[0] == "bc")
    nIn = (len(sides) - nBC) // 2
    S = nBC + nIn
    SideToElem = np.full((5, S), -1, dtype=np.int32, order="F")
    ElemToSide = np.zeros((2, 6, K), dtype=np.int32, order="F")
    BC = np.zeros(nBC, dtype=np.int32)
    ids = {}
    next_bc, next_in = 0, nBC
    for el in range(K):
        for ls in range(1, 7):
            v = sides[(el, ls)]
            if v[0] == "bc":
                sid = next_bc
                next_bc += 1
                BC[sid] = v[1]
                SideToElem[0, sid], SideToElem[2, sid] = el + 1, ls
                ElemToSide[:, ls - 1, el] = (sid + 1, 0)
                continue
            if v not in ids:
                ids[v] = next_in
                next_in += 1
            sid = ids[v]
            _, me, mls, se, sls, fl = v
            SideToElem[:, sid] = (me + 1, se + 1, mls, sls, fl)
            ElemToSide[:, ls - 1, el] = (sid + 1, 0 if (el, ls) == (me, mls) else fl)
    bt = dict(boundary_types or {})
    BoundaryType = np.zeros((6, 3), dtype=np.int32, order="F")
    for k, name in enumerate(BOX_BC_NAMES):
        BoundaryType[k, :2] = bt.get(name, (4, 1))
    return Mesh(nElems=K, nSides=S, NGeo=NGeo, NodeCoords=NodeCoords, ElemToSide=ElemToSide,
                SideToElem=SideToElem, shape=dims, box_lo=tuple(lo), box_len=tuple(length), nBCSides=nBC,
                BC=BC, BoundaryType=BoundaryType, BCNames=list(BOX_BC_NAMES))
