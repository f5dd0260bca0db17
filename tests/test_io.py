"""FLEXI mesh and State files (flexi_amd.io over libfxg_io.so), on the CPU.

Fixtures under tests/golden/hdf5/ are unmodified copies of files the reference holds:
* CART_HEX_PERIODIC_004_mesh.h5 (tutorials/taylorgreenvortex), CART_HEX_PERIODIC_008_mesh.h5
  (regressioncheck/checks/tgv/split): HOPR meshes of the periodic box [0,2pi]^3;
* cartbox3D_mesh.h5, cartbox_State_0000000.000000000.h5, cartbox_State_0000000.000001000.h5 and
  freestream_3D_parameter.ini (= parameter.ini there) (regressioncheck/checks/timedisc/
  freestream_3D): a 2^3 box with Dirichlet boundaries and the states FLEXI wrote at t=0 and
  t=1e-6 for RefState (rho,u,v,w,p) = (1,1,1,1,1), N=3 Gauss nodes.
"""
import os
import re
import subprocess

import numpy as np
import pytest

import oracle as orc
from flexi_amd import testcase
from flexi_amd.mesh import periodic_box, side_to_vol

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
G = os.path.join(ROOT, "tests", "golden", "hdf5")


def io():
    from flexi_amd import io as fio
    return fio


def test_io_library_exports_every_header_symbol():
    fio = io()
    text = open(os.path.join(ROOT, "include", "fxg_io.h")).read()
    syms = sorted(set(re.findall(r"\b(fxg_io_[a-z0-9_]+)\s*\(", text)))
    assert len(syms) == 8
    out = subprocess.run(["nm", "-D", "--defined-only", fio.IO_LIB_PATH], capture_output=True, text=True).stdout
    for s in syms:
        assert re.search(rf"\bT {s}$", out, re.M), s


def test_reads_the_state_flexi_wrote():
    s = io().read_state(os.path.join(G, "cartbox_State_0000000.000000000.h5"))
    assert s["U"].shape == (5, 4, 4, 4, 8) and s["N"] == 3
    assert s["time"] == 0.0 and s["node_type"] == "GAUSS" and s["mesh_file"] == "cartbox3D_mesh.h5"
    # RefState (1,1,1,1,1) primitive -> (rho, m, E) = (1, 1,1,1, p/(kappa-1) + rho|u|^2/2 = 4)
    ref = np.array([1.0, 1.0, 1.0, 1.0, 1.0 / 0.4 + 1.5])
    assert np.max(np.abs(s["U"] - ref[:, None, None, None, None])) <= 1e-14


def test_state_round_trip(tmp_path):
    fio = io()
    rng = np.random.default_rng(3)
    U = np.asfortranarray(rng.normal(size=(5, 6, 6, 6, 7)))
    p = str(tmp_path / "case_State_0000000.125000000.h5")
    fio.write_state(p, U, 5, "gauss-lobatto", 0.125, "case_mesh.h5", project="case")
    s = fio.read_state(p)
    assert np.array_equal(s["U"], U)
    buf = np.zeros_like(U, order="F")
    assert fio.read_state(p, out=buf)["U"] is buf and np.array_equal(buf, U)
    with pytest.raises(ValueError):
        fio.read_state(p, out=np.zeros((5, 6, 6, 6, 6), order="F"))
    assert s["N"] == 5 and s["time"] == 0.125
    assert s["node_type"] == "GAUSS-LOBATTO" and s["mesh_file"] == "case_mesh.h5"
    with pytest.raises(ValueError):
        fio.write_state(p, U, 4, "gauss", 0.0, "m.h5")


def test_rejects_meshes_with_boundary_sides():
    with pytest.raises(RuntimeError, match="boundary sides are not supported"):
        io().read_mesh(os.path.join(G, "cartbox3D_mesh.h5"))


@pytest.mark.parametrize("name,ne", [("CART_HEX_PERIODIC_004_mesh.h5", 4), ("CART_HEX_PERIODIC_008_mesh.h5", 8)])
def test_periodic_mesh_connectivity_matches_the_geometry(name, ne):
    """Brute force: every side has one master (flip 0) and one slave, ElemToSide and SideToElem
    agree, and each side point of the master coincides with the slave's point under the slave's
    flip (SideToVol, mappings.f90:421-431) up to a whole period of the box."""
    m = io().read_mesh(os.path.join(G, name))
    K, S = m.nElems, m.nSides
    assert K == ne ** 3 and S == 3 * K and m.NGeo == 1
    L = np.array(m.box_len)
    assert np.allclose(L, 2 * np.pi) and np.allclose(m.box_lo, 0.0)
    seen = np.zeros(S, dtype=int)
    for e in range(K):
        for ls in range(1, 7):
            sid, flip = m.ElemToSide[:, ls - 1, e]
            col = m.SideToElem[:, sid - 1]
            if flip == 0:
                assert col[0] == e + 1 and col[2] == ls
            else:
                assert col[1] == e + 1 and col[3] == ls and col[4] == flip
            seen[sid - 1] += 1
    assert np.all(seen == 2)
    X = m.NodeCoords
    flips = set()
    for s in range(S):
        em, en, lm, ln, fl = m.SideToElem[:, s]
        flips.add(int(fl))
        for q in range(2):
            for p in range(2):
                a = X[(slice(None),) + side_to_vol(1, 0, p, q, 0, lm) + (em - 1,)]
                b = X[(slice(None),) + side_to_vol(1, 0, p, q, fl, ln) + (en - 1,)]
                k = (a - b) / L
                assert np.allclose(k, np.round(k), atol=1e-12), (s, p, q)
    assert flips - {0}, "the fixture exercises non-zero flips"


def test_rhs_on_read_mesh_equals_generated_box():
    """The oracle's right-hand side of the same TGV state on FLEXI's 4^3 mesh file and on the
    generated box of the same elements (matched by barycentre): equal to round-off, although the
    two number sides, masters and flips differently. (The file's coordinates carry 15
    significant digits, 2pi = 6.28318530717959, so the metric terms differ in the last bits.)"""
    N = 3
    mr = io().read_mesh(os.path.join(G, "CART_HEX_PERIODIC_004_mesh.h5"))
    mb = periodic_box(4, 4, 4, lo=(0.0, 0.0, 0.0))
    out = []
    for m in (mr, mb):
        met = orc.calc_metrics(N, orc.NODE_GL, m)
        S = orc.Solver(orc.params(N, orc.NODE_GL, 1, orc.RIEMANN_ROEENTROPYFIX), m, met)
        U = testcase.tgv(met["Elem_xGP"], Ma=1.0)
        out.append((S.rhs(U), met["Elem_xGP"].reshape(3, -1, m.nElems).mean(axis=1)))
    (Ur, xr), (Ub, xb) = out
    d = np.linalg.norm(xr[:, :, None] - xb[:, None, :], axis=0)
    perm = np.argmin(d, axis=1)
    assert np.max(d[np.arange(mr.nElems), perm]) < 1e-12 and len(set(perm)) == mr.nElems
    err = np.max(np.abs(Ur - Ub[..., perm])) / np.max(np.abs(Ub))
    print(f"max |Ut(file) - Ut(box)| / max |Ut| = {err:.2e}")
    assert err <= 1e-12
