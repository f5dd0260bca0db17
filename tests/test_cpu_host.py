"""CPU tests of the host side: C-ABI exports, mesh connectivity, tables and initial states."""
import ctypes
import os
import re
import subprocess

import numpy as np
import pytest

import oracle as orc
from flexi_amd import testcase, timedisc
from flexi_amd.mesh import build_s2v2, periodic_box, side_to_vol, sine_deform

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_symbols():
    text = open(os.path.join(ROOT, "include", "fxg.h")).read()
    return sorted(set(re.findall(r"\b(fxg_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_header_symbol():
    from flexi_amd import _lib
    lib = _lib.lib()
    syms = header_symbols()
    assert len(syms) >= 19
    assert sorted(_lib.EXPORTS) == syms
    for s in syms:
        assert hasattr(lib, s), s
    out = subprocess.run(["nm", "-D", "--defined-only", _lib.LIB_PATH], capture_output=True, text=True).stdout
    for s in syms:
        assert re.search(rf"\bT {s}$", out, re.M), s


def test_library_rejects_bad_parameters_without_gpu_work():
    from flexi_amd import _lib
    lib = _lib.lib()
    p = _lib.FxgParams(9, 2, 4, 3, 1.4, 287.0, 0.72, 1e-3, 0, 0, 0)  # N=9 is not built
    ctx = ctypes.c_void_p()
    assert lib.fxg_create(ctypes.byref(p), None, ctypes.byref(ctx)) != 0
    assert b"N must be" in lib.fxg_last_error()
    p = _lib.FxgParams(3, 1, 4, 3, 1.4, 287.0, 0.72, 1e-3, 0, 0, 0)  # split form on Gauss nodes
    assert lib.fxg_create(ctypes.byref(p), None, ctypes.byref(ctx)) != 0
    assert b"Gauss-Lobatto" in lib.fxg_last_error()


@pytest.mark.parametrize("shape", [(2, 2, 2), (3, 2, 4), (4, 4, 4)])
def test_periodic_box_connectivity(shape):
    m = periodic_box(*shape, NGeo=2, deform=sine_deform(0.1))
    K, S = m.nElems, m.nSides
    assert S == 3 * K
    seen = np.zeros((S, 2), int)
    for e in range(K):
        for ls in range(1, 7):
            sid, flip = m.ElemToSide[:, ls - 1, e]
            se = m.SideToElem[:, sid - 1]
            if flip == 0:
                assert se[0] == e + 1 and se[2] == ls
                seen[sid - 1, 0] += 1
            else:
                assert se[1] == e + 1 and se[3] == ls and se[4] == flip
                seen[sid - 1, 1] += 1
    assert np.all(seen == 1)
    # geometric check: matching face nodes of master and slave coincide modulo the period
    NG = m.NGeo
    L = np.array(m.box_len)
    for s in range(S):
        e, nb, ls, nbls, flip = m.SideToElem[:, s]
        for q in range(NG + 1):
            for p in range(NG + 1):
                a = m.NodeCoords[(slice(None),) + side_to_vol(NG, 0, p, q, 0, ls) + (e - 1,)]
                b = m.NodeCoords[(slice(None),) + side_to_vol(NG, 0, p, q, flip, nbls) + (nb - 1,)]
                d = (a - b) / L
                assert np.allclose(d - np.round(d), 0.0, atol=1e-12)


def test_s2v2_matches_reference_element_fixture():
    import flexi_fixtures as fx
    np.testing.assert_array_equal(build_s2v2(fx.NREF), fx.element_data()["S2V2"])


def test_rk_tables_match_oracle():
    A, B, C = orc.rk_coefs()
    s = timedisc.CARPENTER_RK4_5
    np.testing.assert_array_equal(A[1:], s.RKA)
    np.testing.assert_array_equal(B, s.RKb)
    np.testing.assert_array_equal(C[1:], s.RKc)
    for N in (1, 3, 5, 7, 12):
        for node in (1, 2):
            assert timedisc.fill_cfl_dfl(N, node, 0.9, 0.8) == orc.fill_cfl_dfl(N, node, 0.9, 0.8)


def lserk_butcher(scheme):
    """Butcher tableau of the 2N-storage scheme R_i = Ut_i - RKA_i R_{i-1}, U += RKb_i dt R_i."""
    s = scheme.nStages
    A_ = np.concatenate([[0.0], -scheme.RKA])
    a = np.zeros((s, s))
    b = np.zeros(s)
    # U_i = U_0 + sum_j a_ij k_j with R_i = sum_j r_ij k_j
    r = np.zeros((s, s))
    u = np.zeros((s + 1, s))  # coefficients of U after stage i
    for i in range(s):
        a[i] = u[i]
        r[i] = A_[i] * (r[i - 1] if i > 0 else 0) + np.eye(s)[i]
        u[i + 1] = u[i] + scheme.RKb[i] * r[i]
    b = u[s]
    return a, b


def test_carpenter_rk45_is_fourth_order_and_consistent():
    # classic order conditions up to order 4 for the tableau implied by the 2N-storage form
    a, b = lserk_butcher(timedisc.CARPENTER_RK4_5)
    c = a.sum(1)
    np.testing.assert_allclose(c[1:], timedisc.CARPENTER_RK4_5.RKc, rtol=1e-12, atol=1e-14)
    assert abs(b.sum() - 1) < 1e-14
    assert abs(b @ c - 1 / 2) < 1e-14
    assert abs(b @ c ** 2 - 1 / 3) < 1e-14
    assert abs(b @ (a @ c) - 1 / 6) < 1e-14
    assert abs(b @ c ** 3 - 1 / 4) < 1e-14
    assert abs((b * c) @ (a @ c) - 1 / 8) < 1e-14
    assert abs(b @ (a @ c ** 2) - 1 / 12) < 1e-14
    assert abs(b @ (a @ (a @ c)) - 1 / 24) < 1e-14


def test_tgv_initial_state_matches_oracle():
    prm = orc.params(3, orc.NODE_G, 0, orc.RIEMANN_LF)
    x = np.asfortranarray(np.random.default_rng(3).uniform(-np.pi, np.pi, (3, 200)))
    for Ma in (0.1, 1.0):
        np.testing.assert_allclose(testcase.tgv(x, Ma=Ma), orc.tgv_ic(prm, x, Ma=Ma), rtol=1e-15, atol=1e-14)


def test_density_perturbation_is_seeded_and_bounded():
    x = np.zeros((3, 50))
    U1 = testcase.tgv_density_perturbed(x, seed=1)
    U2 = testcase.tgv_density_perturbed(x, seed=1)
    U3 = testcase.tgv_density_perturbed(x, seed=2)
    np.testing.assert_array_equal(U1, U2)
    assert not np.array_equal(U1, U3)
    assert np.all(np.abs(U1[0] - 1.0) <= 1e-3) and np.std(U1[0]) > 1e-4


def test_fortran_binding_compiles_links_and_reports_errors(tmp_path):
    """flexi_amd/fortran/fxg_interface.f90 (the module a FLEXI host USEs) against libfxg.so."""
    import shutil
    from flexi_amd import _lib
    fc = shutil.which("amdflang") or shutil.which("flang")
    if fc is None:
        pytest.skip("no Fortran compiler")
    _lib.lib()
    mod = os.path.join(ROOT, "flexi_amd", "fortran", "fxg_interface.f90")
    drv = os.path.join(ROOT, "tests", "fortran", "fxg_abi_check.f90")
    # the check program calls every exported entry point, so a binding that is missing from the
    # module or not PUBLIC fails to compile
    src = open(drv).read()
    missing = [s for s in _lib.EXPORTS if f"{s}(" not in src]
    assert not missing, missing
    libdir = os.path.dirname(_lib.LIB_PATH)
    subprocess.run([fc, "-c", mod, "-o", str(tmp_path / "m.o"), f"-J{tmp_path}"], check=True, cwd=tmp_path)
    subprocess.run([fc, drv, str(tmp_path / "m.o"), f"-I{tmp_path}", f"-L{libdir}", "-lfxg",
                    f"-Wl,-rpath,{libdir}", "-o", str(tmp_path / "chk")], check=True, cwd=tmp_path)
    out = subprocess.run([str(tmp_path / "chk")], capture_output=True, text=True, check=True).stdout
    assert out.strip().endswith("OK"), out
    sizes = dict(line.split("=") for line in out.splitlines() if line.startswith("sizeof"))
    assert int(sizes["sizeof(fxg_params)"]) == ctypes.sizeof(_lib.FxgParams)
    assert int(sizes["sizeof(fxg_mesh)"]) == ctypes.sizeof(_lib.FxgMesh)


@pytest.mark.parametrize("N", [1, 2, 3, 5, 7, 10])
def test_host_quadrature_matches_oracle(N):
    from flexi_amd.basis import nodes_weights
    for node in (orc.NODE_G, orc.NODE_GL):
        x, w = nodes_weights(N, node)
        xo, wo = orc.nodes_weights(N, node)[:2]
        np.testing.assert_allclose(x, xo, rtol=0, atol=1e-14)
        np.testing.assert_allclose(w, wo, rtol=1e-13, atol=1e-15)
