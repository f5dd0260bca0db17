"""Boundary conditions (FLEXI's GetBoundaryFlux / Lifting_GetBoundaryFlux) and meshes with
boundary sides.

tests/golden/cavity/ holds unmodified copies of the reference's
regressioncheck/checks/parabolic/cavity_3D files: the HOPR mesh cavity4x4x4_mesh.h5 (4^3 elements
on the unit cube, six named boundaries), parameter.ini (N=2, Gauss nodes, weak form,
RoeEntropyFix, BR1, mu0=0.01, isothermal walls and a Dirichlet lid moving with u=1, t=0..1), and
reggie_cavity_Re100_State_0000001.000000000.h5, the State FLEXI wrote at t=1, which its regression
check compares with h5diff at 1e-12 absolute.
"""
import os

import numpy as np
import pytest

import oracle as orc
from flexi_amd import equation
from flexi_amd import run as frun

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CAV = os.path.join(ROOT, "tests", "golden", "cavity")
HD = os.path.join(ROOT, "tests", "golden", "hdf5")
CAVITY_BCS = [("BC_wall_left", (4, 1)), ("BC_wall_right", (4, 1)), ("BC_free", (2, 1))]
CAVITY_REF = [[1.0, 1.0, 0.0, 0.0, 71.4285714286], [1.0, 0.0, 0.0, 0.0, 71.4285714286]]


def fio():
    from flexi_amd import io
    return io


def cavity_mesh(boundaries=CAVITY_BCS):
    return fio().read_mesh_bc(os.path.join(CAV, "cavity4x4x4_mesh.h5"), boundaries)


def test_cavity_mesh_boundary_sides():
    m = cavity_mesh()
    assert (m.nElems, m.nSides, m.nBCSides) == (64, 240, 96)
    assert m.BCNames == ["BC_zminus", "BC_zplus", "BC_wall_lower", "BC_free", "BC_wall_left", "BC_wall_right"]
    # the file's types, with the parameter file's BoundaryType for three of them (BC_free was (2,0))
    assert m.BoundaryType[:, 0].tolist() == [4, 4, 4, 2, 4, 4]
    assert m.BoundaryType[:, 1].tolist() == [1, 1, 1, 1, 1, 1]
    assert np.bincount(m.BC, minlength=7)[1:].tolist() == [16] * 6
    s2e = m.SideToElem
    # boundary sides first (setLocalSideIDs): no neighbour; inner sides have both
    assert np.all(s2e[1, :96] == -1) and np.all(s2e[3, :96] == -1) and np.all(s2e[0, :96] >= 1)
    assert np.all(s2e[1, 96:] >= 1) and np.all(s2e[4, 96:] >= 1)
    # every element's side in ElemToSide points back through SideToElem
    for e in range(m.nElems):
        for ls in range(6):
            sid, flip = m.ElemToSide[:, ls, e]
            col = s2e[:, sid - 1]
            if flip == 0:
                assert col[0] == e + 1 and col[2] == ls + 1
            else:
                assert col[1] == e + 1 and col[3] == ls + 1 and col[4] == flip
    # each boundary side lies on the face of the unit cube its name says (corner nodes of the side)
    from flexi_amd.mesh import side_to_vol
    plane = {"BC_zminus": (2, 0.0), "BC_zplus": (2, 1.0), "BC_wall_lower": (1, 0.0), "BC_free": (1, 1.0),
             "BC_wall_left": (0, 0.0), "BC_wall_right": (0, 1.0)}
    for s in range(96):
        d, val = plane[m.BCNames[m.BC[s] - 1]]
        e, ls = s2e[0, s] - 1, s2e[2, s]
        for q in range(2):
            for p in range(2):
                x = m.NodeCoords[(slice(None),) + side_to_vol(1, 0, p, q, 0, ls) + (e,)]
                assert abs(x[d] - val) < 1e-12, (s, m.BCNames[m.BC[s] - 1], x)


def test_boundary_mesh_errors_and_periodic_agreement():
    io = fio()
    with pytest.raises(RuntimeError, match="not found"):
        cavity_mesh([("BC_nowhere", (4, 1))])
    with pytest.raises(RuntimeError, match="periodic"):
        cavity_mesh([("BC_free", (1, 0))])
    with pytest.raises(RuntimeError, match="boundary sides are not supported"):
        io.read_mesh(os.path.join(CAV, "cavity4x4x4_mesh.h5"))
    # a periodic mesh reads the same through both functions
    p = os.path.join(HD, "CART_HEX_PERIODIC_004_mesh.h5")
    a, b = io.read_mesh(p), io.read_mesh_bc(p)
    assert b.nBCSides == 0 and np.array_equal(a.SideToElem, b.SideToElem)
    assert np.array_equal(a.ElemToSide, b.ElemToSide) and np.array_equal(a.NodeCoords, b.NodeCoords)
    # the freestream fixture's Dirichlet box: all 6*4 outer faces are boundary sides
    c = io.read_mesh_bc(os.path.join(HD, "cartbox3D_mesh.h5"))
    assert c.nElems == 8 and c.nBCSides == 24 and c.nSides == 24 + 12


def oracle_solver(mesh, N, node, split, riemann, mu0=0.01, R=1.0, ref=CAVITY_REF):
    prm = orc.params(N, node, split, riemann, kappa=1.4, R=R, Pr=0.72, mu0=mu0)
    met = orc.calc_metrics(N, node, mesh)
    return orc.Solver(prm, mesh, met, equation.ref_state_prim(ref, R)), met


def run_to(S, U, t_end, CFL, DFL, N, node):
    cfl, dfl = orc.fill_cfl_dfl(N, node, CFL, DFL)
    t, R = 0.0, orc.F(S.shape)
    it = 0
    while True:
        dt_min, err, _ = S.calc_timestep(U, cfl, dfl)
        assert err == 0
        dt, _, fin = frun.update_time_step(dt_min, t, t_end, t_end)
        S.step(U, t, dt, R)
        t += dt
        it += 1
        if fin:
            return t, it


def test_oracle_reproduces_flexis_cavity_state():
    """The oracle's lid-driven cavity (isothermal walls, Dirichlet lid) from rest to t=1: FLEXI's
    State at t=1 to its regression tolerance, 1e-12 absolute (observed 1.7e-13 on the energy,
    |E| ~ 180). This pins GetBoundaryState/GetBoundaryFlux/Lifting_GetBoundaryFlux for types 2
    and 4, the weak-form Gauss operator at N=2 and the time loop's dt rules over 353 steps."""
    m = cavity_mesh()
    S, _ = oracle_solver(m, 2, orc.NODE_G, 0, orc.RIEMANN_ROEENTROPYFIX)
    U0 = equation.prim_to_cons(equation.ref_state_prim(CAVITY_REF, 1.0)[:, 1], 1.4)
    U = np.asfortranarray(np.broadcast_to(U0.reshape(5, 1, 1, 1, 1), S.shape)).copy(order="F")
    t, it = run_to(S, U, 1.0, 0.99, 0.4, 2, orc.NODE_G)
    st = fio().read_state(os.path.join(CAV, "reggie_cavity_Re100_State_0000001.000000000.h5"))
    assert st["time"] == 1.0 and st["N"] == 2 and st["node_type"] == "GAUSS"
    err = np.max(np.abs(st["U"] - U))
    print(f"{it} steps, max |U - U_FLEXI| = {err:.2e}")
    assert t == 1.0 and err <= 1e-12


def smooth_state(xgp, amp=0.05, seed=0):
    rng = np.random.default_rng(seed)
    k = rng.uniform(1.0, 3.0, size=(5, 3))
    ph = rng.uniform(0, 2 * np.pi, size=5)
    x, y, z = xgp
    w = [np.sin(k[v, 0] * x + k[v, 1] * y + k[v, 2] * z + ph[v]) for v in range(5)]
    prim = np.stack([1.0 + amp * w[0], 0.3 * w[1], 0.2 * w[2], 0.1 * w[3], 71.43 * (1.0 + amp * w[4])])
    return np.asfortranarray(equation.prim_to_cons(prim, 1.4))


def weighted_sum(Ut, sJ, w):
    W = w[:, None, None] * w[None, :, None] * w[None, None, :]
    return np.einsum("vijke,ijk,ijke->v", Ut, W, 1.0 / sJ)


@pytest.mark.parametrize("wall", [3, 4, 9])
def test_oracle_walls_conserve_mass(wall):
    """A box closed by walls of one type: no mass crosses a wall (Flux(DENS)=0), so the
    quadrature of the density's time derivative vanishes to round-off for any state; at
    adiabatic walls (3) the energy flux is zero too."""
    bcs = [(n, (wall, 1)) for n in ("BC_zminus", "BC_zplus", "BC_wall_lower", "BC_free", "BC_wall_left",
                                    "BC_wall_right")]
    m = cavity_mesh(bcs)
    S, met = oracle_solver(m, 3, orc.NODE_G, 0, orc.RIEMANN_ROE)
    U = smooth_state(met["Elem_xGP"])
    Ut = S.rhs(U)
    w = orc.dg_basis(3, orc.NODE_G)["wGP"]
    tot = weighted_sum(Ut, met["sJ"], w)
    scale = weighted_sum(np.abs(Ut), met["sJ"], w)
    print(wall, tot / scale)
    assert abs(tot[0]) <= 1e-13 * scale[0]
    if wall == 3:
        assert abs(tot[4]) <= 1e-13 * scale[4]
    else:
        assert abs(tot[4]) > 1e-10 * scale[4]  # heat or work crosses these walls


@pytest.mark.parametrize("bc", [(2, 2), (3, 1), (4, 2), (9, 1)])
def test_oracle_rest_state_is_steady(bc):
    """Fluid at rest at the RefState's temperature and pressure, in a box with one boundary type
    all round: no flux anywhere but the pressure (a wall's Riemann pressure at zero normal velocity
    is the inner pressure), so the state is steady to round-off."""
    bcs = [(n, bc) for n in ("BC_zminus", "BC_zplus", "BC_wall_lower", "BC_free", "BC_wall_left", "BC_wall_right")]
    m = cavity_mesh(bcs)
    S, _ = oracle_solver(m, 2, orc.NODE_G, 0, orc.RIEMANN_ROEENTROPYFIX)
    U0 = equation.prim_to_cons(equation.ref_state_prim(CAVITY_REF, 1.0)[:, 1], 1.4)
    U = np.asfortranarray(np.broadcast_to(U0.reshape(5, 1, 1, 1, 1), S.shape)).copy(order="F")
    Ut = S.rhs(U)
    assert np.max(np.abs(Ut)) < 1e-10 * 178.0


def test_cavity_parameter_file():
    prm = frun.parse_parameter_file(os.path.join(CAV, "parameter.ini"))
    c = frun.Case(prm, CAV)
    assert c.boundaries == CAVITY_BCS
    assert (c.N, c.riemann, c.split, c.mu0, c.R, c.tend, c.analyze_dt, c.CFL, c.DFL) == \
        (2, "roeentropyfix", None, 0.01, 1.0, 1.0, 1.0, 0.99, 0.4)
    assert c.ini_exact == 1 and c.ini_ref == 2 and c.ref_states == CAVITY_REF


# ------------------------------------------------------------------------------------------------
# GPU


def gpu_solver(mesh, N, node, split, riemann, precision="fp64", mu0=0.01):
    from flexi_amd.solver import DGSolver
    return DGSolver(mesh, N, node_type=node, split_dg=split, riemann=riemann, kappa=1.4, R=1.0, Pr=0.72, mu0=mu0,
                    precision=precision, ref_state_prim=equation.ref_state_prim(CAVITY_REF, 1.0), CFLScale=0.99,
                    DFLScale=0.4)


CASES = [(2, "gauss", None, "roeentropyfix"), (3, "gauss", None, "lf"), (4, "gauss-lobatto", None, "roe"),
         (5, "gauss-lobatto", "pi", "roe")]
BC_SETS = {
    "cavity": CAVITY_BCS,
    "adiabatic+slip": [("BC_zminus", (9, 1)), ("BC_zplus", (9, 1)), ("BC_wall_lower", (3, 1)), ("BC_free", (2, 1)),
                       ("BC_wall_left", (3, 1)), ("BC_wall_right", (4, 2))],
}


@pytest.mark.gpu
@pytest.mark.parametrize("bcs", list(BC_SETS))
@pytest.mark.parametrize("N,node,split,riemann", CASES)
def test_gpu_rhs_with_boundaries_matches_oracle(N, node, split, riemann, bcs):
    """Right-hand side and lifted gradients on the cavity mesh with every supported boundary type,
    against the oracle, for a smooth non-trivial state: the four-kernel path (Gauss, weak GL) and
    the fused path (GL split form, boundary sides through the outer face slots)."""
    m = cavity_mesh(BC_SETS[bcs])
    onode = orc.NODE_G if node == "gauss" else orc.NODE_GL
    oriem = {"lf": orc.RIEMANN_LF, "roe": orc.RIEMANN_ROE, "roeentropyfix": orc.RIEMANN_ROEENTROPYFIX}[riemann]
    S, met = oracle_solver(m, N, onode, 1 if split else 0, oriem)
    U = smooth_state(met["Elem_xGP"], seed=N)
    ref = S.rhs(U)
    gref = S.gradients()
    g = gpu_solver(m, N, node, split, riemann)
    try:
        assert g.kernel_pipeline() == ("fused" if split else "classic")
        out = g.rhs(U)
        gg = g.gradients()
    finally:
        g.close()
    err = np.max(np.abs(out - ref)) / np.max(np.abs(ref))
    gerr = max(np.max(np.abs(a - b)) / np.max(np.abs(b)) for a, b in zip(gg, gref))
    print(f"N={N} {node} {split} {riemann} {bcs}: rhs {err:.2e} grads {gerr:.2e}")
    assert err <= 1e-12 and gerr <= 1e-12


@pytest.mark.gpu
def test_gpu_cavity_run_matches_flexis_state(tmp_path):
    """The reference's cavity_3D case through the parameter-file driver on the GPU (353 steps to
    t=1): FLEXI's State at t=1 to its h5diff tolerance, 1e-12 absolute; and the FP32-bulk path
    within 1e-6 relative L2 of it."""
    io = fio()
    prm = frun.parse_parameter_file(os.path.join(CAV, "parameter.ini"))
    ref = io.read_state(os.path.join(CAV, "reggie_cavity_Re100_State_0000001.000000000.h5"))
    out = {}
    for prec in ("fp64", "fp32-bulk"):
        d = tmp_path / prec
        d.mkdir()
        case = frun.Case(prm, CAV, node_type="gauss")
        res = frun.run(case, str(d), precision=prec, log=lambda s: None)
        names = [os.path.basename(p) for p in res["states"]]
        assert names == ["cavity_Re100_State_0000000.000000000.h5", "cavity_Re100_State_0000001.000000000.h5"]
        st = io.read_state(res["states"][-1])
        assert st["time"] == 1.0 and st["node_type"] == "GAUSS"
        out[prec] = (st["U"], res["iterations"])
    U64, it = out["fp64"]
    err = np.max(np.abs(U64 - ref["U"]))
    U32 = out["fp32-bulk"][0]
    rel = np.linalg.norm(U32 - ref["U"]) / np.linalg.norm(ref["U"])
    print(f"{it} steps: FP64 max |U - U_FLEXI| = {err:.2e}; FP32-bulk rel L2 {rel:.2e}")
    assert it == 353 and err <= 1e-12
    assert rel <= 1e-6


@pytest.mark.gpu
@pytest.mark.parametrize("method", ["standardrk3-3", "niegemannrk4-14", "toulorgerk4-8c", "toulorgerk3-7c",
                                    "toulorgerk4-8f", "carpenterrk4-5"])
def test_gpu_dirichlet_box_freestream_matches_flexis_state(tmp_path, method):
    """The reference's timedisc/freestream_3D case (2^3 box, N=3 Gauss, every side a Dirichlet
    boundary with BCState 0 = the constant exact function, t = 0..1e-6) through the driver with
    each 2N-storage method of its list: the end state equals FLEXI's file for that case
    (freestream kept to round-off) within 1e-12 absolute."""
    io = fio()
    prm = frun.parse_parameter_file(os.path.join(HD, "freestream_3D_parameter.ini"))
    assert method in prm["timediscmethod"][0].split(",")
    prm["timediscmethod"] = [method]  # the regression runs one method per comma entry
    case = frun.Case(prm, HD, node_type="gauss")
    res = frun.run(case, str(tmp_path), log=lambda s: None)
    st = io.read_state(res["states"][-1])
    ref = io.read_state(os.path.join(HD, "cartbox_State_0000000.000001000.h5"))
    assert os.path.basename(res["states"][-1]) == "cartbox_State_0000000.000001000.h5"
    assert st["time"] == ref["time"] == 1e-6
    err = np.max(np.abs(st["U"] - ref["U"]))
    print(f"{res['iterations']} steps, max |U - U_FLEXI| = {err:.2e}")
    assert err <= 1e-12


@pytest.mark.gpu
def test_gpu_boundary_api_rejects_what_it_does_not_support():
    """fxg_set_boundaries validates on the host before any kernel runs: unsupported types, a
    missing RefState index, an exact-function state (0) and a non-positive wall temperature are
    errors with a message; a mesh with boundary sides refuses a right-hand side until its
    conditions are set."""
    import ctypes as C
    from flexi_amd import _lib
    from flexi_amd.solver import DGSolver
    m = cavity_mesh()
    good = equation.ref_state_prim(CAVITY_REF, 1.0)
    cases = [
        ({"BC_free": (22, 1)}, good, "not supported"),
        ({"BC_free": (27, 2)}, good, "subsonic inflow"),
        ({"BC_free": (2, 0)}, good, "exact function"),
        ({"BC_free": (2, 3)}, good, "RefState index"),
        ({}, np.asfortranarray(good * np.array([1, 1, 1, 1, 1, 0.0])[:, None]), "temperature"),
    ]
    for over, ref, msg in cases:
        bt = np.asfortranarray(m.BoundaryType.copy())
        for name, (ty, st) in over.items():
            i = m.BCNames.index(name)
            bt[i, 0], bt[i, 1] = ty, st
        mm = cavity_mesh()
        mm.BoundaryType = bt
        with pytest.raises(_lib.FxgError, match=msg):
            DGSolver(mm, 2, node_type="gauss", split_dg=None, riemann="roeentropyfix", R=1.0, mu0=0.01,
                     ref_state_prim=ref)
    # a context on a mesh with boundary sides and no conditions: the right-hand side is refused
    s = DGSolver(m, 2, node_type="gauss", split_dg=None, riemann="roeentropyfix", R=1.0, mu0=0.01,
                 ref_state_prim=good)
    try:
        L = _lib.lib()
        bare = C.c_void_p()
        assert L.fxg_create_from_nodes(C.byref(s.params), m.nElems, m.nSides,
                                       np.asfortranarray(m.SideToElem).ctypes.data,
                                       np.asfortranarray(m.ElemToSide).ctypes.data, m.NGeo,
                                       np.asfortranarray(m.NodeCoords).ctypes.data, C.byref(bare)) == 0
        assert L.fxg_dg_time_derivative(bare, 0.0) != 0
        assert "fxg_set_boundaries" in L.fxg_last_error().decode()
        L.fxg_destroy(bare)
    finally:
        s.close()


def test_generated_walled_box_equals_the_hopr_cavity_mesh():
    """flexi_amd.mesh.walled_box(4,4,4) on the unit cube against the reference's HOPR cavity mesh
    of the same box (matched by element barycentre): the oracle's right-hand side of the same
    smooth state with the same boundary conditions agrees to round-off, although the two number
    their sides differently."""
    from flexi_amd.mesh import walled_box
    mr = cavity_mesh()
    mb = walled_box(4, 4, 4, boundary_types={"BC_yplus": (2, 1)})
    assert (mb.nSides, mb.nBCSides) == (mr.nSides, mr.nBCSides)
    out = []
    for m in (mr, mb):
        S, met = oracle_solver(m, 3, orc.NODE_G, 0, orc.RIEMANN_ROEENTROPYFIX)
        U = smooth_state(met["Elem_xGP"], seed=4)
        out.append((S.rhs(U), met["Elem_xGP"].reshape(3, -1, m.nElems).mean(axis=1)))
    (Ur, xr), (Ub, xb) = out
    d = np.linalg.norm(xr[:, :, None] - xb[:, None, :], axis=0)
    perm = np.argmin(d, axis=1)
    assert np.max(d[np.arange(mr.nElems), perm]) < 1e-12 and len(set(perm)) == mr.nElems
    err = np.max(np.abs(Ur - Ub[..., perm])) / np.max(np.abs(Ub))
    print(f"max |Ut(HOPR) - Ut(walled_box)| / max |Ut| = {err:.2e}")
    assert err <= 1e-12


@pytest.mark.gpu
@pytest.mark.parametrize("bcs", list(BC_SETS))
def test_gpu_fused_boundary_steps_match_four_kernel_path(bcs, monkeypatch):
    """Ten CALCTIMESTEP + LSERK steps of the GL split form (N=5, Roe) on the cavity mesh with
    boundary sides: the fused path against the four-kernel path (FXG_PIPELINE=classic), and the
    fused FP32-bulk path against both."""
    m = cavity_mesh(BC_SETS[bcs])
    S, met = oracle_solver(m, 5, orc.NODE_GL, 1, orc.RIEMANN_ROE)
    U0 = smooth_state(met["Elem_xGP"], seed=9)
    out = {}
    for name, env, prec in (("fused", None, "fp64"), ("classic", "classic", "fp64"), ("fused32", None, "fp32-bulk"),
                            ("classic32", "classic", "fp32-bulk")):
        if env:
            monkeypatch.setenv("FXG_PIPELINE", env)
        else:
            monkeypatch.delenv("FXG_PIPELINE", raising=False)
        g = gpu_solver(m, 5, "gauss-lobatto", "pi", "roe", precision=prec)
        try:
            assert g.kernel_pipeline() == ("classic" if env else "fused")
            g.set_state(U0)
            t = g.run(10)
            out[name] = (g.get_state(), t)
        finally:
            g.close()
    (Uf, tf), (Uc, tc), (U32, _), (Uc32, _) = out["fused"], out["classic"], out["fused32"], out["classic32"]
    err = np.linalg.norm(Uf - Uc) / np.linalg.norm(Uc - U0)
    err32 = np.linalg.norm(U32 - Uf) / np.linalg.norm(Uf)
    errc32 = np.linalg.norm(Uc32 - Uc) / np.linalg.norm(Uc)
    print(f"{bcs}: fused vs classic {err:.2e} (of U-U0), dt sum {tf:.6e} vs {tc:.6e}; fp32-bulk vs fp64 rel "
          f"{err32:.2e} fused, {errc32:.2e} four-kernel")
    assert abs(tf - tc) <= 1e-13 * tc and err <= 1e-11
    # this state's pressure varies by 5 % around 71: FP32 round-off of the pressure terms dominates;
    # the fused boundary path must not add to what the four-kernel FP32 path shows
    assert err32 <= 2.0 * errc32 + 1e-7


@pytest.mark.gpu
@pytest.mark.parametrize("N,node,split", [(3, "gauss", None), (5, "gauss-lobatto", "pi")])
def test_gpu_boundaries_with_host_metrics(N, node, split):
    """fxg_create with the host's metric arrays (firstInnerSide = nBCSides + 1): the same
    right-hand side as the oracle on those metrics, on both pipelines."""
    from flexi_amd.solver import DGSolver
    m = cavity_mesh(BC_SETS["adiabatic+slip"])
    onode = orc.NODE_G if node == "gauss" else orc.NODE_GL
    S, met = oracle_solver(m, N, onode, 1 if split else 0, orc.RIEMANN_ROEENTROPYFIX)
    U = smooth_state(met["Elem_xGP"], seed=2)
    ref = S.rhs(U)
    g = DGSolver(m, N, node_type=node, split_dg=split, riemann="roeentropyfix", R=1.0, mu0=0.01, metrics=met,
                 ref_state_prim=equation.ref_state_prim(CAVITY_REF, 1.0))
    try:
        out = g.rhs(U)
    finally:
        g.close()
    err = np.max(np.abs(out - ref)) / np.max(np.abs(ref))
    print(f"N={N} {node}: {err:.2e}")
    assert err <= 1e-12


# ------------------------------------------------------------------------------------------------
# Inflow/outflow boundaries (23, 24, 25, 27), the slip wall with gradient treatment (91) and the
# per-point Dirichlet data (12, 121): the oracle's boundary point against closed forms


KAPPA, RGAS = 1.4, 287.0


def frame(seed):
    """A random orthonormal side frame (n, t1, t2)."""
    rng = np.random.default_rng(seed)
    n = rng.normal(size=3)
    n /= np.linalg.norm(n)
    t1 = np.cross(n, rng.normal(size=3))
    t1 /= np.linalg.norm(t1)
    return n, t1, np.cross(n, t1)


def inner_state(n, mach_n, seed, rho=1.2, p=1.0e5):
    """Primitive (rho,u,v,w,p,T) with normal Mach number mach_n along n and a tangential part."""
    rng = np.random.default_rng(seed + 100)
    c = np.sqrt(KAPPA * p / rho)
    tang = rng.normal(size=3)
    tang -= np.dot(tang, n) * n
    u = mach_n * c * n + 0.2 * c * tang / np.linalg.norm(tang)
    return np.array([rho, *u, p, p / (rho * RGAS)])


def bprm():
    return orc.params(3, orc.NODE_G, 0, orc.RIEMANN_ROE, kappa=KAPPA, R=RGAS, Pr=0.72, mu0=1.8e-5)


def euler_flux_n(P, n):
    rho, u, p = P[0], P[1:4], P[4]
    un = np.dot(u, n)
    E = p / (KAPPA - 1.0) + 0.5 * rho * np.dot(u, u)
    return np.array([rho * un, *(rho * u * un + p * n), (E + p) * un])


def ref6(rho=0.0, u=0.0, v=0.0, w=0.0, p=0.0):
    return np.array([rho, u, v, w, p, p / (rho * RGAS) if rho else 0.0])


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_oracle_pressure_outflow_supersonic_is_the_inner_state(seed):
    """24 (and the inner state of 25's supersonic branch aside): a supersonic outflow (normal Mach
    1.5) takes the inner state, so the inviscid boundary flux is the physical Euler flux of the inner
    state (the Riemann solver is consistent) and the lifting flux vanishes."""
    n, t1, t2 = frame(seed)
    Pm = inner_state(n, 1.5, seed)
    Pb, Fl, Fx = orc.boundary_point(bprm(), 24, ref6(1.0, p=0.5e5), Pm, n, t1, t2)
    np.testing.assert_allclose(Pb, Pm, rtol=1e-14, atol=1e-12)
    np.testing.assert_allclose(Fx, euler_flux_n(Pm, n), rtol=1e-12, atol=1e-9)
    assert np.max(np.abs(Fl)) <= 1e-12


@pytest.mark.parametrize("bc_type", [24, 25])
def test_oracle_subsonic_outflow_keeps_the_speed_of_sound(bc_type):
    """24 and 25, subsonic outflow: the boundary pressure is the prescribed one, the density follows
    from the inner speed of sound (c_b = c_in), the velocity is the inner one."""
    n, t1, t2 = frame(4)
    Pm = inner_state(n, 0.4, 4)
    pout = 0.93e5
    Pb, _, _ = orc.boundary_point(bprm(), bc_type, ref6(1.0, p=pout), Pm, n, t1, t2)
    assert Pb[4] == pout
    np.testing.assert_allclose(KAPPA * Pb[4] / Pb[0], KAPPA * Pm[4] / Pm[0], rtol=1e-14)
    np.testing.assert_allclose(Pb[1:4], Pm[1:4], rtol=1e-14, atol=1e-12)
    np.testing.assert_allclose(Pb[5], Pb[4] / (RGAS * Pb[0]), rtol=1e-15)


def test_oracle_subsonic_outflow_turns_inflow_outwards():
    """25 with the inner flow entering (normal Mach -0.3): the boundary velocity is |u.n| n, no
    tangential part; pressure the prescribed one."""
    n, t1, t2 = frame(5)
    Pm = inner_state(n, -0.3, 5)
    Pb, _, _ = orc.boundary_point(bprm(), 25, ref6(1.0, p=1.01e5), Pm, n, t1, t2)
    np.testing.assert_allclose(Pb[1:4], abs(np.dot(Pm[1:4], n)) * n, rtol=1e-13, atol=1e-12)
    assert Pb[4] == 1.01e5


def test_oracle_mach_outflow_at_the_inner_mach_number_is_the_inner_state():
    """23 with the prescribed Mach number equal to the inner normal Mach number (subsonic): the
    total pressure round trip returns the inner pressure, so the boundary state is the inner one;
    with a lower outlet Mach number the isentropic relation fixes the pressure ratio."""
    n, t1, t2 = frame(6)
    Pm = inner_state(n, 0.35, 6)
    c = np.sqrt(KAPPA * Pm[4] / Pm[0])
    ma = np.dot(Pm[1:4], n) / c
    Pb, _, _ = orc.boundary_point(bprm(), 23, ref6(1.0, u=ma, p=1.0), Pm, n, t1, t2)
    np.testing.assert_allclose(Pb, Pm, rtol=1e-13, atol=1e-12)
    Pb2, _, _ = orc.boundary_point(bprm(), 23, ref6(1.0, u=0.1, p=1.0), Pm, n, t1, t2)
    g = KAPPA / (KAPPA - 1.0)
    np.testing.assert_allclose(Pb2[4] / Pm[4], ((1 + 0.2 * ma * ma) / (1 + 0.2 * 0.01)) ** g, rtol=1e-13)


@pytest.mark.parametrize("seed", [7, 8, 9])
def test_oracle_subsonic_inflow_total_state_and_outgoing_invariant(seed):
    """27 (NASA/TM-2011-217181 with Slater's correction): the boundary state has the prescribed total
    temperature and total pressure (T (1 + (k-1)/2 M^2) = Tt, p (1 + (k-1)/2 M^2)^(k/(k-1)) = pt),
    flows along the prescribed direction a, and carries the outgoing Riemann invariant of the inner
    state, -(u.n) - 2 c/(k-1), unchanged (the quadratic for c_b is built from exactly that)."""
    n, t1, t2 = frame(seed)
    rng = np.random.default_rng(seed)
    a = -n + 0.3 * t1 * rng.uniform(-1, 1) + 0.2 * t2 * rng.uniform(-1, 1)  # into the domain
    a /= np.linalg.norm(a)
    Pm = inner_state(n, -0.2, seed)
    Tt, pt = 1.05 * Pm[5], 1.04 * Pm[4]
    Pb, _, _ = orc.boundary_point(bprm(), 27, np.array([Tt, *a, pt, 0.0]), Pm, n, t1, t2)
    rho, u, p, T = Pb[0], Pb[1:4], Pb[4], Pb[5]
    cb = np.sqrt(KAPPA * p / rho)
    M2 = np.dot(u, u) / cb ** 2
    np.testing.assert_allclose(T * (1 + 0.5 * (KAPPA - 1) * M2), Tt, rtol=1e-12)
    np.testing.assert_allclose(p * (1 + 0.5 * (KAPPA - 1) * M2) ** (KAPPA / (KAPPA - 1)), pt, rtol=1e-12)
    np.testing.assert_allclose(np.cross(u, a), 0.0, atol=1e-10 * np.linalg.norm(u))
    assert np.dot(u, a) > 0
    cm = np.sqrt(KAPPA * Pm[4] / Pm[0])
    np.testing.assert_allclose(-np.dot(u, n) - 2 * cb / (KAPPA - 1), -np.dot(Pm[1:4], n) - 2 * cm / (KAPPA - 1),
                               rtol=1e-12)
    np.testing.assert_allclose(p, rho * RGAS * T, rtol=1e-14)


def test_oracle_slip_wall_91_gradient_treatment():
    """91: the temperature gradient loses its wall-normal part; for the velocity gradient
    G_ab = du_a/dx_b, in the frame E = (n, t1, t2) the entries of Z = E^T G E that are wall-normal
    derivatives of a tangential component or tangential derivatives of the normal component vanish,
    the others are kept. 9: every gradient loses its wall-normal part, (I - n n^T) grad."""
    n, t1, t2 = frame(11)
    E = np.stack([n, t1, t2], axis=1)
    g = np.random.default_rng(11).normal(size=(3, 4))  # g[b, a] = d(var a)/dx_b
    q = orc.slip_gradients(91, n, t1, t2, g)
    G, Gq = g[:, :3].T, q[:, :3].T
    Z, Zq = E.T @ G @ E, E.T @ Gq @ E
    for (c, d) in [(1, 0), (2, 0), (0, 1), (0, 2)]:
        assert abs(Zq[c, d]) <= 1e-14
    for (c, d) in [(0, 0), (1, 1), (1, 2), (2, 1), (2, 2)]:
        np.testing.assert_allclose(Zq[c, d], Z[c, d], rtol=1e-13, atol=1e-14)
    P = np.eye(3) - np.outer(n, n)
    np.testing.assert_allclose(q[:, 3], P @ g[:, 3], atol=1e-14)
    q9 = orc.slip_gradients(9, n, t1, t2, g)
    np.testing.assert_allclose(q9, P @ g, atol=1e-14)


def test_oracle_slip_wall_91_state_and_flux():
    """91 takes the slip wall's state (9): zero normal velocity, inner density, the Riemann
    pressure; the inviscid flux is (0, p_wall n, 0) and there is no mass or energy flux."""
    n, t1, t2 = frame(12)
    Pm = inner_state(n, 0.1, 12)
    Pb, Fl, Fx = orc.boundary_point(bprm(), 91, None, Pm, n, t1, t2,
                                    grads=np.random.default_rng(12).normal(size=(3, 4)))
    Pb9, Fl9, _ = orc.boundary_point(bprm(), 9, None, Pm, n, t1, t2)
    np.testing.assert_array_equal(Pb, Pb9)
    np.testing.assert_array_equal(Fl, Fl9)
    assert abs(np.dot(Pb[1:4], n)) <= 1e-12 * np.linalg.norm(Pb[1:4]) and Pb[0] == Pm[0]
    assert Fx[0] == 0.0


@pytest.mark.parametrize("bc_type", [12, 121])
def test_oracle_boundary_data_equal_to_the_refstate_is_the_dirichlet_boundary(bc_type):
    """12/121 with BCDataPrim equal to RefState 1 at every boundary point: the right-hand side of
    the Dirichlet boundary (2, RefState 1), bit for bit."""
    bcs = [(nme, (2, 1)) for nme in ("BC_zminus", "BC_zplus", "BC_wall_lower", "BC_free", "BC_wall_left",
                                     "BC_wall_right")]
    m2 = cavity_mesh(bcs)
    S2, met = oracle_solver(m2, 3, orc.NODE_G, 0, orc.RIEMANN_ROE)
    U = smooth_state(met["Elem_xGP"])
    ref = S2.rhs(U)
    md = cavity_mesh([(nme, (bc_type, 0)) for nme, _ in bcs])
    data = np.asfortranarray(np.broadcast_to(equation.ref_state_prim(CAVITY_REF, 1.0)[:, :1, None, None],
                                             (6, 4, 4, md.nBCSides)))
    prm = orc.params(3, orc.NODE_G, 0, orc.RIEMANN_ROE, kappa=1.4, R=1.0, Pr=0.72, mu0=0.01)
    Sd = orc.Solver(prm, md, orc.calc_metrics(3, orc.NODE_G, md), equation.ref_state_prim(CAVITY_REF, 1.0),
                    bc_data=data)
    np.testing.assert_array_equal(Sd.rhs(U), ref)


# RefStates of the inflow/outflow set (R = 1, the cavity's gas): 3 subsonic inflow (Tt, direction
# a into the box through x = 0, pt) as InitBC leaves it, 4 outflow pressure, 5 outlet Mach number
_A27 = np.array([1.0, 0.1, 0.0]) / np.linalg.norm([1.0, 0.1, 0.0])
OPEN_REF = CAVITY_REF + [[71.5, *_A27, 71.7], [1.0, 0.0, 0.0, 0.0, 71.0], [1.0, 0.05, 0.0, 0.0, 71.43]]
OPEN_BCS = [("BC_zminus", (91, 0)), ("BC_zplus", (121, 0)), ("BC_wall_lower", (25, 4)), ("BC_free", (24, 4)),
            ("BC_wall_left", (27, 3)), ("BC_wall_right", (23, 5))]


def open_bc_data(N, nBCSides):
    """BCDataPrim(6,0:N,0:N,nBCSides): a state varying with the point and the side (types 121/12)."""
    n = N + 1
    p, q, s = np.meshgrid(np.arange(n), np.arange(n), np.arange(nBCSides), indexing="ij")
    ph = 0.3 * p + 0.7 * q + 0.11 * s
    rho = 1.0 + 0.01 * np.sin(ph)
    prim = np.stack([rho, 0.05 * np.cos(ph), 0.02 * np.sin(2 * ph), 0.0 * ph, 71.43 * (1.0 + 0.01 * np.cos(ph))])
    return np.asfortranarray(np.concatenate([prim, (prim[4] / prim[0])[None]], axis=0))


def test_oracle_open_boundaries_keep_a_uniform_flow_at_its_state():
    """The open set on the cavity box with a state that every one of its boundaries reproduces
    (rest state at the outflow pressure of 24/25, so 23's outlet Mach number 0 and the inflow of
    27 would have to be zero as well): here only the outflow types 24/25 and the slip wall 91, with
    the rest state at p = 71.0 everywhere, give a steady state to round-off."""
    bcs = [("BC_zminus", (91, 0)), ("BC_zplus", (91, 0)), ("BC_wall_lower", (25, 4)), ("BC_free", (24, 4)),
           ("BC_wall_left", (24, 4)), ("BC_wall_right", (25, 4))]
    m = cavity_mesh(bcs)
    S, _ = oracle_solver(m, 2, orc.NODE_G, 0, orc.RIEMANN_ROEENTROPYFIX, ref=OPEN_REF)
    U0 = equation.prim_to_cons(equation.ref_state_prim(OPEN_REF, 1.0)[:, 3], 1.4)
    U = np.asfortranarray(np.broadcast_to(U0.reshape(5, 1, 1, 1, 1), S.shape)).copy(order="F")
    assert np.max(np.abs(S.rhs(U))) < 1e-10 * 178.0


@pytest.mark.gpu
@pytest.mark.parametrize("N,node,split,riemann", CASES)
def test_gpu_rhs_with_open_boundaries_matches_oracle(N, node, split, riemann):
    """Right-hand side and lifted gradients with the inflow/outflow types (23, 24, 25, 27), the slip
    wall with gradient treatment (91) and per-point Dirichlet data (121) on the cavity box, against
    the oracle, on both pipelines. The state is at Mach 0.03 (p = 71.4, |u| <= 0.3), where the
    pressure terms cancel to about 1/20 of their size: the bar is three times the distance between
    the oracle built as FLEXI's Fortran (no FMA contraction) and the same source with FMA
    contraction on the same state (1.7e-12 to 8.5e-12 measured over these four cases), and 1e-12
    where that distance is smaller."""
    m = cavity_mesh(OPEN_BCS)
    onode = orc.NODE_G if node == "gauss" else orc.NODE_GL
    oriem = {"lf": orc.RIEMANN_LF, "roe": orc.RIEMANN_ROE, "roeentropyfix": orc.RIEMANN_ROEENTROPYFIX}[riemann]
    data = open_bc_data(N, m.nBCSides)
    prm = orc.params(N, onode, 1 if split else 0, oriem, kappa=1.4, R=1.0, Pr=0.72, mu0=0.01)
    met = orc.calc_metrics(N, onode, m)
    refp = equation.ref_state_prim(OPEN_REF, 1.0)
    U = smooth_state(met["Elem_xGP"], seed=N)
    F = orc.Solver(prm, m, met, refp, bc_data=data, variant="fast")
    fast, gfast = F.rhs(U).copy(), [a.copy() for a in F.gradients()]
    S = orc.Solver(prm, m, met, refp, bc_data=data)
    ref = S.rhs(U)
    gref = S.gradients()
    assert np.all(np.isfinite(ref))
    floor = np.max(np.abs(fast - ref)) / np.max(np.abs(ref))
    gfloor = max(np.max(np.abs(a - b)) / np.max(np.abs(b)) for a, b in zip(gfast, gref))
    from flexi_amd.solver import DGSolver
    g = DGSolver(m, N, node_type=node, split_dg=split, riemann=riemann, kappa=1.4, R=1.0, Pr=0.72, mu0=0.01,
                 ref_state_prim=refp, bc_data=data)
    try:
        assert g.kernel_pipeline() == ("fused" if split else "classic")
        out = g.rhs(U)
        gg = g.gradients()
    finally:
        g.close()
    err = np.max(np.abs(out - ref)) / np.max(np.abs(ref))
    gerr = max(np.max(np.abs(a - b)) / np.max(np.abs(b)) for a, b in zip(gg, gref))
    print(f"open boundaries N={N} {node} {split} {riemann}: rhs {err:.2e} (floor {floor:.2e}) "
          f"grads {gerr:.2e} (floor {gfloor:.2e})")
    assert err <= max(1e-12, 3.0 * floor) and gerr <= max(1e-12, 3.0 * gfloor)


@pytest.mark.gpu
def test_gpu_boundary_data_is_required_for_types_12_and_121():
    """A context with a type-121 boundary refuses a right-hand side until BCDataPrim is set."""
    from flexi_amd import _lib
    from flexi_amd.solver import DGSolver
    m = cavity_mesh(OPEN_BCS)
    g = DGSolver(m, 2, node_type="gauss", split_dg=None, riemann="roe", R=1.0, mu0=0.01,
                 ref_state_prim=equation.ref_state_prim(OPEN_REF, 1.0))
    try:
        U = smooth_state(g.metrics()["Elem_xGP"])
        with pytest.raises(_lib.FxgError, match="fxg_set_boundary_data"):
            g.rhs(U)
    finally:
        g.close()
