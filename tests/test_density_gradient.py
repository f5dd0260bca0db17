"""The lifted density gradient of FLEXI's default build (PP_OPTLIFT=0: lifting variables
(rho,u,v,w,T), src/equations/navierstokes/idealgas/eos.h:108-117), which fxg_get_gradients returns
with nVarLifting = 5.

The oracle's density gradient (or_get_density_gradient) is pinned on the CPU by:
- the same BR1 operator as the velocity lifting, which the reference's TGV trajectory pins: with
  a velocity field equal to the density field node by node, the two gradients agree to round-off
  (volume and surface parts, periodic curved meshes, Gauss-Lobatto nodes);
- exactness for a polynomial density on an affine box closed by slip walls (both node types);
- closed forms of the boundary values of Lifting_GetBoundaryFlux (getboundaryflux.f90:890-921) on a
  one-element box at rest with a uniform state: only the surface term at the boundary face
  remains, -2 (rho_f - rho_in) / (h w_0) at Gauss-Lobatto face nodes, with rho_f the boundary value:
  0.5*(rho_in + rho_b) at a Dirichlet side, the wall state's density p/(R T_wall) at an isothermal
  wall, rho_in at a slip wall (no gradient).
The GPU is compared with the oracle on the cavity mesh with every boundary type, after a
right-hand side and after an LSERK step, on both pipelines.
"""
import numpy as np
import pytest

import oracle as orc
from flexi_amd import equation, testcase
from flexi_amd.mesh import BOX_BC_NAMES, periodic_box, sine_deform, walled_box

PRM = dict(kappa=1.4, R=1.0, Pr=0.72, mu0=0.01)


@pytest.mark.parametrize("N,node", [(4, orc.NODE_GL), (5, orc.NODE_GL)])
def test_density_lifting_is_the_velocity_lifting(N, node):
    """Gauss-Lobatto nodes: face values are node values, so the operator acting on rho and on u
    is the same linear map (on Gauss nodes the face velocity is formed from the prolonged
    conservative variables, GetPrimitiveStateSurface, and is not the prolonged velocity)."""
    mesh = periodic_box(3, 2, 2, NGeo=2, deform=sine_deform(0.08))
    m = orc.calc_metrics(N, node, mesh)
    x, y, z = m["Elem_xGP"]
    rho = 1.0 + 0.2 * np.sin(x + 2 * y) * np.cos(z)
    prim = np.stack([rho, rho, 0.3 * np.cos(y), 0.1 * np.sin(z), 50.0 + 0.0 * x])
    U = np.asfortranarray(equation.prim_to_cons(prim, PRM["kappa"]))
    S = orc.Solver(orc.params(N, node, 1 if node == orc.NODE_GL else 0, orc.RIEMANN_ROE, **PRM), mesh, m)
    S.rhs(U)
    g4 = S.gradients()
    g5 = S.gradients(5)
    for a, b in zip(g4, g5):
        assert np.array_equal(b[1:], a)  # the (u,v,w,T) rows are those of PP_OPTLIFT=1
        scale = np.max(np.abs(b[1]))
        assert scale > 0.1
        np.testing.assert_allclose(b[0], b[1], rtol=0, atol=1e-13 * scale)  # d rho = d u


def one_element_box(bc):
    """A unit cube of one Gauss-Lobatto element, all six sides boundaries: slip walls except
    BC_xminus, which gets `bc`."""
    types = {name: (9, 1) for name in BOX_BC_NAMES}
    types["BC_xminus"] = bc
    return walled_box(1, 1, 1, boundary_types=types)


@pytest.mark.parametrize("bc,rho_face", [
    ((2, 2), lambda rin, rb, p, T: 0.5 * (rin + rb)),   # Dirichlet against RefState 2
    ((4, 2), lambda rin, rb, p, T: p / (PRM["R"] * T)),  # isothermal wall at RefState 2's T
    ((9, 1), lambda rin, rb, p, T: rin),                 # slip wall
    ((3, 1), lambda rin, rb, p, T: p / (PRM["R"] * 50.0)),  # adiabatic wall: the inner T
])
def test_density_lifting_boundary_values_closed_form(bc, rho_face):
    N = 3
    mesh = one_element_box(bc)
    m = orc.calc_metrics(N, orc.NODE_GL, mesh)
    rho_in, p_in = 1.3, 65.0                      # T_in = p/(rho R) = 50
    refs = equation.ref_state_prim([[1.0, 0.0, 0.0, 0.0, 60.0], [2.1, 0.0, 0.0, 0.0, 80.0]], PRM["R"])
    rho_b, T_b = refs[0, 1], refs[5, 1]
    S = orc.Solver(orc.params(N, orc.NODE_GL, 1, orc.RIEMANN_ROE, **PRM), mesh, m, refs)
    shape = m["sJ"].shape
    prim = np.stack([np.full(shape, rho_in), np.zeros(shape), np.zeros(shape), np.zeros(shape), np.full(shape, p_in)])
    S.rhs(np.asfortranarray(equation.prim_to_cons(prim, PRM["kappa"])))
    gx, gy, gz = (g[0] for g in S.gradients(5))
    w0 = orc.nodes_weights(N, orc.NODE_GL)[1][0]
    want = np.zeros(shape)
    # (rho_f - rho_in) * SurfElem * n_x * L_hat_minus(0) * sJ with SurfElem = (h/2)^2, n_x = -1,
    # L_hat_minus(0) = 1/w_0, sJ = (2/h)^3, h = 1; only the nodes i = 0 of the xi- face
    want[0] = -2.0 * (rho_face(rho_in, rho_b, p_in, T_b) - rho_in) / (1.0 * w0)
    np.testing.assert_allclose(gx, want, rtol=1e-13, atol=1e-13)
    np.testing.assert_allclose(gy, 0.0, atol=1e-13)
    np.testing.assert_allclose(gz, 0.0, atol=1e-13)


@pytest.mark.parametrize("N,node", [(3, orc.NODE_G), (4, orc.NODE_GL)])
def test_density_lifting_is_exact_for_polynomials_on_affine_meshes(N, node):
    """rho a polynomial of degree <= N on a 2x2x2 box closed by slip walls (whose density
    boundary value is the inner one): the face values of neighbours agree, every surface term
    vanishes, and the lifted gradient is the exact derivative at every node."""
    mesh = walled_box(2, 2, 2, length=(2.0, 1.5, 1.0), boundary_types={name: (9, 1) for name in BOX_BC_NAMES})
    m = orc.calc_metrics(N, node, mesh)
    x, y, z = m["Elem_xGP"]
    rho = 2.0 + 0.1 * (x ** 2 + x * y + z ** 3)
    prim = np.stack([rho, 0.0 * x, 0.0 * x, 0.0 * x, 50.0 + 0.0 * x])
    refs = equation.ref_state_prim([[1.0, 0.0, 0.0, 0.0, 60.0]], PRM["R"])
    S = orc.Solver(orc.params(N, node, 0, orc.RIEMANN_ROE, **PRM), mesh, m, refs)
    S.rhs(np.asfortranarray(equation.prim_to_cons(prim, PRM["kappa"])))
    gx, gy, gz = (g[0] for g in S.gradients(5))
    np.testing.assert_allclose(gx, 0.1 * (2 * x + y), rtol=0, atol=1e-12)
    np.testing.assert_allclose(gy, 0.1 * x, rtol=0, atol=1e-12)
    np.testing.assert_allclose(gz, 0.3 * z ** 2, rtol=0, atol=1e-12)


# ------------------------------------------------------------------------------------------------
# GPU

@pytest.mark.gpu
@pytest.mark.parametrize("N,node,split,riemann", [(3, "gauss", None, "lf"), (5, "gauss-lobatto", "pi", "roe"),
                                                  (4, "gauss-lobatto", None, "roeentropyfix")])
def test_gpu_density_gradient_matches_oracle(N, node, split, riemann):
    """gradU{x,y,z}(5,...) of the GPU (FLEXI's default layout) against the oracle's on the cavity
    mesh with walls, a slip wall, a Dirichlet side, after a right-hand side and after one
    LSERK step (the last stage's state, recomputed by the GPU from U - RKb*dt*R)."""
    from test_boundaries import BC_SETS, cavity_mesh, gpu_solver, oracle_solver, smooth_state
    from flexi_amd.timedisc import SCHEMES
    mesh = cavity_mesh(BC_SETS["adiabatic+slip"])
    onode = orc.NODE_G if node == "gauss" else orc.NODE_GL
    oriem = {"lf": orc.RIEMANN_LF, "roe": orc.RIEMANN_ROE, "roeentropyfix": orc.RIEMANN_ROEENTROPYFIX}[riemann]
    S, met = oracle_solver(mesh, N, onode, 1 if split else 0, oriem)
    U = smooth_state(met["Elem_xGP"], seed=N)
    g = gpu_solver(mesh, N, node, split, riemann)
    try:
        S.rhs(U)
        ref = S.gradients(5)
        g.rhs(U)
        got = g.gradients(5)
        got4 = g.gradients(4)
        err = max(np.max(np.abs(a - b)) / np.max(np.abs(b)) for a, b in zip(got, ref))
        assert all(np.array_equal(a[1:], b) for a, b in zip(got, got4))
        # one step
        sch = SCHEMES["carpenterrk4-5"]
        dt = 1e-4
        Uo, R = U.copy(order="F"), orc.F(S.shape)
        S.step(Uo, 0.0, dt, R)
        ref_s = S.gradients(5)
        g.set_state(U)
        g.dt = dt
        g.TimeStepByLSERKW2(0.0)
        got_s = g.gradients(5)
        err_s = max(np.max(np.abs(a - b)) / np.max(np.abs(b)) for a, b in zip(got_s, ref_s))
        print(f"[density] N={N} {node} {riemann} {g.kernel_pipeline()}: rhs {err:.2e} step {err_s:.2e}", flush=True)
    finally:
        g.close()
    assert err <= 1e-12 and err_s <= 1e-11
