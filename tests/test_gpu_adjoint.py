"""Discrete adjoint of the Gauss-Lobatto split-form right-hand side and of the LSERK steps.

The reference has no adjoint, so nothing reference-held pins these results; they are checked
against the forward map itself:
  - transpose test of the right-hand side: <lam, J d> from central differences of the GPU
    right-hand side equals <J^T lam, d> from fxg_rhs_vjp_dev, for random lam, d;
  - gradients of J(U(T)) = kinetic energy at the solution nodes through many steps (frozen
    time steps) against central differences of J in random directions.
"""
import numpy as np
import pytest

import oracle as orc
from flexi_amd import testcase
from flexi_amd.mesh import periodic_box, sine_deform

pytestmark = pytest.mark.gpu


def make(N, riemann, shape=(3, 3, 2), deform=None, NGeo=1, oracle_metrics=False):
    from flexi_amd.solver import DGSolver
    mesh = periodic_box(*shape, NGeo=NGeo, deform=deform)
    m = orc.calc_metrics(N, orc.NODE_GL, mesh) if oracle_metrics else None
    s = DGSolver(mesh, N, node_type="gauss-lobatto", split_dg="pi", riemann=riemann, metrics=m)
    return mesh, s


def perturbed_tgv(s, seed, amp=2e-2):
    rng = np.random.default_rng(seed)
    U = testcase.tgv(s.metrics()["Elem_xGP"], Ma=1.0)
    U = U * (1.0 + amp * rng.uniform(-1, 1, U.shape))  # jumps at the faces
    return np.asfortranarray(U)


def vjp(s, U, lam):
    import torch
    from flexi_amd.adjoint import to_flexi, to_tensor
    Ud, ld = to_tensor(U), to_tensor(lam)
    out = torch.empty_like(Ud)
    torch.cuda.synchronize()
    s.rhs_vjp_dev(Ud.data_ptr(), ld.data_ptr(), out.data_ptr())
    s.sync()
    return to_flexi(out)


@pytest.mark.parametrize("N,riemann,oracle_metrics", [
    (3, "lf", True), (5, "roe", True), (4, "roeentropyfix", True),
    (5, "roe", False), (7, "lf", False), (2, "roe", False)])
def test_rhs_vjp_is_the_transpose_of_the_jacobian(N, riemann, oracle_metrics):
    mesh, s = make(N, riemann, deform=sine_deform(0.1), NGeo=2, oracle_metrics=oracle_metrics)
    U = perturbed_tgv(s, 1)
    rng = np.random.default_rng(2)
    for _ in range(2):
        lam = np.asfortranarray(rng.normal(size=U.shape))
        d = np.asfortranarray(rng.normal(size=U.shape) * np.abs(U))
        eps = 1e-6
        jd = (s.rhs(U + eps * d) - s.rhs(U - eps * d)) / (2 * eps)
        lhs = np.sum(lam * jd)
        rhs = np.sum(vjp(s, U, lam) * d)
        scale = np.sum(np.abs(lam * jd))
        print(f"N={N} {riemann} oracle_metrics={oracle_metrics}: <lam,Jd>={lhs:.12e} <J^T lam,d>={rhs:.12e}")
        assert abs(lhs - rhs) <= 1e-7 * scale


def fd_gradient_check(s, U0, dts, every, n_dirs, seed, tol):
    import torch
    from flexi_amd.adjoint import kinetic_energy, lserk_steps, node_weights, to_tensor
    W = node_weights(s)
    U0t = to_tensor(U0).requires_grad_(True)
    J = kinetic_energy(lserk_steps(U0t, s, dts, every), W)
    J.backward()
    g = U0t.grad.detach()
    rng = np.random.default_rng(seed)
    errs = []
    for _ in range(n_dirs):
        d = torch.from_numpy(rng.normal(size=tuple(U0t.shape))).to(g.device)
        eps = 1e-6
        with torch.no_grad():
            Jp = kinetic_energy(lserk_steps(U0t.detach() + eps * d, s, dts, every), W)
            Jm = kinetic_energy(lserk_steps(U0t.detach() - eps * d, s, dts, every), W)
        fd = ((Jp - Jm) / (2 * eps)).item()
        ad = torch.sum(g * d).item()
        errs.append(abs(fd - ad) / abs(fd))
        print(f"directional derivative: adjoint {ad:.12e}  central difference {fd:.12e}  rel {errs[-1]:.2e}")
    assert max(errs) <= tol


def test_gradient_through_steps_matches_finite_differences():
    from flexi_amd.adjoint import primal_dts
    mesh, s = make(4, "roe", shape=(3, 3, 3))
    U0 = perturbed_tgv(s, 3, amp=1e-3)
    dts = primal_dts(s, U0, 10)
    fd_gradient_check(s, U0, dts, every=3, n_dirs=4, seed=5, tol=1e-6)


def test_kinetic_energy_gradient_tgv16_n5_200_steps():
    """TGV 16^3, N=5, 200 steps with checkpoints every 20 steps and a frozen dt sequence; 8
    seeded Gaussian directions, eps = 1e-6."""
    from flexi_amd.adjoint import primal_dts
    from flexi_amd.solver import DGSolver
    mesh = periodic_box(16, 16, 16)
    s = DGSolver(mesh, 5, node_type="gauss-lobatto", split_dg="pi", riemann="roe", CFLScale=0.9, DFLScale=0.9)
    U0 = testcase.tgv(s.metrics()["Elem_xGP"])
    dts = primal_dts(s, U0, 200)
    fd_gradient_check(s, U0, dts, every=20, n_dirs=8, seed=0, tol=1e-5)


def test_trajectory_mode_equals_checkpoint_mode_bitwise():
    """checkpoint_every=0: the forward stores every stage state and its face states, and the
    adjoint runs the reverse stages without recomputation. The stored stages are the same kernels
    on the same inputs as the in-place forward and the segment recomputation, so U(T) and the
    gradient agree bit for bit with the checkpointed run (checkpoints every 3 of 10 steps)."""
    import torch
    from flexi_amd.adjoint import kinetic_energy, lserk_steps, node_weights, primal_dts, to_tensor
    mesh, s = make(5, "roe", shape=(3, 3, 3))
    U0 = perturbed_tgv(s, 7, amp=1e-3)
    dts = primal_dts(s, U0, 10)
    W = node_weights(s)
    out = {}
    for every in (3, 0):
        U0t = to_tensor(U0).requires_grad_(True)
        UT = lserk_steps(U0t, s, dts, every)
        kinetic_energy(UT, W).backward()
        out[every] = (UT.detach().clone(), U0t.grad.detach().clone())
    n = s.adjoint_storage(10, 0)
    assert n == 10 * 5 * (5 * 6 ** 3 * mesh.nElems + 10 * 6 ** 2 * mesh.nSides)
    assert s.adjoint_storage(10, 3) == 4 * 5 * 6 ** 3 * mesh.nElems
    assert torch.equal(out[0][0], out[3][0])
    assert torch.equal(out[0][1], out[3][1])
    assert torch.count_nonzero(out[0][1]).item() > 0.9 * out[0][1].numel()
    s.close()
