"""Gradients through the low-storage Runge-Kutta time integration, for torch.

`lserk_steps(U0, solver, dts)` advances a device state by len(dts) TimeStepByLSERKW2 steps with a
frozen time-step sequence; its backward runs the library's discrete adjoint
(`fxg_adjoint_n_steps`: the transposed Jacobian of the right-hand side, stage by stage, with the
states recomputed from checkpoints, or, with checkpoint_every=0, read from the trajectory the
forward pass stored). States are torch tensors of shape (nElems, N+1, N+1, N+1, 5),
the memory order of FLEXI's U(5,0:N,0:N,0:N,nElems).
"""
from __future__ import annotations

import numpy as np
import torch

from .basis import nodes_weights
from .solver import DGSolver


def to_tensor(U: np.ndarray, device="cuda") -> torch.Tensor:
    """FLEXI-layout U(5,n,n,n,K) (Fortran order) -> tensor (K,n,n,n,5) on `device`."""
    return torch.from_numpy(np.ascontiguousarray(np.asarray(U).transpose(4, 3, 2, 1, 0))).to(device)


def to_flexi(T: torch.Tensor) -> np.ndarray:
    """Tensor (K,n,n,n,5) -> FLEXI-layout U(5,n,n,n,K) (Fortran order, host)."""
    return np.asfortranarray(T.detach().cpu().numpy().transpose(4, 3, 2, 1, 0))


def node_weights(solver: DGSolver, device="cuda") -> torch.Tensor:
    """w_i w_j w_k J at the solution nodes, shape (K,n,n,n): the quadrature of a volume integral."""
    w = nodes_weights(solver.N, solver.node_type)[1]
    sJ = solver.metrics()["sJ"]  # (n,n,n,K) Fortran order
    W = (w[:, None, None] * w[None, :, None] * w[None, None, :])[..., None] / sJ
    return torch.from_numpy(np.ascontiguousarray(W.transpose(3, 2, 1, 0))).to(device)


def kinetic_energy(U: torch.Tensor, weights: torch.Tensor) -> torch.Tensor:
    """J = sum_nodes w_i w_j w_k J * rho|u|^2/2 (kinetic energy integrated at the solution nodes)."""
    m = U[..., 1:4]
    return (weights * 0.5 * (m * m).sum(-1) / U[..., 0]).sum()


def primal_dts(solver: DGSolver, U0: np.ndarray, n_steps: int) -> np.ndarray:
    """Time steps CALCTIMESTEP chooses along n_steps steps from U0 (the frozen sequence)."""
    solver.set_state(U0)
    dts = np.empty(n_steps)
    for it in range(n_steps):
        dt, err = solver.CalcTimeStep()
        if err != 0:
            raise RuntimeError(f"CALCTIMESTEP errType={err} at step {it}")
        dts[it] = dt
        solver.dt = dt
        solver.TimeStepByLSERKW2(0.0)
    return dts


def _check_state(name: str, T: torch.Tensor, solver: DGSolver) -> None:
    """The library reads and writes a device array of (nElems, N+1, N+1, N+1, 5) doubles through
    the pointer; anything else would be addressed out of bounds."""
    n = solver.N + 1
    shape = (solver.mesh.nElems, n, n, n, 5)
    if T.dtype != torch.float64:
        raise TypeError(f"{name} must be float64, got {T.dtype}")
    if not T.is_cuda:
        raise ValueError(f"{name} must be a device (cuda) tensor")
    if tuple(T.shape) != shape:
        raise ValueError(f"{name} must have shape {shape}, got {tuple(T.shape)}")


class LSERKSteps(torch.autograd.Function):
    """U(T) = LSERK^n(U0) with frozen time steps; backward by the discrete adjoint."""

    @staticmethod
    def forward(ctx, U0: torch.Tensor, solver: DGSolver, dts, checkpoint_every: int):
        _check_state("U0", U0, solver)
        dts = np.asarray(dts, dtype=np.float64)
        every = int(checkpoint_every)
        U = U0.detach().clone().contiguous()
        # checkpoints of the state every `every` steps (every = 0: the whole trajectory of stage
        # states), kept for the backward pass
        ckpt = torch.empty(solver.adjoint_storage(len(dts), every), dtype=U.dtype, device=U.device)
        torch.cuda.synchronize()  # the library runs on its own stream
        solver.forward_n_steps_dev(U.data_ptr(), dts, ckpt.data_ptr(), every)
        solver.sync()
        ctx.save_for_backward(U0.detach().contiguous(), ckpt)
        ctx.solver, ctx.dts, ctx.every = solver, dts, every
        return U

    @staticmethod
    def backward(ctx, gU: torch.Tensor):
        U0, ckpt = ctx.saved_tensors
        _check_state("grad of U(T)", gU, ctx.solver)
        ubar = gU.detach().clone().contiguous()
        torch.cuda.synchronize()
        ctx.solver.adjoint_n_steps_dev(U0.data_ptr(), ubar.data_ptr(), ctx.dts, ctx.every, ckpt.data_ptr())
        return ubar, None, None, None


def lserk_steps(U0: torch.Tensor, solver: DGSolver, dts, checkpoint_every: int = 20) -> torch.Tensor:
    return LSERKSteps.apply(U0, solver, dts, checkpoint_every)
