"""flexi_amd — MI355X (gfx950) DGSEM Navier-Stokes right-hand side and LSERK stepper with FLEXI's
call surface (DGTimeDerivative_weakForm, TimeStepByLSERKW2, CALCTIMESTEP).

The compute path is the HIP library ``libfxg.so`` (C ABI in ``include/fxg.h``); this package holds
the host-side set-up (meshes, initial and reference states, RK tables), the ctypes binding, the
parameter-file driver and FLEXI's output formats.
"""
from .mesh import Mesh, periodic_box, sine_deform  # noqa: F401
from .timedisc import CARPENTER_RK4_5, STANDARD_RK3_3, fill_cfl_dfl  # noqa: F401

__all__ = ["Mesh", "periodic_box", "sine_deform", "CARPENTER_RK4_5", "STANDARD_RK3_3", "fill_cfl_dfl", "DGSolver"]


def __getattr__(name):
    if name == "DGSolver":
        from .solver import DGSolver
        return DGSolver
    raise AttributeError(name)
