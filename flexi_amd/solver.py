"""Host-side mirror of FLEXI's hot-path call surface on top of libfxg.

``DGSolver`` keeps the FLEXI names and meanings:

* ``DGTimeDerivative_weakForm(t)``  -> src/dg/dg.f90:221 (Ut from U, on the device),
* ``TimeStepByLSERKW2(t)``          -> src/timedisc/timestep.f90:82 (uses ``self.dt``),
* ``CalcTimeStep()``                -> src/equations/navierstokes/calctimestep.f90:89
  (returns ``(dt, errType)``; a nonzero errType raises, as FLEXI aborts),

with the state in FLEXI's layout ``U(5,0:N,0:N,0:N,nElems)`` (numpy, Fortran order) on the host.
Device-resident use (torch tensors) goes through ``rhs_dev``/``step_dev``.
"""
from __future__ import annotations

import ctypes as C
from typing import Optional

import numpy as np

from . import _lib
from .mesh import Mesh
from .timedisc import CARPENTER_RK4_5, LSERK, fill_cfl_dfl

NODE_TYPES = {"gauss": 1, "gauss-lobatto": 2, 1: 1, 2: 2}
RIEMANN = {"lf": 1, "roe": 3, "roeentropyfix": 33, 1: 1, 3: 3, 33: 33}
SPLITDG = {None: 0, "none": 0, "pi": 4, 0: 0, 4: 4}
PRECISION = {"fp64": 0, "fp32-bulk": 1, "fp32": 1, 0: 0, 1: 1}
METRIC_ORDER = {"own": 0, "reference": 1, 0: 0, 1: 1}


def _f(a, dtype=np.float64):
    return np.asfortranarray(a, dtype=dtype)


def _ptr(a: np.ndarray) -> int:
    return a.ctypes.data


class DGSolver:
    """One device context of the DG operator and LSERK stepper on a periodic hexahedral mesh."""

    def __init__(self, mesh: Mesh, N: int, node_type="gauss-lobatto", split_dg="pi", riemann="roe",
                 kappa=1.4, R=71.42857, Pr=0.72, mu0=6.25e-4, precision="fp64", metrics: Optional[dict] = None,
                 scheme: LSERK = CARPENTER_RK4_5, CFLScale=0.9, DFLScale=0.9,
                 ref_state_prim: Optional[np.ndarray] = None, bc_data: Optional[np.ndarray] = None,
                 metric_order: str = "own"):
        L = _lib.lib()
        self.mesh = mesh
        self.N = N
        self.node_type = NODE_TYPES[node_type]
        self.split = SPLITDG[split_dg]
        self.riemann = RIEMANN[riemann.lower() if isinstance(riemann, str) else riemann]
        self.precision = PRECISION[precision]
        self.kappa, self.R, self.Pr, self.mu0 = kappa, R, Pr, mu0
        self.scheme = scheme
        self.CFL, self.DFL = fill_cfl_dfl(N, self.node_type, CFLScale, DFLScale, scheme)
        # metric terms computed by the library (no `metrics`): "own" (FXG_METRICS_OWN) or "reference"
        # (FXG_METRICS_REFERENCE, metrics.f90's operation order and FLEXI's shared side normals)
        self.params = _lib.FxgParams(N, self.node_type, self.split, self.riemann, kappa, R, Pr, mu0, 0, 0,
                                     self.precision, METRIC_ORDER[metric_order])
        self.ctx = C.c_void_p()
        s2e, e2s = _f(mesh.SideToElem, np.int32), _f(mesh.ElemToSide, np.int32)
        if metrics is None:
            nc = _f(mesh.NodeCoords)
            _lib.check(L.fxg_create_from_nodes(C.byref(self.params), mesh.nElems, mesh.nSides, _ptr(s2e), _ptr(e2s),
                                               mesh.NGeo, _ptr(nc), C.byref(self.ctx)))
        else:
            keep = {k: _f(metrics[k]) for k in ("Metrics_fTilde", "Metrics_gTilde", "Metrics_hTilde", "sJ",
                                               "NormVec", "TangVec1", "TangVec2", "SurfElem")}
            m = _lib.FxgMesh(mesh.nElems, mesh.nSides, mesh.firstInnerSide, mesh.lastInnerSide, _ptr(s2e), _ptr(e2s),
                             *[_ptr(keep[k]) for k in ("Metrics_fTilde", "Metrics_gTilde", "Metrics_hTilde", "sJ",
                                                       "NormVec", "TangVec1", "TangVec2", "SurfElem")])
            _lib.check(L.fxg_create(C.byref(self.params), C.byref(m), C.byref(self.ctx)))
        if getattr(mesh, "nBCSides", 0):
            # InitBC: BC(nBCSides), BoundaryType(nBCs,3), RefStatePrim(6,nRefState)
            self._bc = (_f(mesh.BC, np.int32), _f(mesh.BoundaryType, np.int32),
                        _f(ref_state_prim if ref_state_prim is not None else np.zeros((6, 1))))
            bc, bt, rp = self._bc
            rc = L.fxg_set_boundaries(self.ctx, bt.shape[0], _ptr(bc), _ptr(bt), rp.shape[1], _ptr(rp))
            if rc == 0 and bc_data is not None:
                # BCDataPrim(6,0:N,0:N,nBCSides) of the boundaries of types 12 and 121
                self._bc_data = _f(bc_data)
                rc = L.fxg_set_boundary_data(self.ctx, _ptr(self._bc_data))
            if rc != 0:
                msg = L.fxg_last_error().decode()
                L.fxg_destroy(self.ctx)
                self.ctx = C.c_void_p()
                raise _lib.FxgError(msg)
        n = N + 1
        self.shape = (5, n, n, n, mesh.nElems)
        self.dt = 0.0
        self.t = 0.0
        self._RKA = (C.c_double * max(1, scheme.nStages - 1))(*scheme.RKA)
        self._RKb = (C.c_double * scheme.nStages)(*scheme.RKb)
        self._RKc = (C.c_double * max(1, scheme.nStages - 1))(*scheme.RKc)

    # -- lifetime ------------------------------------------------------------------------------
    def close(self) -> None:
        if self.ctx:
            _lib.lib().fxg_destroy(self.ctx)
            self.ctx = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- state transfer (FLEXI layout) ----------------------------------------------------------
    def set_state(self, U: np.ndarray) -> None:
        U = _f(U)
        if U.shape != self.shape:
            raise ValueError(f"U must have shape {self.shape}")
        _lib.check(_lib.lib().fxg_set_state(self.ctx, _ptr(U)))

    def get_state(self) -> np.ndarray:
        U = np.zeros(self.shape, order="F")
        _lib.check(_lib.lib().fxg_get_state(self.ctx, _ptr(U)))
        return U

    def get_ut(self) -> np.ndarray:
        Ut = np.zeros(self.shape, order="F")
        _lib.check(_lib.lib().fxg_get_ut(self.ctx, _ptr(Ut)))
        return Ut

    def gradients(self, n_var_lifting: int = 4):
        """gradUx, gradUy, gradUz (n_var_lifting,0:N,0:N,0:N,nElems) of the last right-hand side:
        4 lifting variables (u,v,w,T) as FLEXI's PP_OPTLIFT=1 build holds them, or 5 (rho,u,v,w,T)
        as its default PP_OPTLIFT=0 build does."""
        n = self.N + 1
        g = [np.zeros((n_var_lifting, n, n, n, self.mesh.nElems), order="F") for _ in range(3)]
        _lib.check(_lib.lib().fxg_get_gradients(self.ctx, int(n_var_lifting), *[_ptr(x) for x in g]))
        return g

    def tgv_analyze(self, n_analyze: int) -> np.ndarray:
        """The 15 values of one TGVAnalysis line (AnalyzeTestcase,
        src/testcase/taylorgreenvortex/testcase.f90:241-479), evaluated on the device from the
        state and the gradients ``gradients()`` would return, on n_analyze+1 Gauss-Lobatto
        points per direction."""
        out = np.zeros(15)
        _lib.check(_lib.lib().fxg_tgv_analyze(self.ctx, int(n_analyze), _ptr(out)))
        return out

    def metrics(self) -> dict:
        n, K, S = self.N + 1, self.mesh.nElems, self.mesh.nSides
        out = dict(Metrics_fTilde=np.zeros((3, n, n, n, K), order="F"), Metrics_gTilde=np.zeros((3, n, n, n, K), order="F"),
                   Metrics_hTilde=np.zeros((3, n, n, n, K), order="F"), sJ=np.zeros((n, n, n, K), order="F"),
                   NormVec=np.zeros((3, n, n, S), order="F"), TangVec1=np.zeros((3, n, n, S), order="F"),
                   TangVec2=np.zeros((3, n, n, S), order="F"), SurfElem=np.zeros((n, n, S), order="F"),
                   Elem_xGP=np.zeros((3, n, n, n, K), order="F"))
        _lib.check(_lib.lib().fxg_get_metrics(self.ctx, *[_ptr(v) for v in out.values()]))
        return out

    # -- FLEXI call surface ---------------------------------------------------------------------
    def DGTimeDerivative_weakForm(self, t: float) -> None:
        _lib.check(_lib.lib().fxg_dg_time_derivative(self.ctx, float(t)))

    def TimeStepByLSERKW2(self, t: float) -> None:
        _lib.check(_lib.lib().fxg_timestep_lserkw2(self.ctx, float(t), float(self.dt), self.scheme.nStages,
                                                   self._RKA, self._RKb, self._RKc))

    def CalcTimeStep(self, with_elements: bool = False):
        dt, err = C.c_double(), C.c_int()
        dtElem = np.zeros(self.mesh.nElems) if with_elements else None
        rc = _lib.lib().fxg_calc_timestep(self.ctx, self.CFL, self.DFL, C.byref(dt), C.byref(err),
                                          None if dtElem is None else _ptr(dtElem))
        if rc != 0 and err.value == 0:
            _lib.check(rc)
        if with_elements:
            return dt.value, err.value, dtElem
        return dt.value, err.value

    # -- convenience ----------------------------------------------------------------------------
    def rhs(self, U: np.ndarray, t: float = 0.0) -> np.ndarray:
        self.set_state(U)
        self.DGTimeDerivative_weakForm(t)
        return self.get_ut()

    def run(self, n_steps: int, t: float = 0.0, fixed_dt: Optional[float] = None) -> float:
        """UpdateTimeStep + TimeStepByLSERKW2 loop of TimeDisc (src/timedisc/timedisc.f90:190-212)."""
        for _ in range(n_steps):
            if fixed_dt is None:
                dt, err = self.CalcTimeStep()
                if err != 0:
                    raise _lib.FxgError(f"CALCTIMESTEP errType={err} at t={t}")
                self.dt = dt
            else:
                self.dt = fixed_dt
            self.TimeStepByLSERKW2(t)
            t += self.dt
        self.t = t
        return t

    # -- device-pointer forms -------------------------------------------------------------------
    def rhs_dev(self, U_ptr: int, Ut_ptr: int, t: float = 0.0, stream: int = 0) -> None:
        _lib.check(_lib.lib().fxg_dg_time_derivative_dev(self.ctx, float(t), U_ptr, Ut_ptr, stream or None))

    def step_dev(self, U_ptr: int, R_ptr: int, dt: float, t: float = 0.0, stream: int = 0) -> None:
        _lib.check(_lib.lib().fxg_timestep_lserkw2_dev(self.ctx, float(t), float(dt), self.scheme.nStages, self._RKA,
                                                       self._RKb, self._RKc, U_ptr, R_ptr, stream or None))

    # -- discrete adjoint (Gauss-Lobatto split form, FP64) -----------------------------------------
    def rhs_vjp_dev(self, U_ptr: int, lam_ptr: int, out_ptr: int, t: float = 0.0, stream: int = 0) -> None:
        """out = (dUt/dU)^T lam at U (device pointers, FLEXI layout)."""
        _lib.check(_lib.lib().fxg_rhs_vjp_dev(self.ctx, float(t), U_ptr, lam_ptr, out_ptr, stream or None))

    def _dt_array(self, dts):
        dts = np.ascontiguousarray(np.asarray(dts, dtype=np.float64))
        return dts, dts.ctypes.data_as(C.POINTER(C.c_double))

    def forward_n_steps_dev(self, U_ptr: int, dts, ckpt_ptr: int = 0, checkpoint_every: int = 20,
                            stream: int = 0) -> None:
        """len(dts) LSERK steps on the device state at U_ptr with the given time steps; with
        ckpt_ptr, the state before every checkpoint_every-th step is stored there."""
        arr, p = self._dt_array(dts)
        _lib.check(_lib.lib().fxg_forward_n_steps(self.ctx, len(arr), p, self.scheme.nStages, self._RKA, self._RKb,
                                                  self._RKc, U_ptr, ckpt_ptr or None, int(checkpoint_every),
                                                  stream or None))

    def adjoint_n_steps_dev(self, U0_ptr: int, ubar_ptr: int, dts, checkpoint_every: int = 20, ckpt_ptr: int = 0,
                            stream: int = 0) -> None:
        """ubar <- (dU(T)/dU(0))^T ubar for the steps of forward_n_steps_dev from U0 (from the
        checkpoints at ckpt_ptr when given)."""
        arr, p = self._dt_array(dts)
        _lib.check(_lib.lib().fxg_adjoint_n_steps(self.ctx, len(arr), p, self.scheme.nStages, self._RKA, self._RKb,
                                                  self._RKc, U0_ptr, ckpt_ptr or None, ubar_ptr,
                                                  int(checkpoint_every), stream or None))

    def adjoint_storage(self, n_steps: int, checkpoint_every: int) -> int:
        """Doubles of the ckpt buffer for n_steps steps: checkpoints every checkpoint_every steps,
        or (0) the whole trajectory of stage states that spares the adjoint its recomputation."""
        n = C.c_size_t()
        _lib.check(_lib.lib().fxg_adjoint_storage(self.ctx, int(n_steps), self.scheme.nStages, int(checkpoint_every),
                                                  C.byref(n)))
        return n.value

    def kernel_pipeline(self) -> str:
        """'fused' or 'classic' (which stage kernels this context launches)."""
        return _lib.lib().fxg_pipeline(self.ctx).decode()

    def sync(self) -> None:
        _lib.check(_lib.lib().fxg_sync(self.ctx))

    # -- kernel timing --------------------------------------------------------------------------
    def profile(self, enable: bool = True) -> None:
        _lib.check(_lib.lib().fxg_profile_enable(self.ctx, int(enable)))

    def kernel_times(self) -> dict:
        """{kernel name: (total ms, launches)} since profile(True)."""
        out = {}
        for kid, name in enumerate(_lib.KERNEL_NAMES):
            ms, cnt = C.c_double(), C.c_int()
            _lib.check(_lib.lib().fxg_profile_read(self.ctx, kid, C.byref(ms), C.byref(cnt)))
            out[name] = (ms.value, cnt.value)
        return out
