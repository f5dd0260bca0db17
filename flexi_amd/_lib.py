"""Loads libfxg.so (the gfx950 C-ABI library of include/fxg.h) and declares its signatures.

The library must have been built in-tree (``python -m flexi_amd.build`` or
``__graft_entry__.build()``); a missing library is an error, there is no CPU fallback.
``FXG_LIB=strict`` in the environment loads ``libfxg_strict.so`` instead: the same kernels
compiled without FMA contraction, so that every multiply and add rounds as written.
"""
from __future__ import annotations

import ctypes as C
import os

_VARIANT = os.environ.get("FXG_LIB", "default")
if _VARIANT not in ("default", "strict"):
    raise ImportError(f"FXG_LIB={_VARIANT!r}: expected 'default' or 'strict'")
LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)),
                        "libfxg.so" if _VARIANT == "default" else "libfxg_strict.so")


class FxgParams(C.Structure):
    _fields_ = [("N", C.c_int), ("nodeType", C.c_int), ("splitDG", C.c_int), ("riemann", C.c_int),
                ("kappa", C.c_double), ("R", C.c_double), ("Pr", C.c_double), ("mu0", C.c_double),
                ("doWeakLifting", C.c_int), ("doConservativeLifting", C.c_int), ("precision", C.c_int),
                ("metrics", C.c_int)]


class FxgMesh(C.Structure):
    _fields_ = [("nElems", C.c_int), ("nSides", C.c_int), ("firstInnerSide", C.c_int), ("lastInnerSide", C.c_int),
                ("SideToElem", C.c_void_p), ("ElemToSide", C.c_void_p), ("Metrics_fTilde", C.c_void_p),
                ("Metrics_gTilde", C.c_void_p), ("Metrics_hTilde", C.c_void_p), ("sJ", C.c_void_p),
                ("NormVec", C.c_void_p), ("TangVec1", C.c_void_p), ("TangVec2", C.c_void_p),
                ("SurfElem", C.c_void_p)]


class FxgError(RuntimeError):
    pass


_lib = None


def lib() -> C.CDLL:
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise FxgError(f"{LIB_PATH} is missing: build it with `python -m flexi_amd.build` (no CPU fallback exists)")
    # torch's ROCm wheel carries its own libamdhip64.so.7. Loading it first lets libfxg.so bind to
    # that copy (same SONAME), so one HIP runtime serves both and device pointers can be shared;
    # with /opt/rocm's copy loaded first, torch would start a second runtime that sees no GPU.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    L = C.CDLL(LIB_PATH)
    vp, ip, dp = C.c_void_p, C.POINTER(C.c_int), C.POINTER(C.c_double)
    L.fxg_last_error.restype = C.c_char_p
    L.fxg_version.restype = C.c_char_p
    L.fxg_create.argtypes = [C.POINTER(FxgParams), C.POINTER(FxgMesh), C.POINTER(vp)]
    L.fxg_create_from_nodes.argtypes = [C.POINTER(FxgParams), C.c_int, C.c_int, vp, vp, C.c_int, vp, C.POINTER(vp)]
    L.fxg_destroy.argtypes = [vp]
    L.fxg_set_boundaries.argtypes = [vp, C.c_int, vp, vp, C.c_int, vp]
    L.fxg_set_boundary_data.argtypes = [vp, vp]
    L.fxg_set_state.argtypes = [vp, vp]
    L.fxg_get_state.argtypes = [vp, vp]
    L.fxg_get_ut.argtypes = [vp, vp]
    L.fxg_get_gradients.argtypes = [vp, C.c_int, vp, vp, vp]
    L.fxg_tgv_analyze.argtypes = [vp, C.c_int, vp]
    L.fxg_get_metrics.argtypes = [vp] + [vp] * 9
    L.fxg_dg_time_derivative.argtypes = [vp, C.c_double]
    L.fxg_timestep_lserkw2.argtypes = [vp, C.c_double, C.c_double, C.c_int, dp, dp, dp]
    L.fxg_calc_timestep.argtypes = [vp, C.c_double, C.c_double, dp, ip, vp]
    L.fxg_dg_time_derivative_dev.argtypes = [vp, C.c_double, vp, vp, vp]
    L.fxg_calc_timestep_dev_noalloc.argtypes = [vp, vp, C.c_double, C.c_double, vp, vp, vp]
    L.fxg_timestep_lserkw2_dev.argtypes = [vp, C.c_double, C.c_double, C.c_int, dp, dp, dp, vp, vp, vp]
    L.fxg_rhs_vjp_dev.argtypes = [vp, C.c_double, vp, vp, vp, vp]
    L.fxg_forward_n_steps.argtypes = [vp, C.c_int, dp, C.c_int, dp, dp, dp, vp, vp, C.c_int, vp]
    L.fxg_adjoint_n_steps.argtypes = [vp, C.c_int, dp, C.c_int, dp, dp, dp, vp, vp, vp, C.c_int, vp]
    L.fxg_adjoint_storage.argtypes = [vp, C.c_int, C.c_int, C.c_int, C.POINTER(C.c_size_t)]
    L.fxg_pipeline.argtypes = [vp]
    L.fxg_pipeline.restype = C.c_char_p
    L.fxg_sync.argtypes = [vp]
    L.fxg_sizes.argtypes = [vp, ip, ip, ip]
    L.fxg_profile_enable.argtypes = [vp, C.c_int]
    L.fxg_profile_read.argtypes = [vp, C.c_int, dp, ip]
    _lib = L
    return L


# Every exported symbol of include/fxg.h (checked by the CPU test suite).
EXPORTS = ["fxg_last_error", "fxg_version", "fxg_create", "fxg_create_from_nodes", "fxg_destroy", "fxg_set_boundaries",
           "fxg_set_boundary_data",
           "fxg_set_state",
           "fxg_get_state", "fxg_get_ut", "fxg_get_gradients", "fxg_tgv_analyze", "fxg_get_metrics",
           "fxg_dg_time_derivative",
           "fxg_timestep_lserkw2", "fxg_calc_timestep", "fxg_dg_time_derivative_dev", "fxg_timestep_lserkw2_dev",
           "fxg_rhs_vjp_dev", "fxg_forward_n_steps", "fxg_adjoint_n_steps", "fxg_adjoint_storage", "fxg_pipeline",
           "fxg_sync", "fxg_sizes", "fxg_calc_timestep_dev_noalloc",
           "fxg_profile_enable", "fxg_profile_read"]

KERNEL_NAMES = ["prolong_u", "lift", "flux", "volint", "dt_elem", "dt_reduce", "faces", "visc_faces", "dg_fused"]


def check(rc: int) -> None:
    if rc != 0:
        raise FxgError(lib().fxg_last_error().decode())
