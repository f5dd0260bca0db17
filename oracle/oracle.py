"""ctypes loader for the CPU oracle (TEST INFRASTRUCTURE ONLY).

Only ``tests/``, ``__graft_entry__.smoke()`` and the CPU-baseline leg of ``bench.py`` may import
this module, and only to check results or to time the CPU baseline. The product path
(``flexi_amd``) never imports it.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = os.path.join(_HERE, "liboracle.so")
# builds of the same sources (oracle/Makefile): "parity" rounds every operation as written (no FMA
# contraction), "fast" is the -O3 timing build with contraction (its difference from "parity" is
# the round-off floor of FMA), "native" the timing build for this host (load_native)
_PATHS = {"parity": _LIB, "fast": os.path.join(_HERE, "liboracle_fast.so")}

NODE_G, NODE_GL, NODE_CL, NODE_VISU = 1, 2, 3, 4
RIEMANN_LF, RIEMANN_ROE, RIEMANN_ROEENTROPYFIX = 1, 3, 33


class Params(C.Structure):
    _fields_ = [("N", C.c_int), ("nodeType", C.c_int), ("splitDG", C.c_int), ("riemann", C.c_int),
                ("kappa", C.c_double), ("R", C.c_double), ("Pr", C.c_double), ("mu0", C.c_double)]


def build() -> None:
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


_libs = {}


def lib(variant: str = "parity"):
    if variant not in _libs:
        path = _PATHS[variant]
        if not os.path.exists(path):
            build()
        L = C.CDLL(path)
        _setup(L)
        _libs[variant] = L
    return _libs[variant]


def load_native(out_dir: str) -> str:
    """Build the -O3 -march=native timing oracle for this host into out_dir (make native) and
    register it as variant "native"; returns the library path."""
    subprocess.run(["make", "-s", "-C", _HERE, "native", f"OUT={out_dir}"], check=True)
    _PATHS["native"] = os.path.join(out_dir, "liboracle_native.so")
    _libs.pop("native", None)
    lib("native")
    return _PATHS["native"]


_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="F_CONTIGUOUS")
_ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="F_CONTIGUOUS")


def _setup(L):
    P = C.POINTER(Params)
    v = C.c_void_p
    L.or_nodes_weights.argtypes = [C.c_int, C.c_int, _dp, _dp, _dp]
    L.or_derivative_matrix.argtypes = [C.c_int, C.c_int, _dp]
    L.or_vandermonde.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, _dp]
    L.or_dg_basis.argtypes = [C.c_int, C.c_int] + [_dp] * 12
    L.or_mappings.argtypes = [C.c_int, _ip, _ip, _ip]
    L.or_calc_metrics.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, _dp, C.c_int, _ip] + [_dp] * 10
    L.or_cons_to_prim.argtypes = [P, _dp, _dp]
    L.or_prim_to_cons.argtypes = [P, _dp, _dp]
    L.or_riemann_point.argtypes = [P] + [_dp] * 8
    L.or_split_volume_flux_pi.argtypes = [_dp] * 7
    L.or_eval_diff_flux_point.argtypes = [P] + [_dp] * 7
    L.or_init.restype = v
    L.or_init.argtypes = [P, C.c_int, C.c_int, _ip, _ip] + [_dp] * 8
    L.or_free.argtypes = [v]
    L.or_set_boundaries.restype = C.c_int
    L.or_set_boundaries.argtypes = [v, C.c_int, _ip, C.c_int, _ip, C.c_int, _dp]
    L.or_set_boundary_data.argtypes = [v, _dp]
    L.or_boundary_point.argtypes = [P, C.c_int, v, v, _dp] + [_dp] * 6 + [_dp] * 3
    L.or_slip_gradients.argtypes = [C.c_int] + [_dp] * 9
    L.or_set_threads.argtypes = [C.c_int]
    L.or_dg_time_derivative.argtypes = [v, C.c_double, _dp, _dp]
    L.or_get_gradients.argtypes = [v, _dp, _dp, _dp]
    L.or_get_density_gradient.argtypes = [v, _dp, _dp, _dp]
    L.or_get_faces.argtypes = [v, _dp, _dp, _dp]
    L.or_calc_timestep.restype = C.c_double
    L.or_calc_timestep.argtypes = [v, _dp, C.c_double, C.c_double, C.POINTER(C.c_int), _dp]
    L.or_fill_cfl_dfl.argtypes = [C.c_int, C.c_int, C.c_double, C.c_double, C.POINTER(C.c_double),
                                  C.POINTER(C.c_double)]
    L.or_rk_coefs.argtypes = [_dp, _dp, _dp]
    L.or_timestep_lserkw2.argtypes = [v, C.c_double, C.c_double, _dp, _dp]
    L.or_tgv_ic.argtypes = [P, C.c_double, C.c_int, _dp, _dp]
    L.or_tgv_analyze.argtypes = [v, _dp, _dp, _dp, _dp, C.c_int, _dp]
    L.or_prolong_to_face_raw.argtypes = [C.c_int, C.c_int, C.c_int, _ip, _dp, _dp, C.c_int, _dp, _dp, _dp]
    L.or_surf_int_raw.argtypes = [C.c_int, C.c_int, C.c_int, _ip, _dp, _dp, C.c_int, _dp, _dp, C.c_double, _dp]


def F(shape, dtype=np.float64):
    return np.zeros(shape, dtype=dtype, order="F")


def params(N, nodeType, splitDG, riemann, kappa=1.4, R=71.42857, Pr=0.72, mu0=6.25e-4) -> Params:
    return Params(N, nodeType, splitDG, riemann, kappa, R, Pr, mu0)


def nodes_weights(N, nodeType):
    x, w, wb = F(N + 1), F(N + 1), F(N + 1)
    lib().or_nodes_weights(N, nodeType, x, w, wb)
    return x, w, wb


def derivative_matrix(N, nodeType):
    D = F((N + 1, N + 1))
    lib().or_derivative_matrix(N, nodeType, D)
    return D


def vandermonde(Nin, tin, Nout, tout, modal=False):
    V = F((Nout + 1, Nin + 1))
    lib().or_vandermonde(Nin, tin, Nout, tout, int(modal), V)
    return V


def dg_basis(N, nodeType):
    n = N + 1
    out = dict(xGP=F(n), wGP=F(n), wBary=F(n), L_Minus=F(n), L_Plus=F(n), L_HatMinus=F(n),
               L_HatPlus=F(n), D=F((n, n)), D_T=F((n, n)), D_Hat=F((n, n)), D_Hat_T=F((n, n)),
               DVolSurf=F((n, n)))
    lib().or_dg_basis(N, nodeType, *out.values())
    return out


def mappings(N):
    n = N + 1
    S2V2 = F((2, n, n, 5, 6), np.int32)
    S2V = F((3, n, n, n, 5, 6), np.int32)
    V2S = F((3, n, n, n, 5, 6), np.int32)
    lib().or_mappings(N, S2V2, S2V, V2S)
    return S2V2, S2V, V2S


def calc_metrics(N, nodeType, mesh):
    n = N + 1
    K, S = mesh.nElems, mesh.nSides
    out = dict(Metrics_fTilde=F((3, n, n, n, K)), Metrics_gTilde=F((3, n, n, n, K)),
               Metrics_hTilde=F((3, n, n, n, K)), sJ=F((n, n, n, K)), Elem_xGP=F((3, n, n, n, K)),
               NormVec=F((3, n, n, S)), TangVec1=F((3, n, n, S)), TangVec2=F((3, n, n, S)),
               SurfElem=F((n, n, S)), Face_xGP=F((3, n, n, S)))
    lib().or_calc_metrics(N, nodeType, mesh.NGeo, K, np.asfortranarray(mesh.NodeCoords), S,
                          np.asfortranarray(mesh.ElemToSide), *out.values())
    return out


def fill_cfl_dfl(N, nodeType, CFL, DFL):
    a, b = C.c_double(), C.c_double()
    lib().or_fill_cfl_dfl(N, nodeType, CFL, DFL, C.byref(a), C.byref(b))
    return a.value, b.value


def rk_coefs():
    A, B, Cc = F(5), F(5), F(5)
    lib().or_rk_coefs(A, B, Cc)
    return A, B, Cc


class Solver:
    """The oracle's DG operator on one mesh (FLEXI layout)."""

    def __init__(self, prm: Params, mesh, metrics, ref_state_prim=None, variant="parity", bc_data=None):
        self.L = lib(variant)
        self.p = prm
        self.N = prm.N
        self.mesh = mesh
        self.m = metrics
        f = np.asfortranarray
        self._keep = [f(metrics[k]) for k in ("Metrics_fTilde", "Metrics_gTilde", "Metrics_hTilde", "sJ",
                                               "NormVec", "TangVec1", "TangVec2", "SurfElem")]
        self.ctx = self.L.or_init(C.byref(prm), mesh.nElems, mesh.nSides, f(mesh.SideToElem),
                                 f(mesh.ElemToSide), *self._keep)
        n = self.N + 1
        self.shape = (5, n, n, n, mesh.nElems)
        if getattr(mesh, "nBCSides", 0):
            ref = np.asfortranarray(ref_state_prim if ref_state_prim is not None else np.zeros((6, 1)))
            bt = np.asfortranarray(mesh.BoundaryType, dtype=np.int32)
            rc = self.L.or_set_boundaries(self.ctx, mesh.nBCSides, np.asfortranarray(mesh.BC, dtype=np.int32),
                                         bt.shape[0], bt, ref.shape[1], ref)
            if rc != 0:
                raise ValueError(f"or_set_boundaries: invalid boundary data (code {rc})")
            if bc_data is not None:
                self._bc_data = np.asfortranarray(bc_data, dtype=np.float64)
                self.L.or_set_boundary_data(self.ctx, self._bc_data)

    def __del__(self):
        try:
            self.L.or_free(self.ctx)
        except Exception:
            pass

    def rhs(self, U, t=0.0):
        Ut = F(self.shape)
        self.L.or_dg_time_derivative(self.ctx, t, np.asfortranarray(U), Ut)
        return Ut

    def gradients(self, n_var_lifting=4):
        """gradUx/y/z of the last right-hand side: (u,v,w,T), or with n_var_lifting=5
        (rho,u,v,w,T) as FLEXI's default PP_OPTLIFT=0 build holds them."""
        n = self.N + 1
        g = [F((4, n, n, n, self.mesh.nElems)) for _ in range(3)]
        self.L.or_get_gradients(self.ctx, *g)
        if n_var_lifting == 4:
            return g
        r = [F((n, n, n, self.mesh.nElems)) for _ in range(3)]
        self.L.or_get_density_gradient(self.ctx, *r)
        return [np.asfortranarray(np.concatenate([ri[None], gi], axis=0)) for ri, gi in zip(r, g)]

    def faces(self):
        n = self.N + 1
        f = [F((5, n, n, self.mesh.nSides)) for _ in range(3)]
        self.L.or_get_faces(self.ctx, *f)
        return f

    def calc_timestep(self, U, CFL, DFL):
        err = C.c_int()
        dtElem = F(self.mesh.nElems)
        dt = self.L.or_calc_timestep(self.ctx, np.asfortranarray(U), CFL, DFL, C.byref(err), dtElem)
        return dt, err.value, dtElem

    def step(self, U, t, dt, R=None):
        """One LSERK step in place on ``U`` (Fortran-ordered float64)."""
        if R is None:
            R = F(self.shape)
        self.L.or_timestep_lserkw2(self.ctx, t, dt, U, R)
        return U

    def tgv_analyze(self, U, grads, NAnalyze=10):
        out = F(15)
        self.L.or_tgv_analyze(self.ctx, np.asfortranarray(U), *grads, NAnalyze, out)
        return out


def tgv_ic(prm: Params, xyz, Ma=0.1):
    """TGV initial state at points ``xyz(3,...)`` (Fortran order) -> ``U(5,...)``."""
    pts = np.asfortranarray(xyz.reshape(3, -1, order="F"))
    U = F((5, pts.shape[1]))
    lib().or_tgv_ic(C.byref(prm), Ma, pts.shape[1], pts, U)
    return U.reshape((5,) + xyz.shape[1:], order="F")


def prolong_to_face_raw(N, nodeType, S2E, L_Minus, L_Plus, Uvol):
    nVar = Uvol.shape[0]
    nSides = S2E.shape[1]
    n = N + 1
    Um, Us = F((nVar, n, n, nSides)), F((nVar, n, n, nSides))
    lib().or_prolong_to_face_raw(N, nodeType, nSides, np.asfortranarray(S2E, dtype=np.int32),
                                 np.asfortranarray(L_Minus), np.asfortranarray(L_Plus), nVar,
                                 np.asfortranarray(Uvol), Um, Us)
    return Um, Us


def surf_int_raw(N, nodeType, S2E, L_HatMinus, L_HatPlus, Fm, Fs, nElems, slave_sign=-1.0):
    nVar = Fm.shape[0]
    n = N + 1
    Ut = F((nVar, n, n, n, nElems))
    lib().or_surf_int_raw(N, nodeType, S2E.shape[1], np.asfortranarray(S2E, dtype=np.int32),
                          np.asfortranarray(L_HatMinus), np.asfortranarray(L_HatPlus), nVar,
                          np.asfortranarray(Fm), np.asfortranarray(Fs), slave_sign, Ut)
    return Ut


def boundary_point(prm: Params, bc_type, ref, Pm, nv, t1, t2, grads=None, data=None):
    """One boundary point through GetBoundaryState / Lifting_GetBoundaryFlux / GetBoundaryFlux
    (before SurfElem): returns (Pb(6), lifting flux(4), flux(5))."""
    f = lambda a: np.asfortranarray(np.asarray(a, dtype=np.float64))
    g = grads if grads is not None else np.zeros((3, 4))
    Pb, Fl, Fx = F(6), F(4), F(5)
    refp = f(ref) if ref is not None else None
    datp = f(data) if data is not None else None
    lib().or_boundary_point(C.byref(prm), int(bc_type), refp.ctypes.data if refp is not None else None,
                            datp.ctypes.data if datp is not None else None, f(Pm), f(nv), f(t1), f(t2),
                            f(g[0]), f(g[1]), f(g[2]), Pb, Fl, Fx)
    return Pb, Fl, Fx


def slip_gradients(bc_type, nv, t1, t2, grads):
    """(3,4) gradients of (u,v,w,T) as a slip wall of type 9 or 91 passes them to its viscous flux."""
    f = lambda a: np.asfortranarray(np.asarray(a, dtype=np.float64))
    q = [F(4) for _ in range(3)]
    lib().or_slip_gradients(int(bc_type), f(nv), f(t1), f(t2), f(grads[0]), f(grads[1]), f(grads[2]), *q)
    return np.array(q)


def set_threads(n: int, variant: str = "parity") -> None:
    lib(variant).or_set_threads(n)
