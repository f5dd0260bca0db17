"""FLEXI's time loop over the fxg library: ``python -m flexi_amd.run parameter.ini``.

A FLEXI parameter file (readintools syntax: ``key = value`` lines, ``!`` comments, arrays as
``(/a,b,c/)``, keys case-insensitive, repeated keys such as RefState collected in order) drives
the GPU solver the way FLEXI's TimeDisc does (src/timedisc/timedisc.f90:100-212): the state is
written at the start time (TStart, or the restart file's time), then CALCTIMESTEP and one
TimeStepByLSERKW2 step per iteration with UpdateTimeStep's rules (dt shortened to hit the next
analysis time and TEnd, within dt/100; src/timedisc/timedisc_func.f90:215-268) and
AnalyzeTimeStep's output (src/timedisc/timedisc_func.f90:318-381): State files every nWriteData
analysis intervals and at the end (TEnd, or maxIter iterations), named
``<ProjectName>_State_<TIMESTAMP>.h5`` (TIMESTAMP of src/globals/globals.f90:512-540).

Keys read: ProjectName, MeshFile, N, Riemann (LF, Roe, RoeEntropyFix), SplitDG (PI; absent: weak
form), IniExactFunc with IniRefState/RefState (the Taylor-Green vortex of the TGV testcase, or a
uniform RefState for IniExactFunc = 1), kappa, R, Pr, mu0, TStart, TEnd, CFLScale, DFLScale,
Analyze_dt, nWriteData, maxIter, TimeDiscMethod, nCalcTimeStepMax (only 1), MachNumber,
IniConstDens (T only), NAnalyze and nAnalyzeTestCase of the TGV testcase (whose analysis goes to
``<ProjectName>_TGVAnalysis.csv`` as FLEXI writes it), BoundaryName/BoundaryType (boundary types
2, 23, 24, 25, 27 with a RefState, or 2 with BCState 0 and IniExactFunc = 1; 3, 4, 9, 91). The
node type is a compile-time choice in FLEXI (FLEXI_NODETYPE); here it is ``--node-type``. Other
keys are reported and ignored.

The metric terms are computed on the GPU in metrics.f90's operation order with FLEXI's side
normals shared by both sides (``--metric-order reference``, the default), so that a run follows
FLEXI's own to round-off; ``--metric-order own`` takes the library's metrics, which keep the
freestream exact on curved meshes.
"""
from __future__ import annotations

import argparse
import os
import re
import sys
import time

import numpy as np

from . import analysis, equation, io, testcase
from .solver import DGSolver
from .timedisc import SCHEMES

USED = {"projectname", "meshfile", "n", "riemann", "splitdg", "iniexactfunc", "inirefstate", "refstate", "kappa",
        "r", "pr", "mu0", "tstart", "tend", "cflscale", "dflscale", "analyze_dt", "nwritedata", "maxiter",
        "timediscmethod", "ncalctimestepmax", "machnumber", "iniconstdens", "boundaryname", "boundarytype",
        "nanalyze", "nanalyzetestcase", "nwritestats"}


def parse_parameter_file(path: str) -> dict:
    """{lower-case key: [value strings in order]} of a FLEXI parameter file."""
    out: dict = {}
    for line in open(path, encoding="latin-1"):
        line = line.split("!", 1)[0].strip()
        if "=" not in line:
            continue
        k, v = line.split("=", 1)
        out.setdefault(k.strip().lower(), []).append(v.strip())
    return out


def _array(v: str) -> list[float]:
    m = re.fullmatch(r"\(/(.*)/\)", v.replace(" ", ""))
    if not m:
        raise ValueError(f"not an array: {v}")
    return [float(x.replace("d", "e").replace("D", "e")) for x in m.group(1).split(",")]


def _real(v: str) -> float:
    return float(v.replace("d", "e").replace("D", "e"))


def timestamp(t: float) -> str:
    """FLEXI's TIMESTAMP: F17.9 with the leading blanks as zeros."""
    return f"{t:17.9f}".replace(" ", "0")


def update_time_step(dt_min: float, t: float, t_analyze: float, tend: float) -> tuple[float, bool, bool]:
    """UpdateTimeStep's choice of dt (timedisc_func.f90:248-268): the smallest of the stable step,
    the time to the next analysis and the time to TEnd; a step that would leave less than dt/100
    before the analysis or the end is stretched to reach it. Returns (dt, doAnalyze, doFinalize)."""
    dt_an, dt_end = t_analyze - t, tend - t
    dt = min(dt_min, dt_an, dt_end)
    do_analyze = dt == dt_an or dt == dt_end
    do_finalize = dt == dt_end
    dt = min(x for x in (dt_min, dt_an, dt_end) if x > 0)
    if dt_an - dt < dt / 100.0 and dt_an > 0:
        dt, do_analyze = dt_an, True
    if dt_end - dt < dt / 100.0 and dt_end > 0:
        dt, do_analyze, do_finalize = dt_end, True, True
    return dt, do_analyze, do_finalize


class Case:
    """The run a parameter file describes (defaults as FLEXI's DefineParameters*)."""

    def __init__(self, prm: dict, base_dir: str, node_type: str = "gauss-lobatto"):
        def get(key, default=None, conv=str):
            if key not in prm:
                if default is None:
                    raise KeyError(f"parameter {key} is missing")
                return default
            return conv(prm[key][-1])

        self.project = get("projectname", "fxg")
        mesh_file = get("meshfile")
        self.mesh_path = mesh_file if os.path.isabs(mesh_file) else os.path.join(base_dir, mesh_file)
        self.mesh_file = mesh_file
        self.N = get("n", conv=int)
        self.node_type = node_type
        self.riemann = get("riemann", "roeentropyfix").lower()
        self.split = {"pi": "pi"}[get("splitdg").lower()] if "splitdg" in prm else None
        self.kappa = get("kappa", 1.4, _real)
        self.R = get("r", 287.058, _real)
        self.Pr = get("pr", 0.72, _real)
        self.mu0 = get("mu0", 0.0, _real)
        self.t0 = get("tstart", 0.0, _real)
        self.tend = get("tend", conv=_real)
        self.CFL = get("cflscale", conv=_real)
        self.DFL = get("dflscale", conv=_real)
        self.analyze_dt = get("analyze_dt", 0.0, _real)
        self.n_write = get("nwritedata", 1, int)
        if self.analyze_dt <= 0.0:  # analyze.f90:124-130: analyses only at the end
            self.analyze_dt, self.n_write = self.tend, 1
        self.max_iter = get("maxiter", -1, int)
        method = get("timediscmethod", "carpenterrk4-5").lower()
        if method not in SCHEMES:
            raise ValueError(f"TimeDiscMethod {method}: one of {', '.join(SCHEMES)}")
        self.scheme = SCHEMES[method]
        if get("ncalctimestepmax", 1, int) != 1:
            raise ValueError("nCalcTimeStepMax other than 1 is not supported")
        self.ini_exact = get("iniexactfunc", 0, int)
        if self.ini_exact not in (0, 1):
            raise ValueError("IniExactFunc: only 0 (the Taylor-Green vortex testcase) and 1 (RefState)")
        self.mach = get("machnumber", 0.1, _real)  # TGV testcase, testcase.f90:61
        self.n_analyze_points = get("nanalyze", 2 * (self.N + 1), int)  # analyze.f90:118
        self.n_analyze_testcase = get("nanalyzetestcase", 10, int)  # testcase.f90:59
        if get("iniconstdens", "T").upper().strip(".") not in ("T", "TRUE"):
            raise ValueError("IniConstDens = F (constant temperature TGV) is not supported")
        self.ref_states = [_array(v) for v in prm.get("refstate", [])]
        # BoundaryName / BoundaryType pairs (ReadBCs, mesh_readin.f90:92-100)
        names, types = prm.get("boundaryname", []), prm.get("boundarytype", [])
        if len(names) != len(types):
            raise ValueError("every BoundaryName needs a BoundaryType")
        self.boundaries = [(nm, tuple(int(x) for x in _array(tp))[:2]) for nm, tp in zip(names, types)]
        self.ini_ref = get("inirefstate", 1, int)
        self.ignored = sorted(k for k in prm if k not in USED)

    def initial_state(self, xgp: np.ndarray) -> np.ndarray:
        if self.ini_exact == 1:  # RefStateCons(IniRefState) everywhere (exactfunc.f90:258-259)
            prim = equation.ref_state_prim(self.ref_states, self.R)[:, self.ini_ref - 1]
            cons = equation.prim_to_cons(prim, self.kappa)
            return np.asfortranarray(np.broadcast_to(cons.reshape((5,) + (1,) * (xgp.ndim - 1)),
                                                     (5,) + xgp.shape[1:]))
        # the Taylor-Green vortex of the TGV testcase (testcase.f90:176-219)
        return testcase.tgv(xgp, kappa=self.kappa, R=self.R, Ma=self.mach)


def run(case: Case, out_dir: str, precision: str = "fp64", restart: str | None = None, log=print,
        metric_order: str = "reference") -> dict:
    mesh = io.read_mesh_bc(case.mesh_path, case.boundaries)
    ref_prim = equation.ref_state_prim(case.ref_states, case.R) if case.ref_states else None
    if mesh.nBCSides and np.any((mesh.BoundaryType[:, 0] == 2) & (mesh.BoundaryType[:, 1] == 0)):
        # Dirichlet boundaries with BCState 0 take ExactFunc(IniExactFunc) as their state
        # (getboundaryflux.f90:261-265); for IniExactFunc = 1 that is ConsToPrim(RefStateCons(IniRefState)),
        # passed as one more RefState
        if case.ini_exact != 1:
            raise ValueError("Dirichlet boundaries with BCState 0 (exact function) need IniExactFunc = 1")
        cons = equation.prim_to_cons(ref_prim[:, case.ini_ref - 1], case.kappa)
        ref_prim = np.asfortranarray(np.column_stack([ref_prim, equation.cons_to_prim(cons, case.kappa, case.R)]))
        bt = mesh.BoundaryType.copy(order="F")
        bt[(bt[:, 0] == 2) & (bt[:, 1] == 0), 1] = ref_prim.shape[1]
        mesh.BoundaryType = bt
    if mesh.nBCSides and ref_prim is not None:
        # subsonic inflow (27): InitBC turns the RefState's (alpha, beta) in degrees into the inflow
        # direction vector a(1:3) in place (getboundaryflux.f90:176-186)
        ref_prim = np.asfortranarray(ref_prim.copy())
        for st in sorted({int(x) for x in mesh.BoundaryType[mesh.BoundaryType[:, 0] == 27, 1]}):
            ta, tb = np.tan(np.pi / 180.0 * ref_prim[1, st - 1]), np.tan(np.pi / 180.0 * ref_prim[2, st - 1])
            nrm = np.sqrt(1.0 + ta ** 2 + tb ** 2)
            ref_prim[1:4, st - 1] = (1.0 / nrm, ta / nrm, tb / nrm)
    s = DGSolver(mesh, case.N, node_type=case.node_type, split_dg=case.split, riemann=case.riemann,
                 kappa=case.kappa, R=case.R, Pr=case.Pr, mu0=case.mu0, precision=precision, scheme=case.scheme,
                 CFLScale=case.CFL, DFLScale=case.DFL, ref_state_prim=ref_prim, metric_order=metric_order)
    try:
        return _time_loop(case, s, mesh.nElems, out_dir, restart, log)
    finally:
        s.close()


def _time_loop(case: Case, s: DGSolver, n_elems: int, out_dir: str, restart: str | None, log) -> dict:
    t = case.t0
    if restart:
        st = io.read_state(restart)
        nt = io.NODE_TYPE_NAMES[case.node_type]
        if st["N"] != case.N or st["node_type"] != nt or st["U"].shape[4] != n_elems:
            raise ValueError(f"{restart}: N={st['N']} {st['node_type']} with {st['U'].shape[4]} elements, the run "
                             f"has N={case.N} {nt} with {n_elems} elements (no interpolation on restart)")
        U0, t = st["U"], st["time"]
    else:
        U0 = case.initial_state(s.metrics()["Elem_xGP"])
    s.set_state(U0)
    written = []

    def write(tw):
        p = os.path.join(out_dir, f"{case.project}_State_{timestamp(tw)}.h5")
        io.write_state(p, s.get_state(), case.N, case.node_type, tw, case.mesh_file, project=case.project)
        written.append(p)
        return p

    # the TGV testcase's analysis file (InitTestcase, testcase.f90:132-162) and its first line,
    # from the right-hand side TimeDisc evaluates before the loop (timedisc.f90:136,157)
    tgv = None
    if case.ini_exact == 0:
        tgv = analysis.TGVAnalysisFile(os.path.join(out_dir, f"{case.project}_TGVAnalysis.csv"),
                                       keep=restart is not None)

        def analyze_tgv(tt):
            tgv.append(tt, s.tgv_analyze(case.n_analyze_points))
        s.DGTimeDerivative_weakForm(t)
        analyze_tgv(t)
    # TimeDisc (timedisc.f90:100-212): the state at the start time is written, also on restart
    write(t)
    tend, t_analyze = case.tend, min(t + case.analyze_dt, case.tend)
    it, it_analyze, write_counter, dof = 0, 0, 0, n_elems * (case.N + 1) ** 3
    t_wall = time.perf_counter()
    while True:
        # UpdateTimeStep (timedisc_func.f90:215-268)
        dt_min, err = s.CalcTimeStep()
        if err != 0:
            write(t)
            raise RuntimeError(f"CALCTIMESTEP errType={err} at t={t}")
        dt, do_analyze, do_finalize = update_time_step(dt_min, t, t_analyze, tend)
        s.dt = dt
        s.TimeStepByLSERKW2(t)
        it += 1
        it_analyze += 1
        t += dt
        # AnalyzeTimeStep (timedisc_func.f90:318-381)
        if it == case.max_iter:
            tend = t_analyze = t
            do_analyze = do_finalize = True
        if do_analyze:  # a fresh right-hand side for the analysis (timedisc_func.f90:324-325)
            s.DGTimeDerivative_weakForm(t)
        if tgv is not None and (it % case.n_analyze_testcase == 0 or do_analyze):
            # AnalyzeTestcase (timedisc_func.f90:340): between analyses with the gradients of the
            # last stage's right-hand side, as FLEXI's lifted gradients are then
            analyze_tgv(t)
        if do_analyze:
            s.sync()
            wall = time.perf_counter() - t_wall
            work = it_analyze * case.scheme.nStages * dof
            log(f"Time = {t:.8e}  dt = {dt:.8e}  iter = {it}  PID = {wall / work:.4e} s  "
                f"({work / max(wall, 1e-30):.3e} DOF-stage/s)")
            write_counter += 1
            if write_counter == case.n_write or do_finalize:
                write(t)
                write_counter = 0
            it_analyze, t_wall = 0, time.perf_counter()
            t_analyze = min(t_analyze + case.analyze_dt, tend)
        if do_finalize:
            break
    return {"t": t, "iterations": it, "states": written}


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("parameter_file")
    ap.add_argument("--node-type", default="gauss-lobatto", choices=["gauss", "gauss-lobatto"])
    ap.add_argument("--precision", default="fp64", choices=["fp64", "fp32-bulk"])
    ap.add_argument("--restart", default=None, help="State file to start from")
    ap.add_argument("--out-dir", default=".")
    ap.add_argument("--metric-order", default="reference", choices=["reference", "own"])
    args = ap.parse_args(argv)
    prm = parse_parameter_file(args.parameter_file)
    case = Case(prm, os.path.dirname(os.path.abspath(args.parameter_file)), args.node_type)
    if case.ignored:
        print("ignored parameters: " + ", ".join(case.ignored), file=sys.stderr)
    res = run(case, args.out_dir, args.precision, args.restart, metric_order=args.metric_order)
    print(f"t = {res['t']:.10g} after {res['iterations']} iterations; states: {', '.join(res['states'])}")
    return 0


if __name__ == "__main__":
    sys.exit(main())
