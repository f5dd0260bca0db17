"""Benchmark: DOF-stage updates per second of the DGSEM Navier-Stokes step on one MI355X.

Workload (one "step" = FLEXI's time-loop body: CALCTIMESTEP + TimeStepByLSERKW2 with the
5-stage Carpenter RK4-5, i.e. 5 right-hand sides and 5 register updates):
Taylor-Green vortex (Ma 0.1, Re 1600) on a periodic 32^3 hexahedral mesh, N = 7 Gauss-Lobatto
nodes (16,777,216 DOF), split form with the Pirozzoli flux, Roe Riemann solver, BR1 lifting.
DOF = one volume node (FLEXI's convention); value = DOF * stages * steps / wall time.

Also measured (keys of the same JSON line): the FP32-bulk rate and its error against FP64
("fp32_bulk"), the adjoint/forward time ratio of the kinetic-energy gradient on TGV 16^3 N=5
over 200 steps ("adjoint"), a curvilinear 24^3 N=5 case ("curvilinear"), a 32^3 N=7 cavity box
with walls and a Dirichlet lid ("boundaries"), and FLEXI State file write/read rates
("state_io").

Prints one JSON line. Options: --gpus (only 1 exists on this platform), --steps, --warmup,
--elems (per direction), --no-cpu (skip the CPU baseline), --no-fp32, --no-adjoint,
--no-curvilinear, --no-boundaries, --no-io.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0  # MI355X datasheet HBM3E bandwidth
# Algorithmic HBM bytes per DOF of each kernel launch (FP64, N=7, Gauss-Lobatto: 512 volume
# nodes, 296 boundary nodes and 6*64 own face points per element). Face values carry 40 B
# (5 doubles), viscous face fluxes 32 B, metrics 80 B (Ja^1..3 + sJ).
#   faces      : boundary U read 296*40 + own face states written 384*40       ->  53.1 B/DOF
#   visc_faces : U 512*40 + boundary metrics 296*80 + neighbour face states 384*40
#                + own viscous fluxes written 384*32                              -> 140.25 B/DOF
#   dg_fused   : U 40 + metrics 80 + neighbour states 30 + neighbour fluxes 24 + U written 40
#                + next face states 30 + register R read 40 (not in stage 1) + written 40
#                (stage 5 also forms CALCTIMESTEP's spectral radii from the metric terms, no
#                algorithmic bytes), averaged over the 5 stages: (284 + 4*324)/5  -> 316.0 B/DOF
# The four-kernel path's figures follow the same counting.
ALGO_BYTES_PER_DOF = {"prolong_u": 60.0, "lift": 372.0, "flux": 147.0, "volint": 366.0,
                      "faces": 53.125, "visc_faces": 140.25, "dg_fused": 316.0}
# Per DOF-stage byte model of the stage (U, Ut/R, metrics and gradient traffic of one stage)
B_STAGE = 432.0
# HBM bytes per DOF of one launch of the fused kernels at this workload (FP64, mean over the 5
# stages of a step), from rocprofv3 FETCH_SIZE (x2.0, plain-copy calibration) + WRITE_SIZE runs of
# this benchmark; the summary and the commands are in profiles/fused_stage_tgv32_n7_fp64.md
MEASURED_BYTES_PER_DOF = {"dg_fused": 364.1, "visc_faces": 143.3, "faces": 64.0}
# FP64 floating-point operations per DOF of one launch (mean over the 5 stages of a step): rocprofv3
# SQ_INSTS_VALU_{ADD,MUL,FMA,TRANS}_F64 per dispatch, wave instructions x 64 lanes (every wave of
# these kernels runs with all 64 lanes or skips a phase as a whole), FMA counted as 2; same summary
MEASURED_FLOP_PER_DOF = {"dg_fused": 2048.6, "visc_faces": 391.0}
FP64_PEAK_TFLOPS = 78.6
RHS_EVALS = 1000


def make_solver(nel, precision):
    from flexi_amd.mesh import periodic_box
    from flexi_amd.solver import DGSolver
    mesh = periodic_box(nel, nel, nel)
    s = DGSolver(mesh, 7, node_type="gauss-lobatto", split_dg="pi", riemann="roe", precision=precision,
                 CFLScale=0.8, DFLScale=0.8)
    return mesh, s


def run_gpu(nel, precision, steps, warmup, profile, rhs_evals=RHS_EVALS):
    mesh, s = make_solver(nel, precision)
    U0 = s_tgv(s)
    s.set_state(U0)
    del U0
    t = s.run(warmup)
    s.sync()
    if profile:
        s.profile(True)
    t0 = time.perf_counter()
    t = s.run(steps, t=t)
    s.sync()
    wall = time.perf_counter() - t0
    kt = s.kernel_times() if profile else None
    # RHS-only rate (DGTimeDerivative_weakForm alone), 1000 evaluations
    s.profile(False)
    s.sync()
    r0 = time.perf_counter()
    for _ in range(rhs_evals):
        s.DGTimeDerivative_weakForm(t)
    s.sync()
    rhs_wall = (time.perf_counter() - r0) / rhs_evals if rhs_evals > 0 else None
    # the TGV testcase analysis on the device (NAnalyze = 10 as in the TGV parameter files), after
    # the right-hand sides above, i.e. without a recomputation of the last stage's gradients
    a0 = time.perf_counter()
    for _ in range(10):
        vals = s.tgv_analyze(10)
    ana_wall = (time.perf_counter() - a0) / 10
    dof = mesh.nElems * 8 ** 3
    pipeline = s.kernel_pipeline()
    s.close()
    return dict(wall=wall, kernels=kt, dof=dof, rhs_s=rhs_wall, t=t, pipeline=pipeline, analyze_s=ana_wall,
                ekin=float(vals[2]))


def s_tgv(s):
    from flexi_amd import testcase
    xgp = s.metrics()["Elem_xGP"]
    return testcase.tgv(xgp)


def fp32_accuracy():
    """FP32-bulk against FP64 on TGV 8^3, N=3 Gauss, weak form, LF, density perturbed by 1e-3
    (seed 1), 1000 steps of the FP64 run's first time step: relative L2 error of U and of the
    kinetic-energy dissipation rate -dEkin/dt (from the right-hand side at the final state)."""
    from flexi_amd import testcase
    from flexi_amd.basis import nodes_weights
    from flexi_amd.mesh import periodic_box
    from flexi_amd.solver import DGSolver
    mesh = periodic_box(8, 8, 8)
    res, dt = {}, None
    for prec in ("fp64", "fp32-bulk"):
        s = DGSolver(mesh, 3, node_type="gauss", split_dg=None, riemann="lf", precision=prec)
        m = s.metrics()
        s.set_state(testcase.tgv_density_perturbed(m["Elem_xGP"], amplitude=1e-3, seed=1))
        if dt is None:
            dt = s.CalcTimeStep()[0]
        s.run(1000, fixed_dt=dt)
        U = s.get_state()
        s.DGTimeDerivative_weakForm(1000 * dt)
        res[prec] = (U, s.get_ut(), m["sJ"])
        s.close()
    (U64, Ut64, sJ), (U32, Ut32, _) = res["fp64"], res["fp32-bulk"]
    w = nodes_weights(3, 1)[1]
    W = (w[:, None, None] * w[None, :, None] * w[None, None, :])[..., None] / sJ

    def dek(U, Ut):
        u = U[1:4] / U[0]
        return -np.sum(W * (np.sum(u * Ut[1:4], axis=0) - 0.5 * np.sum(u * u, axis=0) * Ut[0]))
    d64, d32 = dek(U64, Ut64), dek(U32, Ut32)
    return {"rel_l2_U_vs_fp64": float(np.linalg.norm(U32 - U64) / np.linalg.norm(U64)),
            "rel_err_dEkin_dt_vs_fp64": float(abs(d32 - d64) / abs(d64)),
            "accuracy_case": "TGV 8^3, N=3 Gauss, weak, LF, rho perturbed 1e-3, 1000 steps"}


def adjoint_ratio():
    """Gradient of the kinetic energy at T w.r.t. U(0): TGV 16^3, N=5 Gauss-Lobatto split PI, Roe,
    200 steps (dt frozen from the primal run). `ratio` = time of the adjoint (backward) call / time
    of the 200 primal forward steps, with checkpoints every 20 steps, the backward
    recomputing the stage states of each segment; `trajectory` = the same with the whole trajectory
    stored by the forward in HBM (71 GB: every stage state and its face states), no recomputation."""
    import torch
    from flexi_amd import testcase
    from flexi_amd.adjoint import primal_dts, to_tensor
    from flexi_amd.mesh import periodic_box
    from flexi_amd.solver import DGSolver
    every = 20
    mesh = periodic_box(16, 16, 16)
    s = DGSolver(mesh, 5, node_type="gauss-lobatto", split_dg="pi", riemann="roe")
    U0 = testcase.tgv(s.metrics()["Elem_xGP"])
    dts = primal_dts(s, U0, 200)
    U0d = to_tensor(U0)
    ckpt = torch.empty((len(dts) // every,) + tuple(U0d.shape), dtype=U0d.dtype, device=U0d.device)

    def timed(fn):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        fn()
        s.sync()
        torch.cuda.synchronize()
        return time.perf_counter() - t0
    U = U0d.clone()
    timed(lambda: s.forward_n_steps_dev(U.data_ptr(), dts[:5]))  # warm-up
    ubar = torch.ones_like(U0d)
    timed(lambda: s.adjoint_n_steps_dev(U0d.data_ptr(), ubar.data_ptr(), dts[:5], every))
    U = U0d.clone()
    t_fwd = timed(lambda: s.forward_n_steps_dev(U.data_ptr(), dts))
    U = U0d.clone()
    t_fwd_ck = timed(lambda: s.forward_n_steps_dev(U.data_ptr(), dts, ckpt.data_ptr(), every))
    ubar = torch.ones_like(U0d)
    t_bwd = timed(lambda: s.adjoint_n_steps_dev(U0d.data_ptr(), ubar.data_ptr(), dts, every, ckpt.data_ptr()))
    g_ck = ubar.clone()
    del ckpt
    # trajectory mode: the forward stores every stage state and its face states (out of place, no
    # extra bytes moved), the backward runs the reverse stages only
    traj = torch.empty(s.adjoint_storage(len(dts), 0), dtype=U0d.dtype, device=U0d.device)
    U = U0d.clone()
    t_fwd_tr = timed(lambda: s.forward_n_steps_dev(U.data_ptr(), dts, traj.data_ptr(), 0))
    ubar = torch.ones_like(U0d)
    t_bwd_tr = timed(lambda: s.adjoint_n_steps_dev(U0d.data_ptr(), ubar.data_ptr(), dts, 0, traj.data_ptr()))
    same = bool(torch.equal(ubar, g_ck))
    traj_bytes = traj.numel() * 8
    del traj
    state = U0.nbytes
    face_states = 10 * (5 + 1) ** 2 * mesh.nSides * 8  # both slots of every side, 5 doubles per point
    s.close()
    return {"ratio": t_bwd / t_fwd, "mode": "checkpoints every 20 steps, segments recomputed",
            "forward_s": t_fwd, "forward_with_checkpoints_s": t_fwd_ck, "adjoint_s": t_bwd,
            "gradient_total_ratio": (t_fwd_ck + t_bwd) / t_fwd,
            "memory_bytes": {"checkpoints": (len(dts) // every) * state, "segment_stage_states": every * 5 * state,
                             "segment_face_states": every * 5 * face_states, "work": 3 * state},
            "trajectory": {"ratio": t_bwd_tr / t_fwd, "mode": "trajectory stored by the forward (HBM), no "
                           "recomputation", "forward_with_trajectory_s": t_fwd_tr, "adjoint_s": t_bwd_tr,
                           "gradient_total_ratio": (t_fwd_tr + t_bwd_tr) / t_fwd, "trajectory_bytes": traj_bytes,
                           "gradient_bitwise_equal_to_checkpointed": same},
            "case": "TGV 16^3, N=5 GL split PI Roe, 200 steps, dt frozen from the primal run"}


def boundaries(nel=32, steps=20, warmup=3):
    """Lid-driven cavity box: nel^3 elements on the unit cube, N=7 Gauss-Lobatto split PI, Roe,
    BR1, isothermal walls and a Dirichlet lid (y+, u=1, lid Mach 0.1), from rest, on the fused
    path (boundary sides through the outer face slots). DOF-stage/s, kernel times (the boundary
    kernel is timed with the viscous face kernel it precedes) and the roofline of the dominant
    kernel; FXG_PIPELINE=classic measures the four-kernel path instead."""
    from flexi_amd import equation
    from flexi_amd.mesh import walled_box
    from flexi_amd.solver import DGSolver
    mesh = walled_box(nel, nel, nel, boundary_types={"BC_yplus": (2, 1)})
    refs = equation.ref_state_prim([[1.0, 1.0, 0.0, 0.0, 71.4285714286], [1.0, 0.0, 0.0, 0.0, 71.4285714286]], 1.0)
    s = DGSolver(mesh, 7, node_type="gauss-lobatto", split_dg="pi", riemann="roe", R=1.0, mu0=0.01,
                 ref_state_prim=refs, CFLScale=0.8, DFLScale=0.8)
    U0 = equation.prim_to_cons(refs[:, 1], 1.4).reshape(5, 1, 1, 1, 1)
    s.set_state(np.asfortranarray(np.broadcast_to(U0, s.shape)))
    s.run(warmup)
    s.sync()
    s.profile(True)
    t0 = time.perf_counter()
    s.run(steps)
    s.sync()
    wall = time.perf_counter() - t0
    kt = s.kernel_times()
    pipeline = s.kernel_pipeline()
    s.close()
    dof = mesh.nElems * 8 ** 3
    kern = {k: {"avg_ms": v[0] / v[1], "launches": v[1]} for k, v in kt.items() if v[1]}
    main = {k: v for k, v in kern.items() if k in ALGO_BYTES_PER_DOF}
    dom = max(main, key=lambda k: main[k]["avg_ms"] * main[k]["launches"])
    achieved = ALGO_BYTES_PER_DOF[dom] * dof / (main[dom]["avg_ms"] * 1e-3) / 1e9
    return {"value": dof * 5 * steps / wall, "unit": "DOF-stage/s", "ms_per_step": 1e3 * wall / steps,
            "pipeline": pipeline, "boundary_sides": mesh.nBCSides, "sides": mesh.nSides, "kernels": kern,
            "roofline": {"bound": "hbm", "kernel": dom, "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                         "frac": achieved / HBM_PEAK_GBS, "algorithmic_bytes_per_dof": ALGO_BYTES_PER_DOF[dom]},
            "case": f"cavity box {nel}^3, N=7 GL split PI Roe BR1, isothermal walls + Dirichlet lid, {steps} steps"}


def state_io(nel=32):
    """FLEXI State file of the nel^3 N=7 state (5 doubles per DOF) written and read back through
    libfxg_io.so on the host's /tmp: bytes per second each way (host I/O, not HBM)."""
    import shutil
    import tempfile
    from flexi_amd import io
    dof = nel ** 3 * 8 ** 3
    U = np.asfortranarray(np.random.default_rng(0).normal(size=(5, 8, 8, 8, nel ** 3)))
    d = tempfile.mkdtemp(prefix="fxg_io_")
    try:
        p = os.path.join(d, "bench_State_0000000.000000000.h5")
        t0 = time.perf_counter()
        io.write_state(p, U, 7, "gauss-lobatto", 0.0, "bench_mesh.h5", project="bench")
        tw = time.perf_counter() - t0
        t0 = time.perf_counter()
        st = io.read_state(p)  # into a fresh array: includes the first touch of its pages
        tr_fresh = time.perf_counter() - t0
        t0 = time.perf_counter()
        st = io.read_state(p, out=st["U"])  # into the host's existing array
        tr = time.perf_counter() - t0
        same = bool(np.array_equal(st["U"], U))
    finally:
        shutil.rmtree(d, ignore_errors=True)
    nbytes = 5 * 8 * dof
    return {"bytes": nbytes, "write_GBps": nbytes / tw / 1e9, "read_GBps": nbytes / tr / 1e9,
            "read_fresh_array_GBps": nbytes / tr_fresh / 1e9, "round_trip_exact": same,
            "case": f"DG_Solution of {nel}^3 elements, N=7, written and read with FLEXI's attributes"}


def curvilinear():
    """24^3 periodic box mapped by x_i + 0.1 sin x sin y sin z, N=5 Gauss-Lobatto split PI, Roe,
    metrics on the GPU: freestream preservation after 100 steps, then a Gaussian density pulse
    (amplitude 0.1, width 0.5) on that freestream, 50 timed steps."""
    from flexi_amd import testcase
    from flexi_amd.mesh import periodic_box, sine_deform
    from flexi_amd.solver import DGSolver
    N = 5
    mesh = periodic_box(24, 24, 24, NGeo=N, deform=sine_deform(0.1))
    s = DGSolver(mesh, N, node_type="gauss-lobatto", split_dg="pi", riemann="roe")
    n = N + 1
    U0 = testcase.freestream((n, n, n, mesh.nElems), vel=(0.3, 0.2, 0.1), p=1.0 / 1.4)
    s.set_state(U0)
    s.run(100)
    fs_err = float(np.max(np.abs(s.get_state() - U0)))
    s.set_state(testcase.gaussian_pulse(s.metrics()["Elem_xGP"]))
    s.run(5)
    s.sync()
    t0 = time.perf_counter()
    s.run(50)
    s.sync()
    wall = time.perf_counter() - t0
    dof = mesh.nElems * n ** 3
    s.close()
    return {"freestream_max_abs_err_100_steps": fs_err, "pulse_value": dof * 5 * 50 / wall,
            "unit": "DOF-stage/s", "dof": dof, "ms_per_step": 1e3 * wall / 50}


def cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(nel):
    """The CPU oracle (a C++/OpenMP restatement of the reference, parallel over elements) built for
    this host with -O3 -march=native (oracle/Makefile `native`, FMA contraction allowed; the
    parity build is separate) and timed on its cores: after one warm-up, the median of 5 runs of
    each case. Config 2 (N=7 GL split PI, Roe, BR1, TGV): the right-hand side alone on the full
    nel^3 mesh, and whole CarpenterRK4-5 steps (CALCTIMESTEP + 5 stages) on a 16^3 sample of the
    same scheme; the plain method (TGV 8^3, N=3 Gauss, weak form, LF, BR1): the same two figures,
    steps in runs of 50. Also on one core (the N=7 scheme on an 8^3 sample). DOF-stage/s."""
    import statistics
    import tempfile
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as orc
    from flexi_amd import testcase
    from flexi_amd.mesh import periodic_box
    from flexi_amd.timedisc import fill_cfl_dfl
    variant, march = "native", "native"
    try:
        orc.load_native(tempfile.mkdtemp(prefix="fxg_oracle_"))
    except Exception as exc:  # no compiler on this host: the portable -O3 build of the same sources
        print(f"[bench] cpu baseline: native oracle build failed ({exc}); using liboracle_fast.so",
              file=sys.stderr, flush=True)
        variant, march = "fast", "x86-64-v3"
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))

    def med(fn, reps=5):
        fn()  # warm-up
        ts = []
        for _ in range(reps):
            t0 = time.perf_counter()
            fn()
            ts.append(time.perf_counter() - t0)
        return statistics.median(ts)

    def case(N, node, split, riemann, nel_rhs, nel_step, steps_per_run, nthreads):
        orc.set_threads(nthreads, variant)
        prm = orc.params(N, node, split, riemann)
        res = {}
        for what, ne in (("rhs", nel_rhs), ("step", nel_step)):
            mesh = periodic_box(ne, ne, ne)
            m = orc.calc_metrics(N, node, mesh)
            S = orc.Solver(prm, mesh, m, variant=variant)
            U = testcase.tgv(m["Elem_xGP"])
            dof = mesh.nElems * (N + 1) ** 3
            if what == "rhs":
                t = med(lambda: S.rhs(U))
                res["rhs"] = {"value": dof / t, "unit": "DOF/s (one RHS)", "median_s": t, "mesh": f"{ne}^3"}
            else:
                cfl, dfl = fill_cfl_dfl(N, node, 0.8 if split else 0.9, 0.8 if split else 0.9)
                R = orc.F(U.shape)
                Uw = np.asfortranarray(U.copy())

                def run():
                    for _ in range(steps_per_run):
                        dt, err, _ = S.calc_timestep(Uw, cfl, dfl)
                        S.step(Uw, 0.0, dt, R)
                t = med(run)
                res["step"] = {"value": dof * 5 * steps_per_run / t, "unit": "DOF-stage/s", "median_s": t,
                               "mesh": f"{ne}^3", "steps_per_run": steps_per_run}
            del S
        return res

    c2 = case(7, orc.NODE_GL, 1, orc.RIEMANN_ROE, nel, 16, 1, threads)
    c1 = case(3, orc.NODE_G, 0, orc.RIEMANN_LF, 8, 8, 50, threads)
    one = case(7, orc.NODE_GL, 1, orc.RIEMANN_ROE, 8, 8, 1, 1)
    orc.set_threads(threads, variant)
    return dict(value=c2["step"]["value"], unit="DOF-stage/s", cores=threads, kind="port",
                sample=f"TGV, N=7 GL split PI, Roe, BR1: CarpenterRK4-5 steps with CALCTIMESTEP "
                       f"on a 16^3 sample, median of 5 after a warm-up, OpenMP over elements",
                cpu_model=cpu_model(), nproc=os.cpu_count(), build=f"g++ -O3 -march={march} -fopenmp (oracle/Makefile)",
                tgv_n7_split=c2, tgv8_n3_weak_lf=c1,
                single_core={"tgv_n7_split": one, "cores": 1})


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--elems", type=int, default=32)
    ap.add_argument("--rhs-evals", type=int, default=RHS_EVALS,
                    help="right-hand sides timed alone after the steps (default 1000)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-fp32", action="store_true")
    ap.add_argument("--no-adjoint", action="store_true")
    ap.add_argument("--no-curvilinear", action="store_true")
    ap.add_argument("--no-boundaries", action="store_true")
    ap.add_argument("--no-io", action="store_true")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="BYTES",
                    help="only run a plain device-to-device copy of BYTES (calibrates HBM counters)")
    ap.add_argument("--adjoint-only", action="store_true",
                    help="only run the adjoint/forward measurement and print its JSON")
    args = ap.parse_args()
    if args.adjoint_only:
        print(json.dumps(adjoint_ratio()), flush=True)
        return
    if args.copy_probe:
        import torch
        n = args.copy_probe // 8
        a = torch.ones(n, dtype=torch.float64, device="cuda")
        b = torch.empty_like(a)
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_probe_bytes": 8 * n, "dispatches": 3}))
        return
    if args.gpus != 1:
        print(f"bench.py: --gpus {args.gpus} is not supported: this build targets one MI355X per run "
              f"(the multi-GPU face exchange is a design, not an implementation)", file=sys.stderr)
        sys.exit(2)

    print(f"[bench] fp64 {args.elems}^3 N=7: {args.warmup} warm-up + {args.steps} timed steps", file=sys.stderr,
          flush=True)
    r64 = run_gpu(args.elems, "fp64", args.steps, args.warmup, profile=True, rhs_evals=args.rhs_evals)
    stages = 5
    value = r64["dof"] * stages * args.steps / r64["wall"]
    ms_per_step = 1e3 * r64["wall"] / args.steps
    kern = {k: {"avg_ms": (v[0] / v[1] if v[1] else None), "launches": v[1]} for k, v in r64["kernels"].items()
            if v[1]}
    # dominant kernel and its roofline (algorithmic bytes / measured average launch time)
    rhs_k = {k: v for k, v in r64["kernels"].items() if k in ALGO_BYTES_PER_DOF and v[1]}
    dom = max(rhs_k, key=lambda k: rhs_k[k][0])
    avg_s = rhs_k[dom][0] / rhs_k[dom][1] * 1e-3
    achieved = ALGO_BYTES_PER_DOF[dom] * r64["dof"] / avg_s / 1e9
    stage_s = sum(v[0] for k, v in rhs_k.items()) / (args.steps * stages) * 1e-3
    stage_achieved = B_STAGE * r64["dof"] / stage_s / 1e9
    out = {
        "metric": "DOF-stage updates/s at N=7 (FP64 vs FP32-bulk) + adjoint/forward time ratio",
        "value": value,
        "unit": "DOF-stage/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (Taylor-Green vortex initial state, Ma 0.1)",
        "config": {"workload": f"TGV {args.elems}^3 hex, N=7 Gauss-Lobatto, split-form PI, Roe, BR1, "
                               f"CarpenterRK4-5 with CALCTIMESTEP each step",
                   "dof": r64["dof"], "stages_per_step": stages, "parallelism": "single GPU",
                   "pipeline": r64["pipeline"]},
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS,
                     "traffic": (MEASURED_BYTES_PER_DOF[dom] * r64["dof"]
                                 if dom in MEASURED_BYTES_PER_DOF and args.elems == 32 else None),
                     "algorithmic_bytes_per_dof": ALGO_BYTES_PER_DOF[dom], "avg_launch_ms": avg_s * 1e3,
                     "flops": ({"achieved": MEASURED_FLOP_PER_DOF[dom] * r64["dof"] / avg_s / 1e12,
                                "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                                "frac": MEASURED_FLOP_PER_DOF[dom] * r64["dof"] / avg_s / 1e12 / FP64_PEAK_TFLOPS,
                                "flop_per_dof": MEASURED_FLOP_PER_DOF[dom],
                                "source": "rocprofv3 SQ_INSTS_VALU_*_F64 x 64 lanes, FMA = 2"}
                               if dom in MEASURED_FLOP_PER_DOF else None)},
        "roofline_stage": {"bytes_per_dof_stage": B_STAGE, "stage_ms": stage_s * 1e3, "achieved": stage_achieved,
                           "unit": "GB/s", "frac": stage_achieved / HBM_PEAK_GBS,
                           "dof_stage_per_s_at_peak": HBM_PEAK_GBS * 1e9 / B_STAGE},
        "kernels": kern,
        "rhs_only": ({"value": r64["dof"] / r64["rhs_s"], "unit": "DOF/s (one RHS)", "ms": r64["rhs_s"] * 1e3,
                      "evaluations": args.rhs_evals} if r64["rhs_s"] else None),
        "tgv_analysis": {"ms": r64["analyze_s"] * 1e3, "n_analyze": 10, "ekin": r64["ekin"],
                         "case": "AnalyzeTestcase on the device (fxg_tgv_analyze), wall clock per call"},
    }
    if not args.no_fp32:
        print("[bench] fp32-bulk", file=sys.stderr, flush=True)
        r32 = run_gpu(args.elems, "fp32-bulk", args.steps, args.warmup, profile=False, rhs_evals=args.rhs_evals)
        out["fp32_bulk"] = {"value": r32["dof"] * stages * args.steps / r32["wall"], "unit": "DOF-stage/s",
                            "ms_per_step": 1e3 * r32["wall"] / args.steps, **fp32_accuracy()}
    if not args.no_adjoint:
        print("[bench] adjoint (TGV 16^3 N=5, 200 steps)", file=sys.stderr, flush=True)
        adj = adjoint_ratio()
        # the headline ratio is the checkpointed mode (every 20 steps), the trajectory mode is secondary
        out["adjoint_forward_ratio"] = adj["ratio"]
        out["adjoint_forward_ratio_trajectory"] = adj["trajectory"]["ratio"]
        out["adjoint"] = adj
    else:
        out["adjoint_forward_ratio"] = None
    if not args.no_curvilinear:
        print("[bench] curvilinear 24^3 N=5", file=sys.stderr, flush=True)
        out["curvilinear"] = curvilinear()
    if not args.no_boundaries:
        print("[bench] boundaries (cavity box 32^3 N=7)", file=sys.stderr, flush=True)
        out["boundaries"] = boundaries()
    if not args.no_io:
        print("[bench] State file I/O", file=sys.stderr, flush=True)
        out["state_io"] = state_io(args.elems)
    if not args.no_cpu:
        print("[bench] cpu baseline (oracle, native build)", file=sys.stderr, flush=True)
        out["cpu_baseline"] = cpu_baseline(args.elems)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
