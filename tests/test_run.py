"""The parameter-file driver flexi_amd.run (FLEXI's TimeDisc loop with State output).

tests/golden/params/tgv_split_parameter.ini is an unmodified copy of the reference's
regressioncheck/checks/tgv/split/parameter.ini (the TGV at Re 1600, N=7, split PI,
RoeEntropyFix, on the CART_HEX_PERIODIC_008 mesh of tests/golden/hdf5/).
"""
import os

import numpy as np
import pytest

from flexi_amd import run as frun

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PRM = os.path.join(ROOT, "tests", "golden", "params", "tgv_split_parameter.ini")
MESH_DIR = os.path.join(ROOT, "tests", "golden", "hdf5")


def tgv_case(**over):
    case = frun.Case(frun.parse_parameter_file(PRM), MESH_DIR)
    for k, v in over.items():
        setattr(case, k, v)
    return case


def test_reads_the_tgv_regression_parameter_file():
    c = tgv_case()
    assert c.project == "TGV_Re1600_Split" and c.mesh_file == "CART_HEX_PERIODIC_008_mesh.h5"
    assert os.path.exists(c.mesh_path)
    assert (c.N, c.riemann, c.split) == (7, "roeentropyfix", "pi")
    assert (c.mu0, c.R, c.kappa, c.Pr) == (6.25e-4, 71.42857, 1.4, 0.72)
    assert (c.tend, c.CFL, c.DFL, c.analyze_dt, c.n_write, c.max_iter) == (13.0, 0.8, 0.8, 13.0, 1, -1)
    assert c.ini_exact == 0 and c.ini_ref == 1 and c.mach == 0.1
    assert c.ref_states == [[1.0, 1.0, 0.0, 0.0, 17194.8345650329]]
    assert c.scheme.nStages == 5  # CarpenterRK4-5, FLEXI's default TimeDiscMethod
    # keys of the file the driver does not act on are reported, none of the ones it needs
    assert (c.n_analyze_points, c.n_analyze_testcase) == (10, 10) and "eddyvisctype" in c.ignored
    assert not set(c.ignored) & frun.USED


def test_parameter_syntax(tmp_path):
    p = tmp_path / "p.ini"
    p.write_text("! comment\nN = 3 ! trailing\nRefState=(/1.,0.5d0,0.,0.,1E1/)\nrefstate = (/2,0,0,0,3/)\n"
                 "MeshFile = m.h5\nTEnd = 1.5D-1\nCFLScale=0.9\nDFLScale=0.9\nAnalyze_dt=0.\n")
    prm = frun.parse_parameter_file(str(p))
    assert prm["n"] == ["3"] and len(prm["refstate"]) == 2
    c = frun.Case(prm, str(tmp_path))
    assert c.ref_states == [[1.0, 0.5, 0.0, 0.0, 10.0], [2.0, 0.0, 0.0, 0.0, 3.0]]
    assert c.tend == 0.15 and c.mesh_path == str(tmp_path / "m.h5")
    # analyze_dt <= 0: analysis (and output) only at the end (analyze.f90:124-130)
    assert c.analyze_dt == 0.15 and c.n_write == 1
    assert c.split is None and c.riemann == "roeentropyfix" and c.mu0 == 0.0 and c.R == 287.058
    with pytest.raises(KeyError):
        frun.Case({"n": ["3"]}, str(tmp_path))


def test_timestamp_is_flexis():
    # the file name of the State the reference holds for t = 0 (tests/golden/hdf5)
    assert f"cartbox_State_{frun.timestamp(0.0)}.h5" == "cartbox_State_0000000.000000000.h5"
    assert frun.timestamp(13.0) == "0000013.000000000"
    assert frun.timestamp(0.123456789) == "0000000.123456789"
    assert frun.timestamp(1234.5) == "0001234.500000000"


def test_update_time_step_rules():
    f = frun.update_time_step
    # the stable step, far from the analysis and the end
    assert f(0.1, 0.0, 1.0, 2.0) == (0.1, False, False)
    # the next analysis is nearer than the stable step: hit it
    assert f(0.1, 0.95, 1.0, 2.0) == (pytest.approx(0.05), True, False)
    # a step that leaves less than dt/100 before the analysis is stretched to reach it
    dt, an, fin = f(0.1, 0.0, 0.1005, 2.0)
    assert dt == 0.1005 and an and not fin
    # ... but not one that leaves more
    assert f(0.1, 0.0, 0.102, 2.0) == (0.1, False, False)
    # the end: analysis and finalize
    dt, an, fin = f(0.1, 1.97, 2.0, 2.0)
    assert dt == pytest.approx(0.03) and an and fin
    dt, an, fin = f(0.1, 1.9 - 5e-4, 2.0, 2.0)
    assert dt == pytest.approx(0.1005) and an and fin


def test_uniform_refstate_initial_state():
    c = tgv_case(ini_exact=1)
    U = c.initial_state(np.zeros((3, 2, 2, 2, 3)))
    p = 17194.8345650329
    assert np.allclose(U[:, 0, 0, 0, 0], [1.0, 1.0, 0.0, 0.0, p / 0.4 + 0.5])


@pytest.mark.gpu
def test_driver_runs_the_tgv_case_and_restarts_bitwise(tmp_path):
    """The TGV regression case on the GPU for a short time: State files at the start, at every
    analysis time and at the end with FLEXI's names; the initial file holds the TGV initial
    condition; a run restarted from the middle file ends in the same state, bit for bit, as the
    run straight through (the restart reads the state and its time exactly and the dt sequence
    depends only on the state); and maxIter steps equal as many UpdateTimeStep steps of the
    solver (DGSolver.run, CALCTIMESTEP each step)."""
    from flexi_amd import io, testcase
    from flexi_amd.solver import DGSolver

    a = tmp_path / "a"
    b = tmp_path / "b"
    a.mkdir()
    b.mkdir()
    log = []
    case = tgv_case(tend=0.05, analyze_dt=0.02)
    res = frun.run(case, str(a), log=log.append)
    names = [os.path.basename(p) for p in res["states"]]
    print(names, log)
    assert names == [f"TGV_Re1600_Split_State_{frun.timestamp(t)}.h5" for t in (0.0, 0.02, 0.04, 0.05)]
    assert len(log) == 3 and res["t"] == pytest.approx(0.05, abs=1e-15)
    s0 = io.read_state(res["states"][0])
    s2 = io.read_state(res["states"][2])
    assert s0["time"] == 0.0 and s0["node_type"] == "GAUSS-LOBATTO" and s0["N"] == 7
    assert s0["mesh_file"] == "CART_HEX_PERIODIC_008_mesh.h5"
    mesh = io.read_mesh(case.mesh_path)
    ref = DGSolver(mesh, 7, split_dg="pi", riemann="roeentropyfix", mu0=6.25e-4, R=71.42857,
                   CFLScale=0.8, DFLScale=0.8, metric_order="reference")  # the driver's metric terms
    try:
        xgp = ref.metrics()["Elem_xGP"]
        assert np.array_equal(s0["U"], testcase.tgv(xgp, R=71.42857, Ma=0.1))
        # kinetic energy of the TGV decays slowly over this short time
        assert np.all(np.isfinite(s2["U"])) and np.max(np.abs(s2["U"] - s0["U"])) < 0.1

        # restart from the file at t=0.02 and run to 0.05
        resb = frun.run(tgv_case(tend=0.05, analyze_dt=0.02), str(b), restart=res["states"][1], log=log.append)
        assert [os.path.basename(p) for p in resb["states"]] == names[1:]
        for pa, pb in zip(res["states"][1:], resb["states"]):
            ua, ub = io.read_state(pa), io.read_state(pb)
            assert ua["time"] == ub["time"]
            assert np.array_equal(ua["U"], ub["U"]), os.path.basename(pa)

        # maxIter = 12 with no analysis before the end: 12 CALCTIMESTEP steps
        c = tmp_path / "c"
        c.mkdir()
        resc = frun.run(tgv_case(max_iter=12), str(c), log=log.append)
        assert resc["iterations"] == 12 and len(resc["states"]) == 2
        ref.set_state(s0["U"])
        t = ref.run(12)
        sc = io.read_state(resc["states"][-1])
        assert sc["time"] == t
        assert np.array_equal(sc["U"], ref.get_state())
    finally:
        ref.close()


@pytest.mark.gpu
def test_driver_writes_flexis_tgv_analysis_file(tmp_path):
    """The reference's tgv/split case through the driver on the GPU for 21 iterations: the
    TGVAnalysis.csv lines at iterations 0, 10 and 20 (analysis every nAnalyzeTestCase = 10 steps,
    between analyses with the last stage's gradients as in FLEXI) equal the reference CSV's first
    three lines to 1e-9 (the regression check itself uses 1e-4), columns that vanish compared
    absolutely; the file has FLEXI's header."""
    import sys
    sys.path.insert(0, os.path.join(ROOT, "tests", "golden"))
    import flexi_fixtures as fx
    header, data = fx.tgv_split_csv()
    res = frun.run(tgv_case(max_iter=21), str(tmp_path), log=lambda s: None)
    assert res["iterations"] == 21
    path = tmp_path / "TGV_Re1600_Split_TGVAnalysis.csv"
    lines = path.read_text().splitlines()
    assert lines[0] == ",".join(header)
    got = np.array([[float(v) for v in ln.split(",")] for ln in lines[1:]])
    assert got.shape == (4, 16)  # iterations 0, 10, 20 and the last one, 21
    for r in range(3):
        for c in range(16):
            ref = data[r, c]
            tol = 1e-12 if abs(ref) < 1e-6 else 1e-9 * abs(ref)
            assert abs(got[r, c] - ref) <= tol, (r, header[c], got[r, c], ref)


def whole_tgv_split_case(tmp_path, variant):
    """FLEXI's regression case tgv/split through the driver in a process of its own that loads the
    library build `variant` (FXG_LIB): all iterations to t = 13. Returns the TGVAnalysis.csv values,
    their relative deviation from the reference's 554 lines and the header."""
    import subprocess
    import sys
    sys.path.insert(0, os.path.join(ROOT, "tests", "golden"))
    import flexi_fixtures as fx
    header, ref = fx.tgv_split_csv()
    out = tmp_path / variant
    out.mkdir()
    code = ("import sys; from flexi_amd import run as frun; "
            "case = frun.Case(frun.parse_parameter_file(sys.argv[1]), sys.argv[2]); "
            "res = frun.run(case, sys.argv[3], log=lambda s: None); print(res['iterations'])")
    print(f"[case] tgv/split to t=13 with FXG_LIB={variant}", flush=True)
    r = subprocess.run([sys.executable, "-c", code, PRM, MESH_DIR, str(out)], capture_output=True, text=True,
                       timeout=600, env={**os.environ, "FXG_LIB": variant}, cwd=ROOT)
    assert r.returncode == 0, r.stderr[-3000:]
    lines = (out / "TGV_Re1600_Split_TGVAnalysis.csv").read_text().splitlines()
    assert lines[0] == ",".join(header)
    got = np.array([[float(v) for v in ln.split(",")] for ln in lines[1:]])
    assert got.shape == ref.shape == (554, 16)
    rel = np.abs(got - ref) / np.maximum(np.abs(ref), 1e-300)
    ip = header.index("DR_p")
    # DR_p at t = 0 is the round-off zero of an integral that vanishes for the divergence-free
    # initial state (5.6e-18 in the reference): compared absolutely
    assert abs(got[0, ip] - ref[0, ip]) <= 1e-16
    rel[0, ip] = 0.0
    worst = {header[c]: (float(rel[:, c].max()), int((rel[:, c] > 1e-4).sum())) for c in range(16)}
    print(f"[case] {variant}: {r.stdout.strip()} iterations; max rel deviation per column (lines > 1e-4): "
          + ", ".join(f"{k} {v[0]:.1e} ({v[1]})" for k, v in worst.items()), flush=True)
    return got, rel, header


@pytest.mark.gpu
@pytest.mark.parametrize("variant", ["strict", "default"])
def test_whole_tgv_split_regression_case(tmp_path, variant):
    """FLEXI's regression check tgv/split end to end on the GPU through the driver, with the
    library built as FLEXI's Fortran is evaluated by default (libfxg_strict.so, -ffp-contract=off)
    and with the default build (FMA contraction, the benchmark's): all 5527 iterations to t = 13,
    every value of the TGVAnalysis.csv within the check's own tolerance, 1e-4 relative (its
    analyze.ini), on all 554 lines and 16 columns. The driver computes the metric terms in
    metrics.f90's operation order (FXG_METRICS_REFERENCE, measured maximum 2.6e-5 in the Maximum
    Vorticity). This case runs at Mach 0.1, where the pressure is 140 times the dynamic pressure: a
    difference of 1e-13 in the metric terms (the library's own curl form, FXG_METRICS_OWN, with
    coordinates relative to each element's first node) moves every right-hand side by 3e-9 in the
    same direction and the Maximum Vorticity by 2.3e-4 at t = 12.4."""
    got, rel, header = whole_tgv_split_case(tmp_path, variant)
    assert not (rel > 1e-4).any(), [(int(i), header[c], float(rel[i, c])) for i, c in np.argwhere(rel > 1e-4)]
