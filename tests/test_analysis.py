"""The TGV testcase analysis on the device (fxg_tgv_analyze) and its CSV output.

The oracle's AnalyzeTestcase reproduces the reference's tgv/split TGVAnalysis CSV (its own test).
Here the device analysis is checked against the oracle on the device's own state and gradients
(so that only the analysis differs), after a right-hand side and after an LSERK step (the
gradients of the last stage), on straight and curved meshes, both node types, FP64 and
FP32-bulk gradients, the smallest and largest NAnalyze; the CSV number format against the
reference CSV's own text.
"""
import os

import numpy as np
import pytest

import oracle as orc
from flexi_amd import analysis, testcase
from flexi_amd.mesh import periodic_box, sine_deform

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSV = os.path.join(ROOT, "tests", "golden", "tgv", "TGV_Re1600_Split_TGVAnalysis_Reference.csv")


def _close(got, ref, what):
    for c, (a, b) in enumerate(zip(got, ref)):
        # relative 1e-12, with an absolute floor for the divergence columns (DR_p, ED_D), which are
        # sums of O(1) terms p*divU that cancel to 1e-6 and below, their rounding amplified alike
        tol = 1e-12 * abs(b) + 1e-13
        assert abs(a - b) <= tol, (what, analysis.TGV_VARNAMES[c], a, b)


@pytest.mark.gpu
@pytest.mark.parametrize("N,node,split,prec,NA,curved", [
    (3, "gauss-lobatto", "pi", "fp64", 6, False),
    (4, "gauss", None, "fp64", 9, True),
    (7, "gauss-lobatto", "pi", "fp64", 10, True),
    (7, "gauss-lobatto", "pi", "fp32-bulk", 10, False),
    (5, "gauss-lobatto", "pi", "fp64", 31, False),
    (1, "gauss", None, "fp64", 1, False),
])
def test_device_tgv_analysis_matches_oracle(N, node, split, prec, NA, curved):
    from flexi_amd.solver import DGSolver
    print(f"[case] N={N} node={node} prec={prec} NAnalyze={NA} curved={curved}", flush=True)
    mesh = periodic_box(3, 3, 2, lo=(0, 0, 0), NGeo=2 if curved else 1, deform=sine_deform(0.1) if curved else None)
    onode = orc.NODE_GL if node == "gauss-lobatto" else orc.NODE_G
    m = orc.calc_metrics(N, onode, mesh)
    S = orc.Solver(orc.params(N, onode, 1 if split else 0, orc.RIEMANN_ROEENTROPYFIX), mesh, m)
    s = DGSolver(mesh, N, node_type=node, split_dg=split, riemann="roeentropyfix", precision=prec, metrics=m)
    U = testcase.tgv(m["Elem_xGP"], Ma=0.5)
    U = np.asfortranarray(U * (1.0 + 1e-3 * np.sin(3.0 * m["Elem_xGP"][0]))[None])  # break the symmetry
    s.set_state(U)
    s.DGTimeDerivative_weakForm(0.0)
    _close(s.tgv_analyze(NA), S.tgv_analyze(s.get_state(), s.gradients(), NAnalyze=NA), "rhs")
    s.dt = 2e-3
    s.TimeStepByLSERKW2(0.0)
    _close(s.tgv_analyze(NA), S.tgv_analyze(s.get_state(), s.gradients(), NAnalyze=NA), "step")
    for bad in (0, 32):
        with pytest.raises(RuntimeError, match="NAnalyze"):
            s.tgv_analyze(bad)
    s.close()


def test_csv_format_is_flexis():
    lines = open(CSV).read().splitlines()
    assert lines[0] == ",".join(["Time"] + analysis.TGV_VARNAMES)
    for line in lines[1:6]:
        vals = [float(v) for v in line.split(",")]
        assert ", ".join(analysis.fortran_e(v) for v in vals) == line
    assert analysis.fortran_e(0.0) == "0.00000000000000E+00000"
    assert analysis.fortran_e(-10.671740566753) == "-.10671740566753E+00002"
    assert analysis.fortran_e(0.99999999999999999) == "0.10000000000000E+00001"


def test_csv_file_lines(tmp_path):
    p = str(tmp_path / "case_TGVAnalysis.csv")
    f = analysis.TGVAnalysisFile(p)
    f.append(0.5, np.arange(15) * 0.25)
    f2 = analysis.TGVAnalysisFile(p, keep=True)
    f2.append(1.0, -np.ones(15))
    lines = open(p).read().splitlines()
    assert len(lines) == 3 and lines[1].startswith("0.50000000000000E+00000, 0.00000000000000E+00000")
    assert lines[2].startswith("0.10000000000000E+00001, -.10000000000000E+00001")
