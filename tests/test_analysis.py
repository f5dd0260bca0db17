"""The TGV testcase analysis of the product (flexi_amd.analysis) and its CSV output.

The oracle's AnalyzeTestcase reproduces the reference's tgv/split TGVAnalysis CSV (its own test);
here the product's numpy analysis is checked against the oracle on the same state and gradients,
and the CSV number format against the reference CSV's own text.
"""
import os

import numpy as np
import pytest

import oracle as orc
from flexi_amd import analysis, testcase
from flexi_amd.mesh import periodic_box

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSV = os.path.join(ROOT, "tests", "golden", "tgv", "TGV_Re1600_Split_TGVAnalysis_Reference.csv")


@pytest.mark.parametrize("N,node,split", [(3, orc.NODE_GL, 1), (4, orc.NODE_G, 0)])
def test_tgv_analysis_matches_oracle(N, node, split):
    mesh = periodic_box(4, 4, 4, lo=(0, 0, 0))
    prm = orc.params(N, node, split, orc.RIEMANN_ROEENTROPYFIX)
    m = orc.calc_metrics(N, node, mesh)
    S = orc.Solver(prm, mesh, m)
    U = testcase.tgv(m["Elem_xGP"], Ma=0.5)
    U = np.asfortranarray(U * (1.0 + 1e-3 * np.sin(3.0 * m["Elem_xGP"][0]))[None])  # break the symmetry
    S.rhs(U)
    g = S.gradients()
    ref = S.tgv_analyze(U, g, NAnalyze=6)
    got = analysis.tgv_analyze(U, *g, m["sJ"], N, node, 1.4, 71.42857, 6.25e-4, NAnalyze=6, chunk=17)
    for c, (a, b) in enumerate(zip(got, ref)):
        # columns that vanish analytically (divergence terms) are compared absolutely
        tol = 1e-13 if abs(b) < 1e-6 else 1e-12 * abs(b)
        assert abs(a - b) <= tol, (analysis.TGV_VARNAMES[c], a, b)


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
