"""FLEXI's CSV output of the Taylor-Green vortex testcase analysis.

The analysis itself (AnalyzeTestcase, src/testcase/taylorgreenvortex/testcase.f90:241-479) runs on
the device (``DGSolver.tgv_analyze``, the library's ``fxg_tgv_analyze``). ``TGVAnalysisFile``
writes its values as FLEXI's ``<ProjectName>_TGVAnalysis.csv`` (InitOutputToFile / OutputToFile,
src/output/output.f90:647-869): a header line, then one line per analysis in the Fortran format
E23.14E5 separated by ", ".
"""
from __future__ import annotations

import os

TGV_VARNAMES = ["Dissipation Rate Incompressible", "Dissipation Rate Compressible", "Ekin incomp", "Ekin comp",
                "Enstrophy comp", "DR_u", "DR_S", "DR_Sd", "DR_p", "Maximum Vorticity", "Mean Temperature", "uprime",
                "Mean Entropy", "ED_S", "ED_D"]


def fortran_e(x: float, d: int = 14, e: int = 5) -> str:
    """Fortran's E(d+9).dEe edit of x (E23.14E5 by default): 0.ddd...E+eeeee, the leading zero
    dropped for negative numbers to fit the width."""
    x = float(x)
    if x == 0.0:
        mant, ex = "0" * d, 0
    else:
        m = f"{abs(x):.{d - 1}e}"  # d significant digits: D.DDDDe+XX
        digits, exp10 = m.split("e")
        mant = digits.replace(".", "")
        ex = int(exp10) + 1
    es = f"{'-' if ex < 0 else '+'}{abs(ex):0{e}d}"
    return (f"-.{mant}E{es}" if x < 0 else f"0.{mant}E{es}")


class TGVAnalysisFile:
    """``<ProjectName>_TGVAnalysis.csv`` with FLEXI's header and number format; ``append`` adds one
    line (time, 15 values). A new file is written unless ``keep`` asks to append to an existing
    one (a restart)."""

    def __init__(self, path: str, keep: bool = False):
        self.path = path
        if not (keep and os.path.exists(path)):
            with open(path, "w") as f:
                f.write(",".join(["Time"] + TGV_VARNAMES) + "\n")

    def append(self, t: float, values) -> None:
        with open(self.path, "a") as f:
            f.write(", ".join(fortran_e(v) for v in [t, *values]) + "\n")
