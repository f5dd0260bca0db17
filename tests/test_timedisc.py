"""The 2N-storage Runge-Kutta tables of flexi_amd.timedisc (FLEXI's SetTimeDiscCoefs).

The coefficients are data copied value for value; what pins them is the mathematics: each
scheme, run as TimeStepByLSERKW2 runs it (R = f(t+c dt, Y) - A R; Y += b dt R), must converge
at its design order on a non-autonomous ODE with a known solution (which also exercises RKc), and
be consistent (sum of the effective weights = 1). A wrong digit in any coefficient breaks the
order.
"""
import numpy as np
import pytest

from flexi_amd.timedisc import SCHEMES

ORDER = {"carpenterrk4-5": 4, "standardrk3-3": 3, "niegemannrk4-14": 4, "toulorgerk4-8c": 4,
         "toulorgerk3-7c": 3, "toulorgerk4-8f": 4}


def lserk_solve(scheme, f, y0, t0, t1, n):
    dt = (t1 - t0) / n
    y, t = y0, t0
    A = np.concatenate([[0.0], scheme.RKA])
    c = np.concatenate([[0.0], scheme.RKc])
    for _ in range(n):
        R = 0.0
        for s in range(scheme.nStages):
            R = f(t + c[s] * dt, y) - A[s] * R
            y = y + scheme.RKb[s] * dt * R
        t += dt
    return y


@pytest.mark.parametrize("name", list(SCHEMES))
def test_scheme_converges_at_its_order(name):
    sc = SCHEMES[name]
    assert len(sc.RKA) == len(sc.RKc) == sc.nStages - 1
    assert len(sc.cfl_alpha_gauss) == len(sc.cfl_alpha_lobatto) == 15
    # y' = cos(t) y (1 - y/3) has an exact solution through y = 3 e^{s}/(2 + e^{s}), s = sin t
    f = lambda t, y: np.cos(t) * y * (1.0 - y / 3.0)
    exact = 3.0 * np.exp(np.sin(2.0)) / (2.0 + np.exp(np.sin(2.0)))
    errs = [abs(lserk_solve(sc, f, 1.0, 0.0, 2.0, n) - exact) for n in (20, 40, 80)]
    rates = np.log2(np.array(errs[:-1]) / np.array(errs[1:]))
    print(name, errs, rates)
    p = ORDER[name]
    assert np.all(rates > p - 0.3), rates
    # a consistent scheme integrates y' = 1 exactly
    assert abs(lserk_solve(sc, lambda t, y: 1.0, 0.0, 0.0, 1.0, 7) - 1.0) < 1e-13
