"""GPU checks of whole runs: FP32-bulk accuracy against the FP64 run (Taylor-Green vortex, 8^3,
N=3 Gauss, weak form, LF, 1000 steps) and freestream preservation on the curvilinear 24^3 N=5
mesh with metrics computed on the GPU."""
import os

import numpy as np
import pytest

from flexi_amd import testcase
from flexi_amd.mesh import periodic_box, sine_deform

pytestmark = pytest.mark.gpu


def weights(N, node):
    import oracle as orc
    return orc.nodes_weights(N, node)[1]


def ekin_rate(U, Ut, W):
    """dEkin/dt = sum w J (u . (rho u)_t - |u|^2/2 rho_t) with Ekin = sum w J rho |u|^2 / 2."""
    rho = U[0]
    u = U[1:4] / rho
    return np.sum(W * (np.sum(u * Ut[1:4], axis=0) - 0.5 * np.sum(u * u, axis=0) * Ut[0]))


def test_fp32_bulk_tgv_1000_steps_against_fp64():
    from flexi_amd.solver import DGSolver
    import oracle as orc
    N = 3
    mesh = periodic_box(8, 8, 8)
    runs = {}
    dt = None
    for prec in ("fp64", "fp32-bulk"):
        s = DGSolver(mesh, N, node_type="gauss", split_dg=None, riemann="lf", precision=prec, CFLScale=0.9,
                     DFLScale=0.9)
        m = s.metrics()
        U0 = testcase.tgv_density_perturbed(m["Elem_xGP"], amplitude=1e-3, seed=1)
        s.set_state(U0)
        if dt is None:
            dt, err = s.CalcTimeStep()
            assert err == 0
        s.run(1000, fixed_dt=dt)  # the FP64 run's first dt for both, so both reach the same time
        U = s.get_state()
        s.DGTimeDerivative_weakForm(1000 * dt)
        runs[prec] = (U, s.get_ut(), m["sJ"])
        s.close()
    (U64, Ut64, sJ), (U32, Ut32, _) = runs["fp64"], runs["fp32-bulk"]
    w = weights(N, orc.NODE_G)
    W = (w[:, None, None] * w[None, :, None] * w[None, None, :])[..., None] / sJ
    err_U = np.linalg.norm((U32 - U64).ravel()) / np.linalg.norm(U64.ravel())
    d64, d32 = -ekin_rate(U64, Ut64, W), -ekin_rate(U32, Ut32, W)
    err_eps = abs(d32 - d64) / abs(d64)
    print(f"FP32-bulk after 1000 steps (t={1000 * dt:.4f}): rel L2 U {err_U:.3e}, "
          f"-dEkin/dt {d64:.6e} vs {d32:.6e} (rel {err_eps:.3e})")
    assert np.all(np.isfinite(U32))
    assert err_U <= 1e-6
    assert err_eps <= 1e-3


@pytest.mark.parametrize("N", [5, 7])
def test_fp32_bulk_fused_split_against_fp64(N):
    """The fused Gauss-Lobatto split-form stage in FP32-bulk (FP32 LDS, FP32 Riemann/split/
    lifting arithmetic; FP64 state, register, metrics) against FP64: TGV 8^3, split PI, Roe,
    density perturbed by 1e-3, 300 steps of the FP64 run's first dt."""
    from flexi_amd.solver import DGSolver
    mesh = periodic_box(8, 8, 8)
    runs, dt = {}, None
    for prec in ("fp64", "fp32-bulk"):
        s = DGSolver(mesh, N, node_type="gauss-lobatto", split_dg="pi", riemann="roe", precision=prec,
                     CFLScale=0.8, DFLScale=0.8)
        assert s.kernel_pipeline() == "fused"
        s.set_state(testcase.tgv_density_perturbed(s.metrics()["Elem_xGP"], amplitude=1e-3, seed=1))
        if dt is None:
            dt, err = s.CalcTimeStep()
            assert err == 0
        s.run(300, fixed_dt=dt)
        runs[prec] = s.get_state()
        s.close()
        assert np.all(np.isfinite(runs[prec])), prec
    U64, U32 = runs["fp64"], runs["fp32-bulk"]
    err_U = np.linalg.norm((U32 - U64).ravel()) / np.linalg.norm(U64.ravel())
    print(f"fused FP32-bulk N={N} after 300 steps: rel L2 U {err_U:.3e}")
    assert np.all(np.isfinite(U32))
    assert err_U <= 1e-6


@pytest.mark.parametrize("pipeline,node,split", [("fused", "gauss-lobatto", "pi"), ("classic", "gauss-lobatto", "pi"),
                                                 ("classic", "gauss-lobatto", None), ("classic", "gauss", None)])
def test_curvilinear_freestream_24cubed_n5_gpu_metrics(pipeline, node, split, monkeypatch):
    from flexi_amd.solver import DGSolver
    if pipeline == "classic":
        monkeypatch.setenv("FXG_PIPELINE", "classic")
    N = 5
    mesh = periodic_box(24, 24, 24, NGeo=N, deform=sine_deform(0.1))
    s = DGSolver(mesh, N, node_type=node, split_dg=split, riemann="roe")
    assert s.kernel_pipeline() == pipeline
    n = N + 1
    U0 = testcase.freestream((n, n, n, mesh.nElems), vel=(0.3, 0.2, 0.1), p=1.0 / 1.4)
    s.set_state(U0)
    s.run(100)
    U = s.get_state()
    err = np.max(np.abs(U - U0))
    print(f"freestream 24^3 N=5 curvilinear ({pipeline}, {node}, {split}), 100 steps to t={s.t:.4f}: "
          f"max|U-U0| = {err:.3e} (per variable {np.abs(U - U0).reshape(5, -1).max(1)})")
    # With library-computed metrics each element takes its face metric from its own metric terms
    # (the fused stage from its boundary nodes, the four-kernel stage from its own face
    # interpolation of the curl-form terms, the slave side included), so an element's volume and
    # surface terms cancel to flux round-off for a uniform state.
    assert err <= 1e-13


@pytest.mark.parametrize("mesh_source", ["box", "hopr_file"])
def test_gpu_reproduces_flexi_tgv_split_regression_trajectory(mesh_source):
    """regressioncheck/checks/tgv/split of the reference: N=7 GL, split PI, RoeEntropyFix, BR1,
    8^3 on [0,2pi]^3, CFL=DFL=0.8, TGV analysis every 10 steps. The GPU state and lifted
    gradients are analysed with the oracle's AnalyzeTestcase restatement and compared with the
    reference CSV over the first 50 steps (the reference's own tolerance is 1e-4 relative). The
    gradients are those of the last stage's right-hand side, which the library recomputes from
    U - RKb_5*dt*R for the fused path. The mesh is the generated box or the case's own HOPR file
    (CART_HEX_PERIODIC_008_mesh.h5, read by flexi_amd.io with FLEXI's side numbering)."""
    import flexi_fixtures as fx
    import oracle as orc
    from flexi_amd.solver import DGSolver
    header, data = fx.tgv_split_csv()
    if mesh_source == "box":
        mesh = periodic_box(8, 8, 8, lo=(0, 0, 0))
    else:
        from flexi_amd import io as fio
        mesh = fio.read_mesh(os.path.join(os.path.dirname(__file__), "golden", "hdf5", "CART_HEX_PERIODIC_008_mesh.h5"))
    m = orc.calc_metrics(7, orc.NODE_GL, mesh)
    S = orc.Solver(orc.params(7, orc.NODE_GL, 1, orc.RIEMANN_ROEENTROPYFIX, mu0=6.25e-4, R=71.42857), mesh, m)
    g = DGSolver(mesh, 7, node_type="gauss-lobatto", split_dg="pi", riemann="roeentropyfix", metrics=m,
                 mu0=6.25e-4, R=71.42857, CFLScale=0.8, DFLScale=0.8)
    assert g.kernel_pipeline() == "fused"
    g.set_state(testcase.tgv(m["Elem_xGP"]))

    def check(row, t):
        # gradients of the last right-hand side, as FLEXI's analysis every nAnalyzeTestCase
        # iterations reads them (timedisc_func.f90:340)
        out = S.tgv_analyze(g.get_state(), g.gradients())
        worst = 0.0
        for c, (got, ref) in enumerate(zip(out, row[1:])):
            if abs(ref) < 1e-6:
                assert abs(got - ref) < 1e-12, header[c + 1]
            else:
                worst = max(worst, abs(got - ref) / abs(ref))
        print(f"t={t:.6f}: max relative deviation from the reference CSV {worst:.2e}")
        assert worst <= 1e-8

    t = 0.0
    g.DGTimeDerivative_weakForm(t)
    check(data[0], t)
    for it in range(1, 51):
        dt, err = g.CalcTimeStep()
        assert err == 0
        g.dt = dt
        g.TimeStepByLSERKW2(t)
        t += dt
        if it % 10 == 0:
            row = data[it // 10]
            assert abs(t - row[0]) <= 1e-12 * row[0]
            check(row, t)


def test_plain_method_50_steps_against_oracle():
    """TGV on [-pi,pi]^3, 8^3 elements, N=3 Gauss, weak form, LF, BR1, CarpenterRK4-5, CFL=DFL=0.9,
    50 steps with CALCTIMESTEP: GPU (four-kernel path) against the oracle step by step."""
    import oracle as orc
    from flexi_amd.solver import DGSolver
    mesh = periodic_box(8, 8, 8)
    m = orc.calc_metrics(3, orc.NODE_G, mesh)
    S = orc.Solver(orc.params(3, orc.NODE_G, 0, orc.RIEMANN_LF), mesh, m)
    g = DGSolver(mesh, 3, node_type="gauss", split_dg=None, riemann="lf", metrics=m, CFLScale=0.9, DFLScale=0.9)
    U0 = testcase.tgv(m["Elem_xGP"])
    Uo = np.asfortranarray(U0.copy())
    R = orc.F(U0.shape)
    g.set_state(U0)
    t = 0.0
    for it in range(50):
        dt, err = g.CalcTimeStep()
        dto, erro, _ = S.calc_timestep(Uo, g.CFL, g.DFL)
        assert err == 0 and erro == 0 and abs(dt - dto) <= 1e-13 * dto
        g.dt = dt
        g.TimeStepByLSERKW2(t)
        S.step(Uo, t, dt, R)
        t += dt
    Ug = g.get_state()
    err = np.linalg.norm((Ug - Uo).ravel()) / np.linalg.norm((Uo - U0).ravel())
    print(f"50 steps to t={t:.4f}: ||U_gpu - U_oracle|| / ||U - U0|| = {err:.2e}")
    assert err <= 1e-10
