"""fxg_calc_timestep_dev_noalloc (CALCTIMESTEP into caller-owned device memory, scratch in the
library's fixed device array, no allocation) against fxg_calc_timestep on the same state: dt and
errType bit for bit, on two calls in a row, the second on a smaller mesh (another context on the
same GPU, so the array still holds the first call's larger contents), then the first mesh again;
and a mesh above the scratch maximum is rejected with a message."""
import numpy as np
import pytest

from flexi_amd import _lib, testcase
from flexi_amd.mesh import periodic_box

pytestmark = pytest.mark.gpu


def noalloc_dt(solver, U, torch):
    dev = torch.device("cuda:0")
    Ud = torch.from_numpy(U.ravel(order="F").copy()).to(dev)
    dt_d = torch.full((1,), -1.0, dtype=torch.float64, device=dev)
    err_d = torch.full((1,), -7, dtype=torch.int32, device=dev)
    torch.cuda.synchronize()  # the context's stream does not wait for torch's stream
    _lib.check(_lib.lib().fxg_calc_timestep_dev_noalloc(solver.ctx, Ud.data_ptr(), solver.CFL, solver.DFL,
                                                         dt_d.data_ptr(), err_d.data_ptr(), None))
    _lib.check(_lib.lib().fxg_sync(solver.ctx))
    return float(dt_d.cpu()[0]), int(err_d.cpu()[0])


def test_no_alloc_calc_timestep_equals_calc_timestep():
    import torch
    from flexi_amd.solver import DGSolver
    cases = [((6, 5, 4), 5, "gauss-lobatto", "pi"), ((3, 2, 2), 3, "gauss", None)]
    solvers = []
    for shape, N, node, split in cases:
        s = DGSolver(periodic_box(*shape), N, node_type=node, split_dg=split, riemann="roe")
        U = testcase.tgv(s.metrics()["Elem_xGP"], Ma=0.3)
        solvers.append((s, U))
    for k, (s, U) in enumerate(solvers + solvers[:1]):
        print(f"[no-alloc] call {k}: {s.mesh.nElems} elements N={s.N}", flush=True)
        dt, err = noalloc_dt(s, U, torch)
        s.set_state(U)
        dt_ref, err_ref = s.CalcTimeStep()
        assert err == err_ref == 0
        assert dt == dt_ref and dt > 0.0, (dt, dt_ref)
    # a NaN density is reported through errType without failing the call
    s, U = solvers[1]
    Ubad = U.copy(order="F")
    Ubad[0, 1, 1, 1, 3] = np.nan
    dt, err = noalloc_dt(s, Ubad, torch)
    assert err == 1
    for s, _ in solvers:
        s.close()


def test_no_alloc_rejects_meshes_above_its_maximum():
    """524,800 elements, above the 524,288 of the scratch array: the host check refuses the call
    before any launch (the buffers are valid device memory of the right sizes all the same)."""
    import torch
    from flexi_amd.solver import DGSolver
    s = DGSolver(periodic_box(82, 80, 80), 1, node_type="gauss", split_dg=None, riemann="lf")
    assert s.mesh.nElems == 524800
    dev = torch.device("cuda:0")
    Ud = torch.ones(5 * 8 * s.mesh.nElems, dtype=torch.float64, device=dev)
    dt_d = torch.zeros(1, dtype=torch.float64, device=dev)
    err_d = torch.zeros(1, dtype=torch.int32, device=dev)
    torch.cuda.synchronize()
    rc = _lib.lib().fxg_calc_timestep_dev_noalloc(s.ctx, Ud.data_ptr(), 1.0, 1.0, dt_d.data_ptr(), err_d.data_ptr(),
                                                  None)
    msg = _lib.lib().fxg_last_error().decode()
    s.close()
    assert rc != 0 and "exceed its scratch maximum" in msg
