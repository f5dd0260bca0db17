"""The C ABI used from a plain C program (tests/c_host/fxg_c_host.c): a FLEXI-like driver that
reads a HOPR mesh, restarts from a FLEXI State file, runs TimeStepByLSERKW2 steps and writes a
State file, through include/fxg.h and include/fxg_io.h only. On the GPU its result must be
bit-identical to the same steps driven from Python (ctypes) on the same library."""
import os
import shutil
import subprocess

import numpy as np
import pytest

from flexi_amd import testcase

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
MESH = os.path.join(ROOT, "tests", "golden", "hdf5", "CART_HEX_PERIODIC_004_mesh.h5")


def build_host(tmp_path) -> str:
    """gcc against the in-tree libfxg.so and libfxg_io.so (a host program, no GPU code)."""
    cc = shutil.which("gcc") or shutil.which("cc")
    if cc is None:
        pytest.skip("no C compiler")
    lib = os.path.join(ROOT, "flexi_amd")
    exe = str(tmp_path / "fxg_c_host")
    r = subprocess.run([cc, "-O2", "-Wall", "-Werror", "-I", os.path.join(ROOT, "include"),
                        os.path.join(ROOT, "tests", "c_host", "fxg_c_host.c"), "-L", lib, "-lfxg", "-lfxg_io",
                        f"-Wl,-rpath,{lib}", "-o", exe], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    return exe


def test_c_host_builds_against_the_headers_and_libraries(tmp_path):
    exe = build_host(tmp_path)
    r = subprocess.run([exe], capture_output=True, text=True)
    assert r.returncode == 2 and "usage" in r.stderr


@pytest.mark.gpu
def test_c_host_matches_the_python_host_bit_for_bit(tmp_path):
    from flexi_amd import io
    from flexi_amd.solver import DGSolver
    exe = build_host(tmp_path)
    N, steps, dt = 5, 10, 2e-3
    mesh = io.read_mesh(MESH)
    s = DGSolver(mesh, N, node_type="gauss-lobatto", split_dg="pi", riemann="roeentropyfix", mu0=6.25e-4,
                 R=71.42857)
    U0 = testcase.tgv(s.metrics()["Elem_xGP"], Ma=0.5)
    s_in, s_out = str(tmp_path / "in_State.h5"), str(tmp_path / "out_State.h5")
    io.write_state(s_in, U0, N, "gauss-lobatto", 0.0, os.path.basename(MESH), project="tgv")
    print(f"[c host] {exe} N={N} steps={steps}", flush=True)
    r = subprocess.run([exe, MESH, str(N), str(steps), repr(dt), s_in, s_out], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stderr
    print(r.stdout.strip())
    out = io.read_state(s_out)
    s.set_state(U0)
    t = s.run(steps, fixed_dt=dt)
    U = s.get_state()
    s.close()
    assert out["N"] == N and out["node_type"] == "GAUSS-LOBATTO" and abs(out["time"] - t) <= 1e-15
    assert np.all(np.isfinite(U)) and np.max(np.abs(U - U0)) > 1e-6  # the state moved
    assert np.array_equal(out["U"], U)
