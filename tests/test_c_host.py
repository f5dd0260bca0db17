"""The C ABI used from plain C programs, through include/fxg.h and include/fxg_io.h only:
tests/c_host/fxg_c_host.c, a FLEXI-like driver that reads a HOPR mesh, restarts from a FLEXI
State file, runs TimeStepByLSERKW2 steps and writes a State file (bit-identical on the GPU to the
same steps driven from Python), and tests/c_host/fxg_c_cavity.c, FLEXI's cavity_3D regression
case with boundary conditions and the time loop (FLEXI's own State file at t=1 to 1e-12)."""
import os
import shutil
import subprocess

import numpy as np
import pytest

from flexi_amd import testcase

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
MESH = os.path.join(ROOT, "tests", "golden", "hdf5", "CART_HEX_PERIODIC_004_mesh.h5")


def build_host(tmp_path, name="fxg_c_host") -> str:
    """gcc against the in-tree libfxg.so and libfxg_io.so (a host program, no GPU code)."""
    cc = shutil.which("gcc") or shutil.which("cc")
    if cc is None:
        pytest.skip("no C compiler")
    lib = os.path.join(ROOT, "flexi_amd")
    exe = str(tmp_path / name)
    r = subprocess.run([cc, "-O2", "-Wall", "-Werror", "-I", os.path.join(ROOT, "include"),
                        os.path.join(ROOT, "tests", "c_host", name + ".c"), "-L", lib, "-lfxg", "-lfxg_io",
                        f"-Wl,-rpath,{lib}", "-lm", "-o", exe], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    return exe


@pytest.mark.parametrize("name", ["fxg_c_host", "fxg_c_cavity"])
def test_c_host_builds_against_the_headers_and_libraries(tmp_path, name):
    exe = build_host(tmp_path, name)
    r = subprocess.run([exe], capture_output=True, text=True)
    assert r.returncode == 2 and "usage" in r.stderr


@pytest.mark.gpu
def test_c_cavity_host_reproduces_flexis_regression_state(tmp_path):
    """FLEXI's parabolic/cavity_3D case from a C program through the C ABI alone (boundary
    sides, fxg_set_boundaries, CALCTIMESTEP and the dt rules to t=1): FLEXI's State file within
    its h5diff tolerance (1e-12 absolute), and bit for bit the state the Python driver reaches."""
    from flexi_amd import io
    from flexi_amd import run as frun
    cav = os.path.join(ROOT, "tests", "golden", "cavity")
    exe = build_host(tmp_path, "fxg_c_cavity")
    out = str(tmp_path / "cavity_Re100_State_0000001.000000000.h5")
    print(f"[c host] {exe}", flush=True)
    r = subprocess.run([exe, os.path.join(cav, "cavity4x4x4_mesh.h5"), out], capture_output=True, text=True,
                       timeout=200)
    assert r.returncode == 0, r.stderr
    print(r.stdout.strip())
    assert "steps=353" in r.stdout
    st = io.read_state(out)
    ref = io.read_state(os.path.join(cav, "reggie_cavity_Re100_State_0000001.000000000.h5"))
    assert st["time"] == 1.0 and st["node_type"] == "GAUSS"
    assert np.max(np.abs(st["U"] - ref["U"])) <= 1e-12
    d = tmp_path / "py"
    d.mkdir()
    case = frun.Case(frun.parse_parameter_file(os.path.join(cav, "parameter.ini")), cav, node_type="gauss")
    res = frun.run(case, str(d), log=lambda s: None)
    assert np.array_equal(io.read_state(res["states"][-1])["U"], st["U"])


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


def build_fortran_host(tmp_path) -> str:
    """amdflang: flexi_amd/fortran/fxg_interface.f90 (MOD_FXG) and tests/fortran/fxg_f_host.f90
    against the in-tree libfxg.so and libfxg_io.so (host code only)."""
    fc = shutil.which("amdflang") or shutil.which("flang")
    if fc is None:
        pytest.skip("no Fortran compiler")
    lib = os.path.join(ROOT, "flexi_amd")
    mod = os.path.join(ROOT, "flexi_amd", "fortran", "fxg_interface.f90")
    prog = os.path.join(ROOT, "tests", "fortran", "fxg_f_host.f90")
    exe = str(tmp_path / "fxg_f_host")
    r = subprocess.run([fc, "-c", mod, "-o", str(tmp_path / "mod_fxg.o"), f"-J{tmp_path}"], capture_output=True,
                       text=True, cwd=tmp_path)
    assert r.returncode == 0, r.stderr
    r = subprocess.run([fc, prog, str(tmp_path / "mod_fxg.o"), f"-I{tmp_path}", "-L", lib, "-lfxg", "-lfxg_io",
                        f"-Wl,-rpath,{lib}", "-o", exe], capture_output=True, text=True, cwd=tmp_path)
    assert r.returncode == 0, r.stderr
    return exe


def test_fortran_host_builds_against_the_module_and_libraries(tmp_path):
    exe = build_fortran_host(tmp_path)
    r = subprocess.run([exe], capture_output=True, text=True)
    assert r.returncode == 2 and "usage" in r.stdout


@pytest.mark.gpu
def test_fortran_host_matches_the_c_host_bit_for_bit(tmp_path):
    """FLEXI's time loop from Fortran through MOD_FXG (ISO_C_BINDING): 10 x (CALCTIMESTEP,
    TimeStepByLSERKW2, t += dt) on the GPU, the same loop from the C host: every dt and the final
    State file bit for bit equal (timedisc.f90:160-212, timestep.f90:82-152)."""
    from flexi_amd import io
    from flexi_amd.solver import DGSolver
    f_exe = build_fortran_host(tmp_path)
    c_exe = build_host(tmp_path)
    N, steps, cfl, dfl = 5, 10, 0.2, 0.2
    mesh = io.read_mesh(MESH)
    s = DGSolver(mesh, N, node_type="gauss-lobatto", split_dg="pi", riemann="roeentropyfix", mu0=6.25e-4,
                 R=71.42857)
    U0 = testcase.tgv(s.metrics()["Elem_xGP"], Ma=0.5)
    s.close()
    s_in = str(tmp_path / "in_State.h5")
    io.write_state(s_in, U0, N, "gauss-lobatto", 0.0, os.path.basename(MESH), project="tgv")
    out = {}
    for name, cmd in (("f", [f_exe, MESH, str(N), str(steps), repr(cfl), repr(dfl), s_in, str(tmp_path / "f.h5")]),
                      ("c", [c_exe, MESH, str(N), str(steps), "0", s_in, str(tmp_path / "c.h5"), repr(cfl),
                             repr(dfl)])):
        print(f"[{name} host] N={N} steps={steps} CFLScale={cfl}", flush=True)
        r = subprocess.run(cmd, capture_output=True, text=True, timeout=120)
        assert r.returncode == 0, r.stdout + r.stderr
        dts = [float(ln.split("=")[1]) for ln in r.stdout.splitlines() if ln.startswith("dt=")]
        out[name] = (dts, io.read_state(str(tmp_path / f"{name}.h5")))
    (df, sf), (dc, sc) = out["f"], out["c"]
    print("dt:", df[0], "...", df[-1])
    assert len(df) == len(dc) == steps and df == dc and df[0] > 0.0
    assert sf["time"] == sc["time"] and abs(sf["time"] - sum(dc)) <= 1e-15
    assert np.max(np.abs(sf["U"] - U0)) > 1e-6  # the state moved
    assert np.array_equal(sf["U"], sc["U"])
