"""Builds libfxg.so (gfx950) in-tree with hipcc.

``python -m flexi_amd.build`` or ``flexi_amd.build.build()``. Every source is compiled on each call
(the check that the library builds); objects go to ``flexi_amd/_build`` and the shared library to
``flexi_amd/libfxg.so``.

A second library, ``flexi_amd/libfxg_strict.so``, is the same code compiled with
``-ffp-contract=off``: no multiply and add is fused into one FMA, so every operation rounds as
written, as in FLEXI's own build of the Fortran. The loader picks it with ``FXG_LIB=strict``
(``flexi_amd/_lib.py``); it exists for parity runs whose outcome depends on that rounding.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libfxg.so")
SOURCES = ["fxg_kernels.hip", "fxg_fused.hip", "fxg_visc.hip", "fxg_adjoint.hip", "fxg_metrics.hip", "fxg_analyze.hip", "fxg_api.hip"]
ARCH = os.environ.get("FXG_OFFLOAD_ARCH", "gfx950")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLAGS = [
    "-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
    # packed FP32 instructions (v_pk_*_f32) stay off: some of their forms have returned wrong
    # results on MI355X without raising an error
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function",
]


# library variants: file name -> extra compile flags
VARIANTS = {"libfxg.so": [], "libfxg_strict.so": ["-ffp-contract=off"]}


def _compile(src: str, variant: str = "libfxg.so") -> str:
    odir = BUILD if variant == "libfxg.so" else os.path.join(BUILD, os.path.splitext(variant)[0])
    obj = os.path.join(odir, os.path.splitext(src)[0] + ".o")
    cmd = [HIPCC, *FLAGS, *VARIANTS[variant], "-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{r.stderr[-8000:]}")
    return obj


# Mesh and state files (include/fxg_io.h): host code over the HDF5 C library, a separate library
# so that libfxg.so does not depend on HDF5.
HDF5_ROOT = os.environ.get("FXG_HDF5_ROOT", "/opt/conda")
IO_LIB = os.path.join(HERE, "libfxg_io.so")


def build_io(verbose: bool = False) -> str | None:
    """libfxg_io.so with g++ against HDF5_ROOT; None (and a note) when HDF5 is not installed."""
    inc, lib = os.path.join(HDF5_ROOT, "include"), os.path.join(HDF5_ROOT, "lib")
    if not os.path.exists(os.path.join(inc, "hdf5.h")):
        print(f"fxg_io not built: no hdf5.h under {inc} (set FXG_HDF5_ROOT)", file=sys.stderr)
        return None
    tmp = IO_LIB + ".tmp"
    cmd = ["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-Wall", "-I", inc, os.path.join(CSRC, "fxg_io.cpp"),
           "-o", tmp, "-L", lib, "-lhdf5", f"-Wl,-rpath,{lib}"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"g++ failed for fxg_io.cpp:\n{r.stderr[-8000:]}")
    os.replace(tmp, IO_LIB)
    if verbose:
        print("built", IO_LIB)
    return IO_LIB


def build(verbose: bool = False) -> str:
    """Builds libfxg_io.so and every library variant; returns the path of libfxg.so."""
    build_io(verbose)
    jobs = [(src, v) for v in VARIANTS for src in SOURCES]
    for v in VARIANTS:
        os.makedirs(BUILD if v == "libfxg.so" else os.path.join(BUILD, os.path.splitext(v)[0]), exist_ok=True)
    with cf.ThreadPoolExecutor(max_workers=min(len(jobs), os.cpu_count() or 8)) as ex:
        objs = list(ex.map(lambda j: _compile(*j), jobs))
    for v in VARIANTS:
        lib = os.path.join(HERE, v)
        tmp = lib + ".tmp"
        mine = [o for o, (_, jv) in zip(objs, jobs) if jv == v]
        cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *mine]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"link failed:\n{r.stderr[-8000:]}")
        os.replace(tmp, lib)
        if verbose:
            print("built", lib)
    return LIB


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
