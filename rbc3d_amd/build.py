"""Builds librbc3d_gpu.so (gfx950) in-tree with hipcc.

Every HIP source is compiled with packed-FP32 instruction selection disabled and
linked into one shared library exporting the C ABI of include/rbc3d.h.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "librbc3d_gpu.so")
OBJ = os.path.join(HERE, "_obj")
SOURCES = ["ctx.hip", "sh.hip", "pair.hip", "pme.hip", "wall.hip", "gmres.hip", "mlp.hip", "mech.hip", "timeint.hip", "repulsion.hip", "io.hip"]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLAGS = ["--offload-arch=gfx950", "-O3", "-fPIC", "-std=c++17",
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
         "-Wno-unused-result"]


def _compile(src):
    obj = os.path.join(OBJ, src.replace(".hip", ".o"))
    cmd = [HIPCC, *FLAGS, "-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{r.stderr[-4000:]}")
    return obj


FLANG = os.environ.get("FC", "/opt/rocm/bin/amdflang")
FDIR = os.path.join(HERE, "fortran")
FBUILD = os.path.join(FDIR, "build")
DRIVER_SRC = os.path.join(os.path.dirname(HERE), "tests", "fortran", "velocity_driver.f90")
DRIVER = os.path.join(os.path.dirname(HERE), "tests", "fortran", "velocity_driver")


def build_fortran(verbose=True):
    """The ISO_C_BINDING module (rbc3d_bind) and the Fortran host driver the GPU
    tests run; the driver links librbc3d_gpu.so through an $ORIGIN-relative rpath."""
    os.makedirs(FBUILD, exist_ok=True)
    mod_obj = os.path.join(FBUILD, "rbc3d_bind.o")
    cmds = [
        [FLANG, "-O2", "-fPIC", "-c", os.path.join(FDIR, "rbc3d_bind.f90"), "-J", FBUILD,
         "-o", mod_obj],
        [FLANG, "-O2", "-I", FBUILD, DRIVER_SRC, mod_obj, "-L", HERE, "-lrbc3d_gpu",
         "-Wl,-rpath,$ORIGIN/../../rbc3d_amd", "-o", DRIVER],
    ]
    for cmd in cmds:
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"amdflang failed:\n{' '.join(cmd)}\n{r.stderr[-4000:]}")
    if verbose:
        print(f"built {DRIVER}")
    return DRIVER


def build(verbose=True):
    os.makedirs(OBJ, exist_ok=True)
    with cf.ThreadPoolExecutor(max_workers=min(8, len(SOURCES))) as ex:
        objs = list(ex.map(_compile, SOURCES))
    tmp = OUT + ".tmp"
    cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp, *objs]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stderr[-4000:]}")
    os.replace(tmp, OUT)
    if verbose:
        print(f"built {OUT}")
    build_fortran(verbose)
    return OUT


if __name__ == "__main__":
    build()
    sys.exit(0)
