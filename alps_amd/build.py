"""Build of the gfx950 engine library (alps_amd/libalpsg.so) and of the Fortran
binding module.  Run as ``python -m alps_amd.build``; __graft_entry__.build()
calls :func:`build_all`.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libalpsg.so")
FDIR = os.path.join(HERE, "fortran")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
SOURCES = ["alpsg_api.cpp", "alpsg_server.cpp", "alpsg_tables.hip", "alpsg_disp.hip", "alpsg_nhds.hip",
           "alpsg_rel.hip", "alpsg_fitsetup.hip", "alpsg_noalloc.hip"]
# No packed-FP32 instructions; FP64 code is unaffected by this switch.
HIP_FLAGS = ["--offload-arch=gfx950", "-O3", "-fPIC", "-std=c++17",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-munsafe-fp-atomics", "-Wno-unused-result"]


def _run(cmd, cwd=None):
    r = subprocess.run(cmd, cwd=cwd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("command failed: " + " ".join(cmd))
    return r


def build_library(verbose: bool = False) -> str:
    """Compile every HIP/C++ source for gfx950 and link libalpsg.so in-tree."""
    objdir = os.path.join(HERE, "_build")
    os.makedirs(objdir, exist_ok=True)
    objs = []
    for src in SOURCES:
        obj = os.path.join(objdir, src + ".o")
        cmd = [HIPCC, *HIP_FLAGS, "-I", os.path.join(ROOT, "include"), "-c",
               os.path.join(CSRC, src), "-o", obj]
        if src.endswith(".hip"):
            cmd.insert(1, "-xhip")
        if verbose:
            print(" ".join(cmd))
        _run(cmd)
        objs.append(obj)
    tmp = LIB + ".tmp"
    _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-pthread", "-Wl,-z,defs", "-o", tmp,
          *objs])
    os.replace(tmp, LIB)
    return LIB


def build_fortran(verbose: bool = False) -> str | None:
    """Compile the ISO_C_BINDING module and its demo driver with amdflang, when present."""
    fc = shutil.which("amdflang") or ("/opt/rocm/bin/amdflang"
                                     if os.path.exists("/opt/rocm/bin/amdflang") else None)
    src = os.path.join(FDIR, "alpsg_mod.f90")
    if fc is None or not os.path.exists(src):
        return None
    objdir = os.path.join(FDIR, "_build")
    os.makedirs(objdir, exist_ok=True)
    _run([fc, "-O2", "-c", src, "-module-dir", objdir, "-o", os.path.join(objdir, "alpsg_mod.o")])
    drv = os.path.join(FDIR, "alpsg_demo.f90")
    exe = os.path.join(FDIR, "alpsg_demo")
    if os.path.exists(drv):
        _run([fc, "-O2", "-I", objdir, drv, os.path.join(objdir, "alpsg_mod.o"),
              "-L", HERE, "-lalpsg", f"-Wl,-rpath,{HERE}", "-o", exe])
        return exe
    return None


def build_all(verbose: bool = False) -> None:
    build_library(verbose)
    build_fortran(verbose)


if __name__ == "__main__":
    build_all(verbose="-v" in sys.argv)
    print(LIB)
