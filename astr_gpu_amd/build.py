"""Build step of the HIP library (hipcc, gfx950), in-tree.

    python -m astr_gpu_amd.build
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
SOURCES = [os.path.join(HERE, "csrc", "astr_solver.hip")]
HEADERS = [os.path.join(HERE, "csrc", "line_engine.hpp"), os.path.join(ROOT, "include", "astr_hip.h")]
TARGET = os.path.join(HERE, "libastr_hip.so")

FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
    # packed FP32 instructions are kept out of every kernel on this target
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
]


def hipcc() -> str:
    for c in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", "hipcc"):
        if c and (os.path.isabs(c) and os.path.exists(c) or not os.path.isabs(c)):
            return c
    return "hipcc"


def build_library(verbose: bool = False) -> str:
    """Compile every HIP source into libastr_hip.so (always recompiles)."""
    tmp = TARGET + ".tmp"
    cmd = [hipcc(), *FLAGS, "-o", tmp, *SOURCES]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError("hipcc failed: " + " ".join(cmd))
    if verbose:
        sys.stderr.write(res.stderr)
    os.replace(tmp, TARGET)
    return TARGET


FORTRAN_DIR = os.path.join(HERE, "fortran")
FORTRAN_EXE = os.path.join(FORTRAN_DIR, "tgv_hip")


def build_fortran_example() -> str | None:
    """Compile the iso_c_binding module and the Fortran Taylor-Green driver
    against libastr_hip.so (needs amdflang; returns None without it)."""
    fc = os.environ.get("FC_ASTR", "/opt/rocm/bin/amdflang")
    if not os.path.exists(fc):
        return None
    import tempfile
    with tempfile.TemporaryDirectory() as tmp:
        mod = os.path.join(FORTRAN_DIR, "astr_hip_mod.f90")
        obj = os.path.join(tmp, "astr_hip_mod.o")
        subprocess.run([fc, "-O2", "-J", tmp, "-c", mod, "-o", obj], check=True)
        subprocess.run([fc, "-O2", "-J", tmp, "-o", FORTRAN_EXE,
                        os.path.join(FORTRAN_DIR, "tgv_hip.f90"), obj, "-L" + HERE, "-lastr_hip",
                        "-Wl,-rpath,$ORIGIN/.."], check=True)
    return FORTRAN_EXE


if __name__ == "__main__":
    print(build_library(verbose="-v" in sys.argv))
    print(build_fortran_example())
