"""Build the HIP engine in-tree: ``python -m w2dynamics_amd.build``.

Produces ``w2dynamics_amd/libw2dyn_ctqmc.so`` for gfx950 with hipcc (all HIP code
compiled with packed-FP32 instructions disabled) and, with amdflang, the Fortran
host module MCTQMC: ``w2dynamics_amd/libw2dyn_ctqmc_fortran.so`` with its module
file in ``w2dynamics_amd/fortran/mod/`` and the test driver
``tests/fortran/host_run``.
"""

from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
SOURCES = [os.path.join(HERE, "csrc", "ctqmc_engine.hip"), os.path.join(HERE, "csrc", "model.cpp")]
TARGET = os.path.join(HERE, "libw2dyn_ctqmc.so")
FORTRAN_HOST = os.path.join(HERE, "fortran", "ctqmc_host.F90")
FORTRAN_LIB = os.path.join(HERE, "libw2dyn_ctqmc_fortran.so")
FORTRAN_MOD = os.path.join(HERE, "fortran", "mod")
HOST_DRIVER_SRC = os.path.join(ROOT, "tests", "fortran", "host_run.F90")
HOST_DRIVER = os.path.join(ROOT, "tests", "fortran", "host_run")


def build(verbose: bool = False) -> str:
    cmd = ["hipcc", "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
           "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
           "-I" + os.path.join(ROOT, "include"), *SOURCES, "-o", TARGET + ".tmp"]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(TARGET + ".tmp", TARGET)
    build_fortran(verbose)
    return TARGET


def build_fortran(verbose: bool = False) -> str:
    """MCTQMC host module over the C-ABI (iso_c_binding), and its test driver."""
    os.makedirs(FORTRAN_MOD, exist_ok=True)
    lib = ["amdflang", "-cpp", "-fPIC", "-O2", "-shared", "-module-dir", FORTRAN_MOD, FORTRAN_HOST,
           "-L" + HERE, "-lw2dyn_ctqmc", "-Wl,-rpath,$ORIGIN", "-o", FORTRAN_LIB + ".tmp"]
    drv = ["amdflang", "-cpp", "-O2", "-I" + FORTRAN_MOD, HOST_DRIVER_SRC, "-L" + HERE,
           "-lw2dyn_ctqmc_fortran", "-lw2dyn_ctqmc", "-Wl,-rpath,$ORIGIN/../../w2dynamics_amd",
           "-o", HOST_DRIVER + ".tmp"]
    for cmd in (lib, drv):
        if verbose:
            print(" ".join(cmd), flush=True)
        subprocess.run(cmd, check=True, cwd=FORTRAN_MOD)
        if cmd is lib:
            os.replace(FORTRAN_LIB + ".tmp", FORTRAN_LIB)
    os.replace(HOST_DRIVER + ".tmp", HOST_DRIVER)
    return FORTRAN_LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
