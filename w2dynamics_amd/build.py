"""Build the HIP engine in-tree: ``python -m w2dynamics_amd.build``.

Produces ``w2dynamics_amd/libw2dyn_ctqmc.so`` for gfx950 with hipcc (all HIP code
compiled with packed-FP32 instructions disabled) and, with amdflang, the Fortran
host module MCTQMC: ``w2dynamics_amd/libw2dyn_ctqmc_fortran.so`` with its module
file in ``w2dynamics_amd/fortran/mod/`` and the test driver
``tests/fortran/host_run``; then the Python extension ``w2dynamics_amd/CTQMC*.so``
(module ``CTQMC.mctqmc``, the name and layout ``cthyb_solver.py`` imports) made by
f2py from the Fortran host.
"""

from __future__ import annotations

import os
import subprocess
import sys
import sysconfig

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
SOURCES = [os.path.join(HERE, "csrc", "ctqmc_engine.hip"), os.path.join(HERE, "csrc", "model.cpp")]
TARGET = os.path.join(HERE, "libw2dyn_ctqmc.so")
FORTRAN_HOST = os.path.join(HERE, "fortran", "ctqmc_host.F90")
FORTRAN_LIB = os.path.join(HERE, "libw2dyn_ctqmc_fortran.so")
FORTRAN_MOD = os.path.join(HERE, "fortran", "mod")
HOST_DRIVER_SRC = os.path.join(ROOT, "tests", "fortran", "host_run.F90")
HOST_DRIVER = os.path.join(ROOT, "tests", "fortran", "host_run")
F2PY_DIR = os.path.join(HERE, "fortran", "f2py")
F2PY_MAP = os.path.join(HERE, "fortran", "f2cmap")
F2PY_EXT = os.path.join(HERE, "CTQMC" + (sysconfig.get_config_var("EXT_SUFFIX") or ".so"))


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
    build_f2py(verbose)
    return FORTRAN_LIB


def build_f2py(verbose: bool = False) -> str:
    """The Python extension of the Fortran host, made the way
    src/ctqmc_fortran/CMakeLists.txt makes CTQMC from CTQMC.F90: f2py writes
    the C module and Fortran wrappers (kinds mapped to double by f2cmap), gcc
    compiles the C parts, amdflang the wrappers against the MCTQMC module file,
    and the result links the Fortran host library."""
    import numpy
    os.makedirs(F2PY_DIR, exist_ok=True)
    npdir = os.path.dirname(numpy.__file__)
    inc = ["-I" + os.path.join(npdir, "f2py", "src"), "-I" + numpy.get_include(),
           "-I" + sysconfig.get_paths()["include"]]
    steps = [
        [sys.executable, "-m", "numpy.f2py", FORTRAN_HOST, "-m", "CTQMC", "--lower", "--f2cmap", F2PY_MAP,
         "--build-dir", F2PY_DIR],
        ["gcc", "-O2", "-fPIC", "-c", os.path.join(F2PY_DIR, "CTQMCmodule.c"), *inc, "-o",
         os.path.join(F2PY_DIR, "CTQMCmodule.o")],
        ["gcc", "-O2", "-fPIC", "-c", os.path.join(npdir, "f2py", "src", "fortranobject.c"), *inc, "-o",
         os.path.join(F2PY_DIR, "fortranobject.o")],
        ["amdflang", "-O2", "-fPIC", "-c", os.path.join(F2PY_DIR, "CTQMC-f2pywrappers2.f90"), "-I" + FORTRAN_MOD,
         "-o", os.path.join(F2PY_DIR, "CTQMC-f2pywrappers2.o")],
        ["amdflang", "-shared", "-fPIC", *(os.path.join(F2PY_DIR, o) for o in
                                            ("CTQMCmodule.o", "fortranobject.o", "CTQMC-f2pywrappers2.o")),
         "-L" + HERE, "-lw2dyn_ctqmc_fortran", "-lw2dyn_ctqmc", "-Wl,-rpath,$ORIGIN", "-o", F2PY_EXT + ".tmp"],
    ]
    for cmd in steps:
        if verbose:
            print(" ".join(cmd), flush=True)
        subprocess.run(cmd, check=True, cwd=F2PY_DIR,
                       stdout=None if verbose else subprocess.DEVNULL)
    os.replace(F2PY_EXT + ".tmp", F2PY_EXT)
    return F2PY_EXT


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
