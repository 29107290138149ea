"""Build the HIP engine in-tree: ``python -m w2dynamics_amd.build``.

Produces ``w2dynamics_amd/libw2dyn_ctqmc.so`` for gfx950 with hipcc.  All HIP code
is compiled with packed-FP32 instructions disabled.
"""

from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
SOURCES = [os.path.join(HERE, "csrc", "ctqmc_engine.hip"), os.path.join(HERE, "csrc", "model.cpp")]
TARGET = os.path.join(HERE, "libw2dyn_ctqmc.so")


def build(verbose: bool = False) -> str:
    cmd = ["hipcc", "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
           "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
           "-I" + os.path.join(ROOT, "include"), *SOURCES, "-o", TARGET + ".tmp"]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(TARGET + ".tmp", TARGET)
    return TARGET


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
