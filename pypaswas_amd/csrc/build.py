"""Builds libpypaswas_hip.so for gfx950 with hipcc, in-tree.

Run as ``python -m pypaswas_amd.csrc.build`` or through ``__graft_entry__.build()``.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.dirname(HERE)
OUT = os.path.join(PKG, "libpypaswas_hip.so")
SOURCES = [os.path.join(HERE, "sw_capi.hip")]
DEPENDS = SOURCES + [os.path.join(HERE, "sw_device.hip"), os.path.join(HERE, "sw_server.hip"), os.path.join(HERE, "sw_kernels.h"),
                     os.path.join(os.path.dirname(PKG), "include", "pypaswas_hip.h")]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-shared",
    # no packed FP32 (v_pk_*_f32) instructions anywhere
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    # strict IEEE single precision: no contraction, no fast math
    "-ffp-contract=off",
    "-fno-fast-math",
    # every value is finite (pw_set_scoring rejects non-finite or huge
    # constants), so a maximum needs no signalling-NaN quieting of its inputs
    "-fno-honor-nans",
    "-Wall",
]


def build(verbose: bool = False) -> str:
    cmd = [HIPCC] + FLAGS + SOURCES + ["-o", OUT + ".tmp"]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(OUT + ".tmp", OUT)
    return OUT


if __name__ == "__main__":
    build(verbose="-v" in sys.argv)
    print(OUT)
