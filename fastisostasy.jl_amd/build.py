"""Build libfastiso_hip.so in-tree for gfx950 (hipcc). Usage: python fastisostasy.jl_amd/build.py"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
SRCS = [os.path.join(HERE, "csrc", "fi_engine.hip"), os.path.join(HERE, "csrc", "fno.hip")]
OUT = os.path.join(HERE, "lib", "libfastiso_hip.so")
FLAGS = ["--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
         "-Wno-unused-result"]


def build(verbose=True):
    os.makedirs(os.path.dirname(OUT), exist_ok=True)
    hipcc = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
    cmd = [hipcc] + FLAGS + ["-I", os.path.join(HERE, "..", "include"), "-o", OUT + ".tmp"] + SRCS
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.check_call(cmd)
    os.replace(OUT + ".tmp", OUT)
    return OUT


if __name__ == "__main__":
    build()
    sys.exit(0)
