"""Build libfastiso_hip.so in-tree for gfx950 (hipcc). Usage: python fastisostasy.jl_amd/build.py

Each source compiles to an object in parallel (the engine's kernel instances dominate the time),
then one link step makes the shared library.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
SRCS = [os.path.join(HERE, "csrc", "fi_engine.hip"), os.path.join(HERE, "csrc", "fi_batch256.hip"),
        os.path.join(HERE, "csrc", "fno.hip")]
OUT = os.path.join(HERE, "lib", "libfastiso_hip.so")
FLAGS = ["--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC",
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
         "-Wno-unused-result"]


def build(verbose=True):
    os.makedirs(os.path.dirname(OUT), exist_ok=True)
    hipcc = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
    inc = ["-I", os.path.join(HERE, "..", "include")]
    objs, procs = [], []
    for src in SRCS:
        obj = os.path.join(os.path.dirname(OUT), os.path.splitext(os.path.basename(src))[0] + ".o")
        cmd = [hipcc] + FLAGS + inc + ["-c", "-o", obj, src]
        if verbose:
            print(" ".join(cmd), flush=True)
        objs.append(obj)
        procs.append(subprocess.Popen(cmd))
    codes = [p.wait() for p in procs]
    if any(codes):
        raise subprocess.CalledProcessError(max(codes), "hipcc")
    cmd = [hipcc, "--offload-arch=gfx950", "-fPIC", "-shared", "-o", OUT + ".tmp"] + objs
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.check_call(cmd)
    os.replace(OUT + ".tmp", OUT)
    return OUT


if __name__ == "__main__":
    build()
    sys.exit(0)
