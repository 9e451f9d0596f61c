"""Builds the in-tree native library ``ems_amd/libems_gpu.so`` for gfx950.

Flags: -ffp-contract=off (no FMA contraction: results are bit-identical to
the reference CPU engine), IEEE denormals (hipcc's default, no
-fgpu-flush-denormals-to-zero), and packed FP32 instructions disabled.
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
SOURCES = [os.path.join(HERE, "csrc", "engine.hip"), os.path.join(HERE, "csrc", "convnet.hip"),
           os.path.join(HERE, "csrc", "host.cpp")]
OUT = os.path.join(HERE, "libems_gpu.so")

HIPCC_FLAGS = ["--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
               "-ffp-contract=off", "-fno-fast-math",
               "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
               "-Wno-unused-result", "-Wno-unused-value", "-pthread"]


def hipcc() -> str:
    for c in ("/opt/rocm/bin/hipcc", "hipcc"):
        if os.path.sep not in c or os.path.exists(c):
            return c
    return "hipcc"


def build(verbose: bool = False) -> str:
    """Compiles every source to an object in parallel (one hipcc process
    each), then links the shared library."""
    objs = [os.path.join(HERE, "csrc", os.path.basename(src) + ".o") for src in SOURCES]
    compile_flags = [f for f in HIPCC_FLAGS if f != "-shared"]
    procs = []
    for src, obj in zip(SOURCES, objs):
        cmd = [hipcc(), *compile_flags, "-c", "-o", obj, src]
        procs.append((cmd, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT,
                                            text=True)))
    failed = []
    for cmd, p in procs:
        out, _ = p.communicate()
        if p.returncode != 0:
            sys.stderr.write(out)
            failed.append(" ".join(cmd))
    if failed:
        raise RuntimeError("hipcc failed: " + "; ".join(failed))
    tmp = OUT + ".tmp"
    cmd = [hipcc(), *HIPCC_FLAGS, "-o", tmp, *objs]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("hipcc link failed: " + " ".join(cmd))
    os.replace(tmp, OUT)
    for obj in objs:
        os.remove(obj)
    if verbose:
        print("built", OUT)
    return OUT


if __name__ == "__main__":
    build(verbose=True)
