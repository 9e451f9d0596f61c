"""Builds the in-tree native libraries for gfx950.

``ems_amd/libems_gpu.so``: -ffp-contract=off (no FMA contraction: results
are bit-identical to the reference CPU engine), IEEE denormals (hipcc's
default, no -fgpu-flush-denormals-to-zero), packed FP32 instructions
disabled.

``ems_amd/libems_gpu_fma.so``: the same sources with the device code of the
step engine compiled with -ffp-contract=fast (multiply-adds fused into
FMAs); selected with EMS_AMD_FP_CONTRACT=fast. Its results differ from the
reference engine's by rounding only.
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
OUT_FMA = os.path.join(HERE, "libems_gpu_fma.so")

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
    each, the engine twice: exact and with FMA contraction), then links
    both shared libraries."""
    objs = [os.path.join(HERE, "csrc", os.path.basename(src) + ".o") for src in SOURCES]
    fma_obj = os.path.join(HERE, "csrc", "engine_fma.o")
    compile_flags = [f for f in HIPCC_FLAGS if f != "-shared"]
    fma_flags = ["-ffp-contract=fast" if f == "-ffp-contract=off" else f for f in compile_flags]
    jobs = [[hipcc(), *compile_flags, "-c", "-o", obj, src] for src, obj in zip(SOURCES, objs)]
    jobs.append([hipcc(), *fma_flags, "-c", "-o", fma_obj, SOURCES[0]])
    procs = [(cmd, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT,
                                    text=True)) for cmd in jobs]
    failed = []
    for cmd, p in procs:
        out, _ = p.communicate()
        if p.returncode != 0:
            sys.stderr.write(out)
            failed.append(" ".join(cmd))
    if failed:
        raise RuntimeError("hipcc failed: " + "; ".join(failed))
    for out_path, link_objs in ((OUT, objs), (OUT_FMA, [fma_obj] + objs[1:])):
        tmp = out_path + ".tmp"
        cmd = [hipcc(), *HIPCC_FLAGS, "-o", tmp, *link_objs]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            sys.stderr.write(r.stdout + r.stderr)
            raise RuntimeError("hipcc link failed: " + " ".join(cmd))
        os.replace(tmp, out_path)
    for obj in objs + [fma_obj]:
        os.remove(obj)
    if verbose:
        print("built", OUT, "and", OUT_FMA)
    return OUT


if __name__ == "__main__":
    build(verbose=True)
