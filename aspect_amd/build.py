"""Build recipe for the gfx950 shared library ``aspect_amd/libaspect_mfs.so``.

Compiles the HIP sources with hipcc for gfx950 only.  Packed FP32 VALU
instructions are disabled (``-packed-fp32-ops``) for every HIP translation unit.
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libaspect_mfs.so")
SOURCES = ["kernels.hip", "kernels_gather.hip", "kernels_planes.hip", "handler.hip", "tria.hip", "probe.hip"]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")


def objects():
    return [os.path.join(CSRC, s) for s in SOURCES]


def build(verbose: bool = False, jobs: int = 3) -> str:
    flags = [
        "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC",
        "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
        "-fno-signed-zeros", "-Wno-unused-result", "-I", os.path.join(HERE, "..", "include"),
    ]
    objs = []
    procs = []
    for src in objects():
        obj = src + ".o"
        objs.append(obj)
        cmd = [HIPCC, *flags, "-c", src, "-o", obj]
        if verbose:
            print(" ".join(cmd), flush=True)
        procs.append((cmd, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)))
    failed = False
    for cmd, p in procs:
        out, _ = p.communicate()
        if p.returncode != 0:
            failed = True
            sys.stderr.write(out)
        elif verbose and out.strip():
            sys.stderr.write(out)
    if failed:
        raise RuntimeError("hipcc compilation failed")
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp, *objs]
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB)
    for o in objs:
        os.remove(o)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
