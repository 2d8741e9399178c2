"""Build recipe of the native library ``libpvglu.so`` (in-tree, gfx950).

``python -m phdthesis_albertojimenezruiz_amd.build`` compiles every C++ and
HIP source under ``csrc/`` with hipcc for gfx950 and links one shared library
next to this file.  Objects go to ``_build/`` (git-ignored).
"""
from __future__ import annotations

import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
OUT = os.path.join(PKG, "libpvglu.so")
BUILD = os.path.join(ROOT, "_build")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

COMMON = ["-O3", "-fPIC", "-std=c++17", "-Wall", "-Wno-unused-function",
          "-ffp-contract=off", f"-I{os.path.join(ROOT, 'include')}", f"-I{CSRC}"]
# Device code: gfx950 only; packed FP32 ops disabled (not generated by hand
# either); contraction off so each + - * / rounds once, as in the CPU GLU.
DEVICE = ["-x", "hip", "--offload-arch=gfx950", "-Xclang", "-target-feature",
          "-Xclang", "-packed-fp32-ops"]


def sources():
    out = []
    for f in sorted(os.listdir(CSRC)):
        if f.endswith(".cpp") or f.endswith(".hip"):
            out.append(os.path.join(CSRC, f))
    return out


def _compile(src: str, verbose: bool) -> str:
    obj = os.path.join(BUILD, os.path.basename(src) + ".o")
    if src.endswith(".hip"):
        cmd = [HIPCC, *COMMON, *DEVICE, "-c", src, "-o", obj]
    else:
        cmd = [HIPCC, *COMMON, "-x", "c++", "-c", src, "-o", obj]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    return obj


def build(verbose: bool = False, jobs: int | None = None) -> str:
    """Compiles every source (in parallel, one process per file) and links."""
    os.makedirs(BUILD, exist_ok=True)
    srcs = sources()
    jobs = jobs or min(len(srcs), max(1, min(8, os.cpu_count() or 1)))
    with ThreadPoolExecutor(jobs) as ex:
        objs = list(ex.map(lambda s: _compile(s, verbose), srcs))
    tmp = OUT + ".tmp"
    cmd = [HIPCC, "-shared", "-fPIC", "--offload-arch=gfx950", "-o", tmp, *objs,
           "-lpthread"]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, OUT)
    return OUT


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
