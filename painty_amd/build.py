"""Build recipe for libpainty_hip.so (gfx950) -- plain hipcc, in-tree.

    python -m painty_amd.build          # compile + link
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(ROOT, "_build")
LIB = os.path.join(HERE, "libpainty_hip.so")
SOURCES = ["kernels.hip", "persistent.hip", "fused.hip", "batch.hip", "batchwave.hip", "grad.hip", "serve.hip", "display.hip", "ingest.hip", "texture.hip", "image.hip", "api.hip"]
ARCH = os.environ.get("PAINTY_HIP_ARCH", "gfx950")

HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
FLAGS = [
    "--offload-arch=" + ARCH, "-O3", "-fPIC", "-std=c++17",
    "-Wall", "-Wno-unused-function",
    # packed FP32 (v_pk_*_f32) stays off: hipcc may emit it from plain float
    # code and some of its forms have given wrong results on these parts.
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I" + os.path.join(ROOT, "include"), "-I" + CSRC,
]


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(" ".join(cmd) + "\n" + r.stdout + r.stderr)
        raise RuntimeError("build step failed: %s" % cmd[-1])
    return r


def build(verbose: bool = False, out: str = LIB, objdir: str = BUILD, extra=()) -> str:
    """Compiles every source and links `out`.  `extra`: more hipcc flags for
    each compile (a sanitizer build of the host code passes them, each
    behind -Xarch_host), with its objects in `objdir`."""
    os.makedirs(objdir, exist_ok=True)
    objs = []

    def compile_one(src):
        obj = os.path.join(objdir, os.path.splitext(src)[0] + ".o")
        cmd = [HIPCC] + FLAGS + list(extra) + ["-c", os.path.join(CSRC, src), "-o", obj]
        if verbose:
            print(" ".join(cmd), flush=True)
        _run(cmd)
        return obj

    with ThreadPoolExecutor(max_workers=len(SOURCES)) as ex:
        objs = list(ex.map(compile_one, SOURCES))
    tmp = out + ".tmp"
    _run([HIPCC, "--offload-arch=" + ARCH, "-shared", "-fPIC", "-o", tmp] + objs + ["-lz"])
    os.replace(tmp, out)
    return out


if __name__ == "__main__":
    print(build(verbose=True))
