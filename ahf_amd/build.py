"""Build the native parts of ahf_amd in-tree.

  libahf_amd.so      HIP kernels for gfx950 + the C ABI (include/ahf_gpu.h)
  libahf_snapgen.so  seeded synthetic snapshot generator (host C, OpenMP)

Every call recompiles.  hipcc cross-compiles for gfx950 without a GPU.
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "lib")

HIP_SOURCES = ["api.hip", "amr_gpu.hip", "unbind_gpu.hip", "gather_gpu.hip", "profiles_gpu.hip", "gridinfo_gpu.hip"]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
# parity-critical FP64/FP32 arithmetic: no contraction into FMA; packed FP32
# instructions disabled for every device function
HIP_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-fPIC", "-std=c++17", "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-I", os.path.join(ROOT, "include"),
]


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(" ".join(cmd) + "\n" + r.stdout + r.stderr)
        raise RuntimeError(f"build step failed: {cmd[0]} ({r.returncode})")
    return r


def build_gpu(verbose=False):
    os.makedirs(LIB, exist_ok=True)
    objs = []
    for src in HIP_SOURCES:
        path = os.path.join(CSRC, src)
        if not os.path.exists(path):
            continue
        obj = os.path.join(LIB, src.replace(".hip", ".o"))
        _run([HIPCC, *HIP_FLAGS, "-c", path, "-o", obj])
        objs.append(obj)
    for csrc in ("dropin.c", "halo_io.c"):
        path = os.path.join(CSRC, csrc)
        if not os.path.exists(path):
            continue
        obj = os.path.join(LIB, csrc.replace(".c", ".o"))
        _run(["gcc", "-std=c11", "-O2", "-fPIC", "-fopenmp", "-ffp-contract=off", "-Wall", "-Wextra",
              "-I", os.path.join(ROOT, "include"), "-I", CSRC, "-c", path, "-o", obj])
        objs.append(obj)
    out = os.path.join(LIB, "libahf_amd.so")
    tmp = out + ".tmp"
    # the drop-in's host loops use GNU OpenMP, the runtime AHF itself links
    _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp, *objs, "-lpthread", "-lm",
          "-l:libgomp.so.1"])
    os.replace(tmp, out)
    for o in objs:
        os.remove(o)
    return out


def build_snapgen():
    os.makedirs(LIB, exist_ok=True)
    out = os.path.join(LIB, "libahf_snapgen.so")
    _run(["gcc", "-std=c11", "-O2", "-fPIC", "-shared", "-fopenmp", "-Wall", "-Wextra",
          "-o", out, os.path.join(CSRC, "snapgen.c"), "-lm"])
    return out


def build_all():
    return [build_snapgen(), build_gpu()]


if __name__ == "__main__":
    for p in build_all():
        print(p)
