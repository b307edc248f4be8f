"""Build liblauematch.so (HIP, gfx950) and the drop-in host programs.

Run ``python -m lauematching_amd.build``.  Outputs land in-tree:
``lauematching_amd/liblauematch.so`` and ``bin/LaueMatchingGPU``,
``bin/LaueMatchingGPUStream`` (the names the reference's Python wrapper
resolves, ``laue_index/indexer.py:33-45``).
"""
from __future__ import annotations

import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
INCLUDE = os.path.join(ROOT, "include")
OBJ = os.path.join(PKG, "_build")
LIB = os.path.join(PKG, "liblauematch.so")
BIN = os.path.join(ROOT, "bin")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"
# Packed-FP32 instructions are disabled for all device code; float64 code is
# built without FP contraction so every operation rounds like the reference.
DEVICE_FLAGS = ["--offload-arch=" + ARCH, "-Xclang", "-target-feature", "-Xclang",
                "-packed-fp32-ops"]
COMMON = ["-O3", "-fPIC", "-std=c++17", "-ffp-contract=off", "-Wall", "-Wno-unused-result",
          "-I", INCLUDE, "-I", CSRC]

LIB_SOURCES = ["lm_api.cpp", "lm_fwd.hip", "lm_score.hip", "lm_band.hip", "lm_merge.cpp",
               "lm_merge_dev.hip", "lm_prep.hip", "lm_median.hip",
               "lm_serve.cpp", "lm_tune.cpp", "lm_hosttable.cpp"]
HOST_COMMON = ["host/lm_params.cpp", "host/lm_hostio.cpp", "host/lm_postproc.cpp"]
HOST_PROGRAMS = {"LaueMatchingGPU": ["host/lm_host_gpu.cpp", *HOST_COMMON],
                 "LaueMatchingGPUStream": ["host/lm_host_stream.cpp", *HOST_COMMON]}


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("build failed: " + " ".join(cmd) + "\n" + r.stdout + r.stderr)
    return r


def _compile(src):
    path = os.path.join(CSRC, src)
    out = os.path.join(OBJ, os.path.basename(src) + ".o")
    lang = ["-x", "hip"] if src.endswith(".hip") else []
    _run([HIPCC, "-c", *COMMON, *DEVICE_FLAGS, *lang, path, "-o", out])
    return out


def build(verbose: bool = False) -> None:
    os.makedirs(OBJ, exist_ok=True)
    os.makedirs(BIN, exist_ok=True)
    srcs = [s for s in LIB_SOURCES if os.path.exists(os.path.join(CSRC, s))]
    with ThreadPoolExecutor(4) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = LIB + ".tmp"
    _run([HIPCC, "-shared", *DEVICE_FLAGS, "-o", tmp, *objs, "-lpthread"])
    os.replace(tmp, LIB)
    if verbose:
        print("built", LIB)
    for name, srcs in HOST_PROGRAMS.items():
        paths = [os.path.join(CSRC, s) for s in srcs]
        if not all(os.path.exists(p) for p in paths):
            continue
        exe = os.path.join(BIN, name)
        _run(["g++", "-O3", "-std=c++17", "-fopenmp", "-ffp-contract=off", "-Wall",
              "-I", INCLUDE, "-I", os.path.join(CSRC, "host"), *paths, "-o", exe + ".tmp",
              "-L", PKG, "-llauematch", "-Wl,-rpath,$ORIGIN/../lauematching_amd",
              "-lpthread", "-lm"])
        os.replace(exe + ".tmp", exe)
        if verbose:
            print("built", exe)


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
