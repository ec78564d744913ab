"""Build the gfx950 shared library holomap_amd/libholomap_gpu.so.

Each HIP source is compiled with hipcc to an object in holomap_amd/_build/
(in parallel) and linked into one shared library that exports the C ABI of
include/holomap_gpu.h; the JSON-lines front door (csrc/hm_jsonl.cpp) is
linked against it as holomap_amd/holomap_gpu_jsonl.

Numerics flags: -ffp-contract=off keeps every a*b+c as two roundings (the
reference's f64 distance folds and f32 calibration depend on it; the MFMA
key computation uses explicit fma builtins), correctly rounded f32
division/sqrt, and no fast-math. Packed-FP32 instructions are disabled.
"""
from __future__ import annotations

import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libholomap_gpu.so")
JSONL = os.path.join(HERE, "holomap_gpu_jsonl")  # the JSON-lines front door
SOURCES = ["hm_prep.hip", "hm_gram.hip", "hm_knn.hip", "hm_calib.hip", "hm_sort.hip",
           "hm_fuzzy.hip", "hm_graph.hip", "hm_runtime.hip", "hm_serve.hip"]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("HM_OFFLOAD_ARCH", "gfx950")
CFLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC",
    "-ffp-contract=off", "-fhip-fp32-correctly-rounded-divide-sqrt", "-fno-fast-math",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-pass-failed",
]


def _compile(src: str) -> str:
    obj = os.path.join(BUILD, src.replace(".hip", ".o"))
    cmd = [HIPCC, *CFLAGS, "-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{r.stderr}")
    return obj


def build_jsonl() -> str:
    """The JSON-lines front door (host C++ over the library's C ABI)."""
    tmp = JSONL + ".tmp"
    cmd = [HIPCC, "-O2", "-std=c++17", "-Wall", os.path.join(CSRC, "hm_jsonl.cpp"),
           f"-L{HERE}", "-lholomap_gpu", "-Wl,-rpath,$ORIGIN", "-o", tmp]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hm_jsonl.cpp failed:\n{r.stderr}")
    os.replace(tmp, JSONL)
    return JSONL


def build(verbose: bool = False) -> str:
    os.makedirs(BUILD, exist_ok=True)
    with ThreadPoolExecutor(max_workers=min(8, len(SOURCES))) as ex:
        objs = list(ex.map(_compile, SOURCES))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stderr}")
    os.replace(tmp, LIB)
    build_jsonl()
    if verbose:
        print(f"built {LIB} and {JSONL}")
    return LIB


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
