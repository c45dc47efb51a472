"""Compiles the HIP engine for gfx950 into qutree_amd/libqutree_amd.so."""
from __future__ import annotations

import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libqutree_amd.so")
SOURCES = ["runtime.cpp", "gemm_ops.hip", "node_ops.hip", "tree_ops.hip", "cmf_ops.hip", "model.cpp", "engine.cpp", "engine_gemm.cpp", "launch_set.cpp", "ttno.cpp", "cmf.cpp",
           "capi_engine.cpp", "formats.cpp", "seam.cpp", "smallmat.hip", "grad_ops.hip", "gradient.cpp", "tuning.cpp",
           "scf.cpp", "scf_ops.hip", "tree_prog.cpp", "apply_fit.cpp"]
# -amdgpu-mfma-vgpr-form: MFMA accumulators in ordinary VGPRs (gfx950 has
# one register file), so loop-carried accumulators are not copied to and
# from AGPRs around every MFMA; the grouped mode product then fits 7
# wavefronts per SIMD instead of 5.
# -amdgpu-kernarg-preload-count=16: the first 16 argument dwords arrive in
# SGPRs with the dispatch, instead of by scalar loads at the start of every
# wavefront (the sweep of latency-bound launches measured 288 -> 283 us).
FLAGS = ["--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-Wall", "-Wno-unused-function",
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", "-mllvm", "-amdgpu-mfma-vgpr-form",
         "-mllvm", "-amdgpu-kernarg-preload-count=16"]


def _compile(src: str) -> str:
    obj = os.path.join(BUILD, os.path.splitext(src)[0] + ".o")
    cmd = ["hipcc", "-x", "hip", *FLAGS, "-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError(f"hipcc failed on {src}")
    return obj


def build(verbose: bool = False) -> str:
    os.makedirs(BUILD, exist_ok=True)
    with ThreadPoolExecutor(max_workers=min(8, len(SOURCES))) as ex:
        objs = list(ex.map(_compile, SOURCES))
    tmp = LIB + ".tmp"
    cmd = ["hipcc", "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp, *objs]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("link failed")
    os.replace(tmp, LIB)
    if verbose:
        print("built", LIB)
    return LIB


if __name__ == "__main__":
    build(verbose=True)
