"""Builds librdrt.so (gfx950 kernels + host scene builder) in-tree.

Host-only sources are compiled with g++, HIP sources with hipcc for gfx950,
and everything is linked with hipcc into ``rendiation_amd/librdrt.so``.
Every float operation of the path must be a single IEEE rounding, so all
code is built with ``-ffp-contract=off``; packed FP32 instructions are
disabled for device code.
"""

from __future__ import annotations

import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

PKG = Path(__file__).resolve().parent
CSRC = PKG / "csrc"
BUILD = PKG / "_build"
LIB = PKG / "librdrt.so"

ARCH = os.environ.get("RDRT_OFFLOAD_ARCH", "gfx950")
HOST_SOURCES = ["bvh_build.cpp", "scene.cpp"]
HIP_SOURCES = ["trace_fast_256.hip", "trace_fast_128.hip", "trace_fast_64.hip", "trace_quad_256.hip",
               "trace_quad_128.hip", "trace_quad_64.hip", "trace_qquad_256.hip", "trace_qquad_128.hip",
               "trace_qquad_64.hip", "trace_kernels.hip", "pt_kernels.hip", "batch_kernels.hip", "api.hip",
               "api_pt.hip", "api_batch.hip", "refit_kernels.hip", "device_build.hip",
               "attrib_kernels.hip", "task_graph.hip", "ao_kernels.hip", "api_ao.hip", "det_sum.hip", "api_static.hip",
               "pipeline_kernels.hip", "api_pipeline.hip", "pt_tail_0.hip", "pt_tail_1.hip", "pt_tail_2.hip",
               "api_external.hip"]

COMMON = ["-O3", "-std=c++17", "-fPIC", "-ffp-contract=off", "-fno-fast-math", "-Wall"]
HIP_FLAGS = [
    f"--offload-arch={ARCH}",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-fhip-fp32-correctly-rounded-divide-sqrt",
    "-Wno-unused-command-line-argument",
]


def _hipcc() -> str:
    for cand in (shutil.which("hipcc"), "/opt/rocm/bin/hipcc"):
        if cand and Path(cand).exists():
            return cand
    raise RuntimeError("hipcc not found; the ROCm toolchain is required to build librdrt")


def _run(cmd: list[str]) -> None:
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"build step failed ({res.returncode}): {' '.join(cmd)}\n{res.stdout}\n{res.stderr}")


def build(verbose: bool = False) -> Path:
    """Compile every source and link librdrt.so; returns the library path."""
    BUILD.mkdir(exist_ok=True)
    hipcc = _hipcc()
    jobs = []
    objs = []
    for src in HOST_SOURCES:
        obj = BUILD / (Path(src).stem + ".o")
        jobs.append(["g++", *COMMON, "-fopenmp-simd", "-c", str(CSRC / src), "-o", str(obj)])
        objs.append(obj)
    for src in HIP_SOURCES:
        obj = BUILD / (Path(src).stem + ".hip.o")
        jobs.append([hipcc, *COMMON, *HIP_FLAGS, "-c", str(CSRC / src), "-o", str(obj)])
        objs.append(obj)
    with ThreadPoolExecutor(max_workers=min(8, len(jobs))) as ex:
        for cmd, _ in zip(jobs, ex.map(_run, jobs)):
            if verbose:
                print(" ".join(cmd), file=sys.stderr)
    tmp = LIB.with_suffix(".so.tmp")
    _run([hipcc, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", str(tmp), *map(str, objs), "-lpthread"])
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
