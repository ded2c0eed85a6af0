"""Builds tests/stages/libpipeline_stages.so, the stage kernels of the
pipeline tests, with the library's own compiler flags."""

from __future__ import annotations

import os
import sys
from pathlib import Path

HERE = Path(__file__).resolve().parent
ROOT = HERE.parent.parent
LIB = HERE / "libpipeline_stages.so"


def build() -> Path:
    if str(ROOT) not in sys.path:
        sys.path.insert(0, str(ROOT))
    from rendiation_amd import build as rb

    hipcc = rb._hipcc()
    tmp = LIB.with_suffix(".so.tmp")
    rb._run([hipcc, *rb.COMMON, *rb.HIP_FLAGS, "-shared", "-o", str(tmp), str(HERE / "pipeline_stages.hip")])
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build())
