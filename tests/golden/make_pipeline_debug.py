"""Writes pipeline_debug_256.npz: the image (256 x 256 RGBA8) and the
closest primitive per pixel of the pipeline of shader/ray-tracing/src/
test.rs:33-141 on the scene of naive/test.rs:9-225 (TLAS 4 bound), computed
by the oracle's closest-hit traversal and the FP32 restatement of the test
stages (tests/pipeline_expect.py).  Run from the repository root:
    python tests/golden/make_pipeline_debug.py
"""

from __future__ import annotations

import sys
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))

from rendiation_amd.acceleration import BuildOptions  # noqa: E402
from tests import pipeline_expect as PE  # noqa: E402
from tests.scenes import oracle_only, reference_default_lists  # noqa: E402


def make(w: int = 256, h: int = 256):
    orc = oracle_only(*reference_default_lists(4), BuildOptions())
    hits, _, _ = orc.trace(PE.camera_rays(w, h, 0.1, 100.0), flags=PE.CULL_BACK)
    return PE.debug_image(hits).reshape(h, w, 4), hits[:, 0].copy()


if __name__ == "__main__":
    rgba, prim = make()
    out = Path(__file__).resolve().parent / "pipeline_debug_256.npz"
    np.savez_compressed(out, rgba=rgba, prim=prim)
    print(out, out.stat().st_size, "bytes")
