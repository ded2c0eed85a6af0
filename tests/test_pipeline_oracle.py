"""CPU checks of the programmable pipeline's expectations and host side.

* The committed fixture of the reference's pipeline test (test.rs:33-141 on
  naive/test.rs:9-225) is what the oracle and the FP32 restatement of the
  test stages give today.
* The oracle's hits behind that image are pinned by geometry: rays passing
  within 4.8 of the sphere's centre (radius 5, 32 x 16 grid, inscribed
  radius above 4.8) hit it on the near side, t within 0.25 of the analytic
  sphere's; rays farther than 5 miss; culling back faces changes nothing
  (every hit is on an outward face).
* The image restatement: misses are (0, 0, 150, 255); a hit is (round(255
  u), round(255 v), 150, 255), so red + green <= 256.
* The chain's payload restatement on hand-made hit records.
* The binding table's hit-group index (sbt.rs:30) and the C-ABI's argument
  checks (on a context without a device).
"""

import ctypes as C
from pathlib import Path

import numpy as np
import pytest

from rendiation_amd import _lib as L
from rendiation_amd.acceleration import AccelerationStructureSystem, BuildOptions
from rendiation_amd.rt_pipeline import ShaderBindingTable
from tests import pipeline_expect as PE
from tests.scenes import oracle_only, reference_default_lists

GOLDEN = Path(__file__).resolve().parent / "golden" / "pipeline_debug_256.npz"


@pytest.fixture(scope="module")
def debug_run():
    orc = oracle_only(*reference_default_lists(4), BuildOptions())
    rays = PE.camera_rays(256, 256, 0.1, 100.0)
    hits, _, _ = orc.trace(rays, flags=PE.CULL_BACK)
    return orc, rays, hits


def test_fixture_matches_oracle(debug_run):
    _, _, hits = debug_run
    g = np.load(GOLDEN)
    assert np.array_equal(hits[:, 0], g["prim"])
    assert np.array_equal(PE.debug_image(hits), g["rgba"].reshape(-1, 4))


def test_debug_hits_pinned_by_sphere_geometry(debug_run):
    orc, rays, hits = debug_run
    d = rays[:, 4:7].astype(np.float64)
    c = np.array([0.0, 0.0, -10.0])
    cd = d @ c
    dist = np.linalg.norm(c[None] - cd[:, None] * d, axis=1)
    hit = hits[:, 0] != PE.MISS
    inner, outer = dist < 4.8, dist > 5.0
    assert inner.sum() > 10000 and outer.sum() > 10000
    assert hit[inner].all() and not hit[outer].any()
    t = hits[:, 1].view(np.float32).astype(np.float64)
    near = cd - np.sqrt(np.maximum(25.0 - dist ** 2, 0.0))
    assert np.abs(t[inner] - near[inner]).max() < 0.25
    nocull, _, _ = orc.trace(rays)
    assert np.array_equal(nocull, hits)


def test_debug_image_restatement(debug_run):
    _, _, hits = debug_run
    img = PE.debug_image(hits)
    hit = hits[:, 0] != PE.MISS
    assert (img[~hit] == np.array([0, 0, 150, 255], np.uint8)).all()
    assert (img[:, 2] == 150).all() and (img[:, 3] == 255).all()
    uv = hits[hit, 2:4].view(np.float32).astype(np.float64)
    assert (np.abs(img[hit, :2] - uv * 255.0) <= 0.5 + 1e-9).all()
    assert (img[hit, 0].astype(int) + img[hit, 1] <= 256).all()


def test_chain_payload_restatement():
    M = PE.MISS
    h = np.full((4, 3, 4), 0xDEAD, dtype=np.uint32)
    h[0, 0, 0] = M                                   # camera ray misses
    h[1, 0, 0], h[1, 1, 0] = 5, M                    # hit, then miss
    h[2, :, 0] = [1, 2, 3]                           # three hits
    h[3, 0, 0], h[3, 1, 0], h[3, 2, 0] = 0xFFFFFFF0, 7, M  # wraps mod 2^32
    p = PE.chain_payload(h)
    assert p.tolist() == [[0x10000, 0, 0], [6 + 0x10000, 1, 1], [2 + 3 + 4, 3, 2],
                          [(0xFFFFFFF1 + 8 + 0x10000) & 0xFFFFFFFF, 2, 2]]
    assert PE.chain_depths(h).tolist() == [1, 2, 3, 3]


def test_sbt_hit_group_index():
    sbt = ShaderBindingTable(ray_type_count=2)
    sbt.config_hit_group(1, 0, 1, 7)   # 1 + 1 * 2 + 0
    sbt.config_hit_group(0, 1, 0, 8)   # 0 + 0 + 1
    assert sbt.hit_groups == {3: 7, 1: 8}
    sbt.config_hit_group(1, 0, 1, None)
    assert sbt.hit_groups == {1: 8}


def test_pipeline_create_argument_checks():
    system = AccelerationStructureSystem(device=None)
    lib, ctx = system._L, system._h
    out = C.c_void_p()

    def create(**kw):
        d = L.RdrtPipelineDesc()
        d.ray_gen, d.payload_words = 0x1000, 1
        for k, v in kw.items():
            setattr(d, k, v)
        return lib.rdrt_pipeline_create(ctx, C.byref(d), C.byref(out))

    assert create(ray_gen=0) == -1
    assert create(payload_words=0) == -1
    assert create(payload_words=17) == -1
    assert create(nclosest=5) == -1
    assert create(nclosest=4, nmiss=3) == -1
    assert create(max_in_flight=65) == -1
    assert create(hit_context=4) == -1
    assert create(tile_pixels=(1 << 24) + 1) == -1
    assert create(nclosest=4, nmiss=2, payload_words=16, hit_context=3) == -6  # valid, but no device
    assert out.value is None
    assert lib.rdrt_pipeline_destroy(None) == 0


def test_pipeline_record_layouts():
    """ctypes records against the C compiler's layout of include/rdrt.h."""
    import subprocess
    import tempfile

    root = Path(__file__).resolve().parent.parent
    src = r"""
#include <stddef.h>
#include <stdio.h>
#include "rdrt.h"
int main(void) {
  printf("%zu %zu %zu %zu %zu %zu %zu\n", sizeof(rdrt_trace_call), sizeof(rdrt_pipeline_desc),
         offsetof(rdrt_pipeline_desc, nclosest), offsetof(rdrt_pipeline_desc, hit_context),
         sizeof(rdrt_stage_args), offsetof(rdrt_stage_args, user), offsetof(rdrt_stage_args, capacity));
  return 0;
}
"""
    with tempfile.TemporaryDirectory() as d:
        c = Path(d) / "layout.c"
        c.write_text(src)
        exe = Path(d) / "layout"
        subprocess.run(["gcc", "-std=c99", "-I", str(root / "include"), str(c), "-o", str(exe)], check=True)
        got = [int(x) for x in subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split()]
    assert got[0] == C.sizeof(L.RdrtTraceCall) == 64
    assert got[1] == C.sizeof(L.RdrtPipelineDesc)
    assert got[2] == L.RdrtPipelineDesc.nclosest.offset
    assert got[3] == L.RdrtPipelineDesc.hit_context.offset
    # rdrt_stage_args: 17 pointers, then launch_size[2], payload_words, capacity
    assert got[4:] == [17 * 8 + 16, 16 * 8, 17 * 8 + 12]


def test_external_buffer_argument_checks():
    system = AccelerationStructureSystem(device=None)
    lib, ctx = system._L, system._h
    out, ptr = C.c_void_p(), C.c_void_p()
    assert lib.rdrt_external_buffer_import(ctx, -1, 4096, 0, 4096, C.byref(out), C.byref(ptr)) == -1
    assert lib.rdrt_external_buffer_import(ctx, 3, 0, 0, 4096, C.byref(out), C.byref(ptr)) == -1
    assert lib.rdrt_external_buffer_import(ctx, 3, 4096, 4000, 200, C.byref(out), C.byref(ptr)) == -1
    assert lib.rdrt_external_buffer_import(ctx, 3, 4096, 0, 4096, C.byref(out), C.byref(ptr)) == -6  # no device
    assert out.value is None and ptr.value is None
    assert lib.rdrt_external_buffer_release(None) == 0
