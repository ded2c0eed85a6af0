"""Programmable pipelines through the C-ABI (rdrt_pipeline_*), with the
caller's stage kernels of tests/stages/pipeline_stages.hip.

* The pipeline of shader/ray-tracing/src/test.rs (ray gen -> trace ->
  closest hit / miss -> ray-gen continuation writing an RGBA8 image) on the
  scene of the reference's tests (naive/test.rs:9-225, TLAS 4 bound): the
  closest hits match the oracle's under the closest-hit contract, the image
  is the stages' function of those hits, and it matches the committed
  oracle fixture except on near-tie pixels.
* Nested traces (closest hit -> trace -> closest hit, three deep, payloads
  handed down and back up): per depth, the rays match the FP32 restatement
  of the stages (from the oracle's hit attributes of the device hits), the
  hits match the oracle's, and the payloads returned to the ray gen are the
  function of the hits the stages saw.
* Tiling, max_in_flight and the round hint do not change the result; every
  task finishes; the ray count is what the stages traced.
"""

import ctypes as C
from pathlib import Path

import numpy as np
import pytest

from rendiation_amd.acceleration import BuildOptions
from rendiation_amd.rt_pipeline import (
    HIT_ATTRIBUTES,
    HIT_TRANSFORMS,
    GPURaytracingPipelineAndBindingSource,
    RayTracingPipeline,
    ShaderBindingTable,
)
from tests import pipeline_expect as PE
from tests.parity import compare
from tests.scenes import reference_default_scene

pytestmark = pytest.mark.gpu

STAGES = Path(__file__).resolve().parent / "stages" / "libpipeline_stages.so"
GOLDEN = Path(__file__).resolve().parent / "golden" / "pipeline_debug_256.npz"


def _say(msg):
    print(f"[gpu-case] {msg}", flush=True)


@pytest.fixture(scope="module")
def stages(device):
    if not STAGES.exists():
        raise FileNotFoundError(f"{STAGES} is missing: run tests/stages/build_stages.py")
    lib = C.CDLL(str(STAGES))
    lib.test_stage.restype = C.c_void_p
    lib.test_stage.argtypes = [C.c_uint32]
    return [lib.test_stage(k) for k in range(13)]


def _pipeline(system, stages, base, payload_words=1, hit_context=0, **kw):
    src = GPURaytracingPipelineAndBindingSource(payload_words=payload_words, hit_context=hit_context, **kw)
    rg = src.register_ray_gen(stages[base])
    ch = src.register_ray_closest_hit(stages[base + 1])
    ms = src.register_ray_miss(stages[base + 2])
    sbt = ShaderBindingTable(ray_type_count=1)
    sbt.config_ray_generation(rg)
    sbt.config_hit_group(0, 0, 0, ch)
    sbt.config_missing(0, ms)
    return RayTracingPipeline(system, src, sbt)


def _run_debug(system, stages, w, h, **kw):
    import torch

    n = w * h
    buf = torch.full((n * 16 + n * 4,), 0xAB, dtype=torch.uint8, device="cuda:0")
    torch.cuda.synchronize()
    p = _pipeline(system, stages, 0, **kw)
    p.trace_ray(w, h, buf.data_ptr())
    system.sync()
    out = buf.cpu().numpy()
    states = p.read_back_execution_states()
    rays = p.ray_count()
    p.close()
    hits = out[: n * 16].view(np.uint32).reshape(n, 4)
    img = out[n * 16:].reshape(n, 4)
    return hits, img, states, rays


@pytest.fixture(scope="module")
def ref_scene(device):
    return reference_default_scene(BuildOptions())


def test_debug_pipeline_reference_scene(stages, ref_scene):
    system, orc = ref_scene
    w = h = 256
    _say(f"debug pipeline {w}x{h}")
    hits, img, states, nrays = _run_debug(system, stages, w, h)
    assert states.is_empty() and not any(states.errors), states
    assert nrays == w * h
    rays = PE.camera_rays(w, h, 0.1, 100.0)
    ref, _, _ = orc.trace(rays, flags=PE.CULL_BACK)
    d = compare(hits, ref)
    assert d.ok() and d.value_mismatch_bits == 0, d
    assert (hits[:, 0] != PE.MISS).sum() > 0.2 * w * h  # the sphere covers much of the view
    # the image is the stages' function of the hits they saw
    assert np.array_equal(img, PE.debug_image(hits))
    # and the oracle fixture's wherever the hit is the oracle's
    g = np.load(GOLDEN)
    same = hits[:, 0] == g["prim"]
    assert same.sum() >= w * h - d.near_ties
    assert np.array_equal(img[same], g["rgba"].reshape(-1, 4)[same])


@pytest.mark.parametrize("tile_pixels,max_in_flight,round_hint", [(4096, 1, 3), (256 * 7 + 5, 3, 12), (1 << 20, 2, 0)])
def test_debug_pipeline_tiling_invariant(stages, ref_scene, tile_pixels, max_in_flight, round_hint):
    system, _ = ref_scene
    w, h = 256, 200
    _say(f"debug pipeline tiles {tile_pixels} in-flight {max_in_flight} rounds {round_hint}")
    base_hits, base_img, _, _ = _run_debug(system, stages, w, h)
    hits, img, states, nrays = _run_debug(system, stages, w, h, tile_pixels=tile_pixels, max_in_flight=max_in_flight,
                                          execution_round_hint=round_hint)
    assert states.is_empty() and not any(states.errors), states
    assert nrays == w * h
    assert np.array_equal(hits, base_hits)
    assert np.array_equal(img, base_img)


def test_graph_replay_equals_eager_calls(stages, ref_scene):
    """A captured call replayed three times (the target cleared between
    calls) gives the eager call's hits and image; a new launch size
    recaptures."""
    system, _ = ref_scene
    _say("pipeline graph replay vs eager")
    eager = _run_debug(system, stages, 200, 150, use_graph=False)
    import torch

    n = 200 * 150
    buf = torch.full((n * 20,), 0xAB, dtype=torch.uint8, device="cuda:0")
    torch.cuda.synchronize()
    p = _pipeline(system, stages, 0, use_graph=True)
    for _ in range(3):
        buf.fill_(0xAB)
        torch.cuda.synchronize()
        p.trace_ray(200, 150, buf.data_ptr())
        system.sync()
        out = buf.cpu().numpy()
        assert np.array_equal(out[: n * 16].view(np.uint32).reshape(n, 4), eager[0])
        assert np.array_equal(out[n * 16:].reshape(n, 4), eager[1])
        st = p.read_back_execution_states()
        assert st.is_empty() and not any(st.errors)
    small = _run_debug(system, stages, 64, 48, use_graph=False)
    m = 64 * 48
    p.trace_ray(64, 48, buf.data_ptr())
    system.sync()
    out = buf.cpu().numpy()
    assert np.array_equal(out[: m * 16].view(np.uint32).reshape(m, 4), small[0])
    p.close()


def test_debug_pipeline_too_few_rounds_leaves_tasks(stages, ref_scene):
    # a trace needs three rounds (ray gen, trace + closest hit, trace
    # resolves + ray gen): with two, the ray-gen tasks are still asleep
    system, _ = ref_scene
    _say("debug pipeline, 2 rounds")
    hits, img, states, nrays = _run_debug(system, stages, 64, 64, execution_round_hint=2)
    assert not states.is_empty()
    assert nrays == 64 * 64
    assert not any(states.errors)


@pytest.mark.parametrize("bind", [3, 0])
def test_chain_pipeline(stages, device, bind):
    import torch

    system, orc = reference_default_scene(BuildOptions(), bind=bind)
    w, h, D = 192, 160, PE.DEPTH
    n = w * h
    _say(f"chain pipeline tlas {bind} {w}x{h}")
    nbytes = n * D * (32 + 16 + 16) + n * 3 * 4
    buf = torch.full((nbytes,), 0xFF, dtype=torch.uint8, device="cuda:0")
    torch.cuda.synchronize()
    p = _pipeline(system, stages, 3, payload_words=3, hit_context=HIT_ATTRIBUTES)
    p.trace_ray(w, h, buf.data_ptr())
    system.sync()
    out = buf.cpu().numpy()
    states = p.read_back_execution_states()
    nrays = p.ray_count()
    p.close()
    off = 0
    rays = out[off: off + n * D * 32].view(np.float32).reshape(n, D, 8)
    off += n * D * 32
    hits = out[off: off + n * D * 16].view(np.uint32).reshape(n, D, 4)
    off += n * D * 16
    meta = out[off: off + n * D * 16].view(np.uint32).reshape(n, D, 4)
    off += n * D * 16
    pay = out[off:].view(np.uint32).reshape(n, 3)

    assert states.is_empty() and not any(states.errors), states
    depth = PE.chain_depths(hits)
    assert nrays == int(depth.sum())
    assert (depth >= 2).sum() > 0.1 * n and (depth == 1).sum() > 0.05 * n  # some chains go deep, some miss at once
    if bind == 3:
        assert (depth == 3).sum() > 0.05 * n
    assert np.array_equal(pay, PE.chain_payload(hits))
    # depth 0: the camera rays
    assert np.array_equal(rays[:, 0].view(np.uint32), PE.camera_rays(w, h, 0.0, 100.0, PE.CHAIN_ORIGIN).view(np.uint32))
    for d in range(D):
        live = depth > d
        ref, _, _ = orc.trace(rays[live, d])
        diff = compare(hits[live, d], ref)
        assert diff.ok() and diff.value_mismatch_bits == 0, (d, diff)
        if d + 1 < D:
            nxt = live & (hits[:, d, 0] != PE.MISS)
            attr, _ = orc.hit_attributes(rays[nxt, d], hits[nxt, d], meta[nxt, d])
            expect = PE.mirror_rays(rays[nxt, d], attr)
            assert np.array_equal(rays[nxt, d + 1].view(np.uint32), expect.view(np.uint32)), d


def test_hit_group_and_miss_dispatch(stages, device):
    """Rays with per-pixel sbt offset / stride / miss index on a scene with a
    two-geometry BLAS and instance hit-group offsets: each ray runs the
    closest-hit stage of hit group offset + stride * geometry + instance
    offset (ctx.rs:53-55; sbt.rs:18-39) or the miss stage of its miss index,
    and a group or index without a stage returns the payload as traced.  The
    hit context handed to the closest-hit stages equals the device fetch of
    rdrt_fetch_hit_transforms for the same hits."""
    import torch

    from tests.scenes import affine, build_pair
    from rendiation_amd import synth

    sph_p, sph_i = synth.uv_sphere(24, 12)
    tor_p, tor_i = synth.torus(32, 8)
    blas_list = [[(sph_p, sph_i, 1), (tor_p * np.float32(1.5), tor_i, 1)], [(sph_p * np.float32(0.7), sph_i, 1)]]
    insts = []
    for k in range(9):
        x, y = (k % 3 - 1) * 2.5, (k // 3 - 1) * 2.5
        insts.append({"transform": affine(None, (1.0, 1.0, 1.0), (x, y, -8.0 - 0.3 * k)), "blas": k % 2,
                      "sbt": [0, 1, 2][k % 3]})
    system, orc = build_pair(blas_list, [insts], [0], BuildOptions())
    w, h = 160, 120
    n = w * h
    _say(f"hit-group dispatch {w}x{h}")
    buf = torch.zeros((n * (4 + 32 + 16 + 16 + 160),), dtype=torch.uint8, device="cuda:0")
    torch.cuda.synchronize()
    src = GPURaytracingPipelineAndBindingSource(payload_words=1, hit_context=HIT_TRANSFORMS)
    rg = src.register_ray_gen(stages[6])
    chs = [src.register_ray_closest_hit(stages[7 + k]) for k in range(4)]
    mss = [src.register_ray_miss(stages[11 + k]) for k in range(2)]
    sbt = ShaderBindingTable(ray_type_count=1)
    sbt.config_ray_generation(rg)
    for k in range(4):
        sbt.config_hit_group(k, 0, 0, chs[k])  # hit group index k (stride 1 per geometry)
    for k in range(2):
        sbt.config_missing(k, mss[k])
    p = RayTracingPipeline(system, src, sbt)
    p.trace_ray(w, h, buf.data_ptr())
    system.sync()
    out = buf.cpu().numpy()
    states = p.read_back_execution_states()
    p.close()
    assert states.is_empty() and not any(states.errors), states
    o = 0
    pay = out[o:o + 4 * n].view(np.uint32); o += 4 * n
    rays = out[o:o + 32 * n].view(np.float32).reshape(n, 8); o += 32 * n
    hits = out[o:o + 16 * n].view(np.uint32).reshape(n, 4); o += 16 * n
    meta = out[o:o + 16 * n].view(np.uint32).reshape(n, 4); o += 16 * n
    xf = out[o:o + 160 * n].view(np.uint32).reshape(n, 40)
    cam = PE.camera_rays(w, h, 0.0, 100.0, PE.CHAIN_ORIGIN)
    ids = np.arange(n)
    inst_sbt = np.array([d["sbt"] for d in insts], dtype=np.int64)

    def group_of(m):  # hit group of a hit (geometry, instance) for pixel ids
        return (ids % 2) + ((ids // 2) % 2) * m[:, 0].astype(np.int64) + inst_sbt[m[:, 1] % len(insts)]

    miss_index = (ids // 4) % 3
    ran = pay != 0xFFFF
    hit = hits[:, 0] != PE.MISS
    # a stage ran: it recorded exactly the camera ray, and it was the stage of its hit group / miss index
    assert np.array_equal(rays[ran].view(np.uint32), cam[ran].view(np.uint32))
    g = group_of(meta)
    assert np.array_equal(pay[ran & hit], g[ran & hit])
    assert np.array_equal(pay[ran & ~hit], 100 + miss_index[ran & ~hit])
    # the hits the stages saw are the oracle's (closest-hit contract)
    ref, ref_meta, _ = orc.trace(cam, with_meta=True)
    d = compare(hits[ran], ref[ran])
    assert d.ok() and d.value_mismatch_bits == 0, d
    # no stage ran exactly where the oracle's hit has no stage (near ties aside)
    ref_hit = ref[:, 0] != PE.MISS
    ref_has_stage = np.where(ref_hit, group_of(ref_meta) < 4, miss_index < 2)
    assert (ran == ref_has_stage).mean() >= 1.0 - 1e-3
    for k in list(range(4)) + [100, 101]:
        assert (pay == k).sum() > 50, k  # every stage ran
    assert (~ran & ref_hit).sum() > 50 and (~ran & ~ref_hit).sum() > 50  # and both no-stage cases occur
    # the hit context the closest-hit stages saw is the device fetch's
    hk = ran & hit
    rd = torch.from_numpy(np.ascontiguousarray(cam[hk])).cuda()
    hd = torch.from_numpy(np.ascontiguousarray(hits[hk]).view(np.int32)).cuda()
    md = torch.from_numpy(np.ascontiguousarray(meta[hk]).view(np.int32)).cuda()
    fx = torch.zeros((int(hk.sum()), 40), dtype=torch.int32, device="cuda:0")
    system.hit_transforms(rd, hd, md, fx)
    system.sync()
    assert np.array_equal(xf[hk], fx.cpu().numpy().view(np.uint32))


def test_trace_pool_exhaustion_reports_and_drains(stages, device):
    """Nested traces need two trace slots per launch entry at once; with
    max_in_flight = 1 the trace group runs out of slots.  The spawns that
    find no slot set the group's error bit and their tasks resolve without
    tracing, so the call still drains (every task finished) and traces fewer
    rays than the full chain would."""
    import torch

    system, _ = reference_default_scene(BuildOptions(), bind=3)
    w, h, D = 96, 80, PE.DEPTH
    n = w * h
    _say("chain pipeline, trace pool of one slot per entry")
    nbytes = n * D * (32 + 16 + 16) + n * 3 * 4
    counts = []
    errors = []
    for mif in (1, 4):
        buf = torch.full((nbytes,), 0xFF, dtype=torch.uint8, device="cuda:0")
        torch.cuda.synchronize()
        p = _pipeline(system, stages, 3, payload_words=3, hit_context=HIT_ATTRIBUTES, max_in_flight=mif,
                      execution_round_hint=12)
        p.trace_ray(w, h, buf.data_ptr())
        system.sync()
        st = p.read_back_execution_states()
        counts.append(p.ray_count())
        errors.append(st.errors[0])
        p.close()
        assert st.is_empty(), (mif, st)
        assert not any(st.errors[1:]), (mif, st)
    assert errors[0] & 1 and errors[1] == 0, errors
    assert counts[0] < counts[1]
