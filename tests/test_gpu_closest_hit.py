"""GPU parity of the closest-hit kernels against the CPU oracle.

* RDRT_TRAVERSAL_REFERENCE must equal the oracle bit for bit: hit records,
  hit context and the per-ray node / triangle / instance visit counters.
* RDRT_TRAVERSAL_FAST must equal it except on near-ties (|dt| <= 1e-5 t),
  with bit-identical t, u, v wherever the primitive ids agree, over both
  node layouts (child pairs, child quads, quantized child quads).
"""

import numpy as np
import pytest

from oracle.oracle import OracleScene
from rendiation_amd import synth
from rendiation_amd.acceleration import (
    RAY_FLAG_ACCEPT_FIRST_HIT_AND_END_SEARCH,
    RAY_FLAG_CULL_BACK_FACING_TRIANGLES,
    RAY_FLAG_CULL_FRONT_FACING_TRIANGLES,
    RAY_FLAG_CULL_OPAQUE,
    RAY_FLAG_FORCE_NON_OPAQUE,
    RAY_FLAG_SKIP_TRIANGLES,
    TRAVERSAL_FAST,
    TRAVERSAL_REFERENCE,
    AccelerationStructureSystem,
    BuildOptions,
    TopLevelAccelerationStructureSourceInstance,
    TriangleGeometry,
)
from tests.parity import compare
from tests.scenes import build_pair as _build_pair, instanced_scene

pytestmark = pytest.mark.gpu


def _say(msg):
    print(f"[gpu-case] {msg}", flush=True)




def _gpu_trace(sys, rays_np, device, meta_np=None, traversal=TRAVERSAL_FAST, **kw):
    import torch

    rays = torch.from_numpy(np.ascontiguousarray(rays_np)).to(device)
    n = rays_np.shape[0]
    hits = torch.empty((n, 4), dtype=torch.int32, device=device)
    hm = torch.empty((n, 4), dtype=torch.int32, device=device)
    cnt = torch.empty((n, 3), dtype=torch.int32, device=device)
    meta = None if meta_np is None else torch.from_numpy(np.ascontiguousarray(meta_np).view(np.int32)).to(device)
    sys.trace(rays, hits, ray_meta=meta, hit_meta=hm, counters=cnt, traversal=traversal, **kw)
    sys.sync()
    return (hits.cpu().numpy().view(np.uint32), hm.cpu().numpy().view(np.uint32), cnt.cpu().numpy().view(np.uint32))


@pytest.fixture(scope="module")
def closest_hit_scene():
    """100k random triangles (leaf <= 4), 1M rays, seed 0."""
    pos = synth.random_triangles(0, synth.CLOSEST_HIT_TRIS)
    rays = synth.random_rays(0, synth.CLOSEST_HIT_RAYS)
    opts = BuildOptions(blas_bin_size=synth.CLOSEST_HIT_LEAF)
    sys, orc = _build_pair([[(pos, None, 1)]], [[{"transform": np.eye(4), "blas": 0}]], [0], opts)
    ref = orc.trace(rays, with_meta=True, with_counters=True)
    return sys, rays, ref


def test_reference_order_kernel_is_bit_exact(closest_hit_scene, device):
    sys, rays, (rh, rm, rc) = closest_hit_scene
    _say("closest-hit 100k tris x 1M rays, reference-order kernel")
    gh, gm, gc = _gpu_trace(sys, rays, device, traversal=TRAVERSAL_REFERENCE)
    assert np.array_equal(gh, rh), compare(gh, rh)
    assert np.array_equal(gm, rm)
    assert np.array_equal(gc, rc)


@pytest.mark.parametrize("wide", [0, 1, 2])
def test_fast_kernel_matches_except_near_ties(closest_hit_scene, wide, device):
    sys, rays, (rh, rm, _) = closest_hit_scene
    _say(f"closest-hit 100k tris x 1M rays, fast kernel, wide={wide}")
    sys.set_tune(wide=wide)
    try:
        gh, gm, _ = _gpu_trace(sys, rays, device, traversal=TRAVERSAL_FAST)
    finally:
        sys.set_tune()
    d = compare(gh, rh)
    print(d)
    assert d.ok(), d
    assert d.value_mismatch_bits == 0, d
    same = gh[:, 0] == rh[:, 0]
    assert np.array_equal(gm[same], rm[same])


@pytest.mark.parametrize("wide,nrays", [(0, None), (1, None), (2, None), (0, 30_000), (1, 30_000), (2, 30_000)])
def test_cooperative_kernel_matches_except_near_ties(closest_hit_scene, wide, nrays, device):
    """Every launch through the cooperative kernel (8 lanes per ray; 16 for
    the launches of at most 32,768 rays, here the first 30,000 rays): the
    same hits as the oracle except near ties, as the one-ray-per-lane
    kernel."""
    sys, rays, (rh, rm, _) = closest_hit_scene
    if nrays is not None:
        rays, rh, rm = rays[:nrays], rh[:nrays], rm[:nrays]
    _say(f"closest-hit 100k tris x {rays.shape[0]} rays, cooperative kernel, wide={wide}")
    sys.set_tune(wide=wide, coop=1 << 22)
    try:
        gh, gm = _gpu_trace_nc(sys, rays, device)
    finally:
        sys.set_tune()
    d = compare(gh, rh)
    print(d)
    assert d.ok(), d
    assert d.value_mismatch_bits == 0, d
    same = gh[:, 0] == rh[:, 0]
    assert np.array_equal(gm[same], rm[same])


def _gpu_trace_nc(sys, rays_np, device, meta_np=None, **kw):
    """A fast launch without visit counters (counting keeps one ray per lane)."""
    import torch

    rays = torch.from_numpy(np.ascontiguousarray(rays_np)).to(device)
    n = rays_np.shape[0]
    hits = torch.empty((n, 4), dtype=torch.int32, device=device)
    hm = torch.empty((n, 4), dtype=torch.int32, device=device)
    meta = None if meta_np is None else torch.from_numpy(np.ascontiguousarray(meta_np).view(np.int32)).to(device)
    sys.trace(rays, hits, ray_meta=meta, hit_meta=hm, traversal=TRAVERSAL_FAST, **kw)
    sys.sync()
    return hits.cpu().numpy().view(np.uint32), hm.cpu().numpy().view(np.uint32)


@pytest.mark.parametrize("traversal,wide,coop", [(TRAVERSAL_REFERENCE, 0, 0), (TRAVERSAL_FAST, 0, 0),
                                                 (TRAVERSAL_FAST, 1, 0), (TRAVERSAL_FAST, 2, 0),
                                                 (TRAVERSAL_FAST, 1, 1 << 22), (TRAVERSAL_FAST, 2, 1 << 22)])
def test_instanced_scene_with_flags(traversal, wide, coop, device):
    blas_list, tlas_list, binding, rays, meta = instanced_scene(seed=3, nrays=60_000)
    opts = BuildOptions()
    sys, orc = _build_pair(blas_list, tlas_list, binding, opts)
    _say(f"instanced scene, traversal={traversal} wide={wide} coop={coop}")
    rh, rm, rc = orc.trace(rays, meta=meta, with_meta=True, with_counters=True)
    sys.set_tune(wide=wide, coop=coop)
    if coop:
        gh, gm = _gpu_trace_nc(sys, rays, device, meta_np=meta)
        gc = None
    else:
        gh, gm, gc = _gpu_trace(sys, rays, device, meta_np=meta, traversal=traversal)
    end_search = (meta[:, 0] & RAY_FLAG_ACCEPT_FIRST_HIT_AND_END_SEARCH) != 0
    if traversal == TRAVERSAL_REFERENCE:
        assert np.array_equal(gh, rh), compare(gh, rh)
        assert np.array_equal(gm, rm)
        assert np.array_equal(gc, rc)
    else:
        # ACCEPT_FIRST_HIT_AND_END_SEARCH returns any hit: compare hit/miss only
        assert np.array_equal(gh[end_search, 0] == 0xFFFFFFFF, rh[end_search, 0] == 0xFFFFFFFF)
        d = compare(gh[~end_search], rh[~end_search])
        print(d)
        assert d.ok(), d
        keep = ~end_search & (gh[:, 0] == rh[:, 0])
        assert np.array_equal(gm[keep], rm[keep])


@pytest.mark.parametrize("flags", [0, RAY_FLAG_CULL_BACK_FACING_TRIANGLES, RAY_FLAG_CULL_FRONT_FACING_TRIANGLES,
                                   RAY_FLAG_SKIP_TRIANGLES, RAY_FLAG_CULL_OPAQUE,
                                   RAY_FLAG_FORCE_NON_OPAQUE | RAY_FLAG_CULL_OPAQUE])
def test_uniform_ray_flags(flags, device):
    pos = synth.random_triangles(5, 4000, spread=0.05)
    rays = synth.random_rays(5, 20_000)
    sys, orc = _build_pair([[(pos, None, 1)]], [[{"transform": np.eye(4), "blas": 0}]], [0], BuildOptions())
    _say(f"uniform flags {flags:#x}")
    rh, _, rc = orc.trace(rays, flags=flags, with_counters=True)
    for trav, wide in ((TRAVERSAL_REFERENCE, 0), (TRAVERSAL_FAST, 0), (TRAVERSAL_FAST, 1), (TRAVERSAL_FAST, 2)):
        sys.set_tune(wide=wide)
        gh, _, gc = _gpu_trace(sys, rays, device, traversal=trav, flags=flags)
        if trav == TRAVERSAL_REFERENCE:
            assert np.array_equal(gh, rh)
            assert np.array_equal(gc, rc)
        else:
            assert compare(gh, rh).ok()
            sys.set_tune(wide=wide, coop=1 << 22)  # the cooperative kernel
            gh, _ = _gpu_trace_nc(sys, rays, device, flags=flags)
            assert compare(gh, rh).ok()
    sys.set_tune()


def test_edge_cases(device):
    import torch

    pos = synth.random_triangles(7, 100)
    sys, orc = _build_pair([[(pos, None, 1)]], [[{"transform": np.eye(4), "blas": 0}]], [0], BuildOptions())
    _say("edge cases: empty launch, unbound tlas index")
    empty = torch.empty((0, 8), dtype=torch.float32, device=device)
    out = torch.empty((0, 4), dtype=torch.int32, device=device)
    sys.trace(empty, out)
    sys.sync()
    rays = synth.random_rays(7, 1000)
    for trav in (TRAVERSAL_REFERENCE, TRAVERSAL_FAST):
        gh, _, _ = _gpu_trace(sys, rays, device, traversal=trav, tlas_idx=5)
        assert np.all(gh[:, 0] == 0xFFFFFFFF)
        rh, _, _ = orc.trace(rays, tlas_idx=5)
        assert np.array_equal(gh, rh)


def test_launch_shapes_agree_and_autotune(closest_hit_scene, device):
    """Every launch shape of the search space traverses in the same order, so
    hits are bit-identical to the default shape; the autotuner keeps one."""
    import torch

    sys, rays_np, _ = closest_hit_scene
    _say("launch shapes + autotune, 100k tris x 1M rays")
    base = {}
    for wide in (0, 1, 2):  # each layout has its own visit order
        sys.set_tune(wide=wide)
        base[wide], _, _ = _gpu_trace(sys, rays_np, device)
    rays = torch.from_numpy(np.ascontiguousarray(rays_np)).to(device)
    hits = torch.empty((rays_np.shape[0], 4), dtype=torch.int32, device=device)
    for shape in [(64, 8, 0, 0, 0, 0), (128, 16, 8, 1, 0, 0), (256, 8, 8, 0, 1, 0), (256, 16, 0, 1, 1, 0),
                  (64, 16, 8, 0, 1, 0), (64, 8, 0, 0, 0, 1), (128, 16, 8, 1, 0, 1), (256, 8, 8, 0, 1, 1),
                  (256, 16, 0, 1, 1, 1), (64, 16, 0, 0, 1, 2), (128, 8, 8, 1, 0, 2), (256, 16, 0, 1, 1, 2),
                  (64, 16, 6, 1, 0, 1), (256, 8, 6, 0, 1, 2), (128, 16, 6, 0, 0, 0)]:
        sys.set_tune(*shape)
        sys.trace(rays, hits)
        sys.sync()
        assert np.array_equal(hits.cpu().numpy().view(np.uint32), base[shape[5]]), shape
    choice = sys.autotune(rays, hits, reps=2)
    print("autotune:", choice)
    assert choice["block_size"] in (64, 128, 256) and choice["lds_entries"] in (8, 16) and choice["ms"] > 0
    sys.trace(rays, hits)
    sys.sync()
    assert np.array_equal(hits.cpu().numpy().view(np.uint32), base[choice["wide"]])
    sys.set_tune()
