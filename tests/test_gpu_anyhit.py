"""GPU parity of any-hit behaviour tables (rdrt_blas_set_anyhit) with the
oracle, on the instanced scene with per-ray flags (non-opaque geometry,
instances and rays forcing opacity either way):

* accept / ignore tables: the reference-order kernel bit for bit (hits,
  hit context, visit counters); every fast layout and the cooperative
  kernel under the closest-hit contract;
* tables with end-search behaviours: the reference-order kernel bit for
  bit (the reference's own visit order decides where the search stops);
* the path tracer's opaque frames are untouched (no table: null pointer).
"""

import numpy as np
import pytest

from rendiation_amd import synth
from rendiation_amd.acceleration import (
    RAY_FLAG_ACCEPT_FIRST_HIT_AND_END_SEARCH,
    TRAVERSAL_FAST,
    TRAVERSAL_REFERENCE,
    BuildOptions,
)
from tests.parity import compare
from tests.scenes import build_pair, instanced_scene
from tests.test_gpu_closest_hit import _gpu_trace, _gpu_trace_nc

pytestmark = pytest.mark.gpu


def _tables(blas_list, seed, with_end):
    out = {}
    for b, geos in enumerate(blas_list):
        for g, (pos, idx, _) in enumerate(geos):
            n = np.asarray(idx).reshape(-1, 3).shape[0] if idx is not None else pos.shape[0] // 3
            r = synth.splitmix64(seed, 100 + 10 * b + g, n)
            if with_end:
                t = (r % np.uint64(16)).astype(np.uint8)
                beh = np.where(t < 9, 1, np.where(t < 13, 0, np.where(t < 15, 3, 2))).astype(np.uint8)
            else:
                beh = ((r % np.uint64(10)) < np.uint64(6)).astype(np.uint8)  # 60 % accept, 40 % ignore
            out[(b, g)] = beh
    return out


_SCENES = {}


def _scene(with_end):
    if with_end not in _SCENES:
        blas_list, tlas_list, binding, rays, meta = instanced_scene(seed=4, nrays=60_000)
        tables = _tables(blas_list, 4, with_end)
        sys, orc = build_pair(blas_list, tlas_list, binding, BuildOptions(), anyhit=tables)
        ref = orc.trace(rays, meta=meta, with_meta=True, with_counters=True)
        _SCENES[with_end] = (with_end, sys, rays, meta, ref)
    return _SCENES[with_end]


@pytest.fixture(scope="module", params=[False, True], ids=["accept_ignore", "with_end_search"])
def anyhit_scene(request, device):
    return _scene(request.param)


@pytest.fixture(scope="module")
def accept_ignore_scene(device):
    return _scene(False)


def test_reference_order_kernel_is_bit_exact(anyhit_scene, device):
    with_end, sys, rays, meta, (rh, rm, rc) = anyhit_scene
    print(f"[gpu-case] any-hit tables (end search {with_end}), reference-order kernel", flush=True)
    gh, gm, gc = _gpu_trace(sys, rays, device, meta_np=meta, traversal=TRAVERSAL_REFERENCE)
    assert np.array_equal(gh, rh), compare(gh, rh)
    assert np.array_equal(gm, rm)
    assert np.array_equal(gc, rc)
    assert (rh[:, 0] != 0xFFFFFFFF).sum() > 1000


@pytest.mark.parametrize("wide,coop", [(0, 0), (1, 0), (2, 0), (1, 1 << 22)])
def test_fast_kernels_match(accept_ignore_scene, wide, coop, device):
    """Accept / ignore tables (where an end-search behaviour stops depends on
    the visit order, which the reference-order kernel shares with the
    oracle)."""
    _, sys, rays, meta, (rh, rm, _) = accept_ignore_scene
    print(f"[gpu-case] any-hit tables, fast kernel wide={wide} coop={coop}", flush=True)
    sys.set_tune(wide=wide, coop=coop)
    try:
        if coop:
            gh, gm = _gpu_trace_nc(sys, rays, device, meta_np=meta)
        else:
            gh, gm, _ = _gpu_trace(sys, rays, device, meta_np=meta, traversal=TRAVERSAL_FAST)
    finally:
        sys.set_tune()
    end_ray = (meta[:, 0] & RAY_FLAG_ACCEPT_FIRST_HIT_AND_END_SEARCH) != 0
    assert np.array_equal(gh[end_ray, 0] == 0xFFFFFFFF, rh[end_ray, 0] == 0xFFFFFFFF)
    d = compare(gh[~end_ray], rh[~end_ray])
    assert d.ok() and d.value_mismatch_bits == 0, d
    keep = ~end_ray & (gh[:, 0] == rh[:, 0])
    assert np.array_equal(gm[keep], rm[keep])
