"""GPU parity of batched multi-scene tracing (rdrt_trace_batched).

Rays of several independent scenes share one queue; each round traces the
live rays, spawns a continuation ray at every hit and compacts the survivors.

* Every round's hits equal the oracle's on the same input rays (round 0: the
  primary rays; round r: the rays the GPU spawned), except near-ties.
* Spawned rays equal the oracle's bit for bit (both compute the spawn in
  FP32 with the same operations).
* Rays that ended carry miss records in every later round, and the per-round
  queue sizes equal the number of hits of the previous round.
* The persistent graph-captured mode and the per-scene sequential mode give
  bit-identical results; graph replays are identical to the first run.
"""

import numpy as np
import pytest

from rendiation_amd import synth
from rendiation_amd._lib import RdrtError
from rendiation_amd.acceleration import BATCH_PERSISTENT, BATCH_SEQUENTIAL, BuildOptions
from tests.parity import MISS, compare
from tests.scenes import batched_pair

pytestmark = pytest.mark.gpu

BOUNCES = 3
SEED = 11


def _say(msg):
    print(f"[gpu-case] {msg}", flush=True)


@pytest.fixture(scope="module")
def batch_case():
    scenes, rays, meta, offs = synth.batched_scenes(seed=5, nscenes=6, tris=(300, 6000), rays=(1000, 20000),
                                                    spread=0.08)
    sys, orc = batched_pair(scenes, BuildOptions(blas_bin_size=4))
    return sys, orc, scenes, rays, meta, offs


def _run(sys, rays, meta, device, mode, offs=None, bounces=BOUNCES, seed=SEED):
    import torch

    n = rays.shape[0]
    d_rays = torch.from_numpy(np.ascontiguousarray(rays)).to(device)
    d_meta = torch.from_numpy(np.ascontiguousarray(meta).view(np.int32)).to(device)
    hits = torch.zeros((bounces + 1) * n * 4, dtype=torch.int32, device=device)
    spawned = torch.zeros(max(bounces, 1) * n * 8, dtype=torch.float32, device=device)
    sys.trace_batched(d_rays, d_meta, hits, bounces=bounces, seed=seed, spawned=spawned if bounces else None,
                      scene_offsets=offs, mode=mode)
    st = sys.batch_stats()
    h = hits.cpu().numpy().view(np.uint32).reshape(bounces + 1, n, 4)
    sp = spawned.cpu().numpy().reshape(-1, n, 8)
    return h, sp, st


def _dead_rays(rays, dead):
    """Inputs for the oracle at paths that are not traced: an empty range."""
    out = rays.copy()
    out[dead] = np.array([0, 0, 0, 1, 0, 0, 1, 0], dtype=np.float32)
    return out


@pytest.mark.parametrize("wide", [0, 1, 2])
def test_rounds_match_oracle(batch_case, wide, device):
    sys, orc, _, rays, meta, _ = batch_case
    n = rays.shape[0]
    _say(f"batched persistent: 6 scenes, {n} rays, {BOUNCES} bounces, wide={wide}")
    sys.set_tune(wide=wide)
    try:
        h, sp, st = _run(sys, rays, meta, device, BATCH_PERSISTENT)
    finally:
        sys.set_tune()
    assert st.rounds == BOUNCES + 1
    live = np.ones(n, dtype=bool)
    for r in range(BOUNCES + 1):
        inp = rays if r == 0 else sp[r - 1]
        oh, osp, _ = orc.batch_round(_dead_rays(inp, ~live), meta, r + 1, SEED)
        assert st.rays[r] == int(live.sum()), (r, st.rays)
        d = compare(h[r][live], oh[live])
        assert d.ok(), (r, d)
        assert d.value_mismatch_bits == 0, (r, d)
        ended = h[r][~live]
        assert np.all(ended[:, 0] == MISS) and np.all(ended[:, 1:] == 0), r
        if r == BOUNCES:
            break
        nxt = live & (h[r][:, 0] != MISS)
        same = nxt & (h[r][:, 0] == oh[:, 0])
        # the oracle computes the spawn in FP32 with the device's operations
        assert np.array_equal(sp[r][same], osp[same]), r
        live = nxt
    assert st.rays[1] > 0 and st.rays[2] > 0, st.rays


def test_spawned_rays_are_cosine_distributed(batch_case, device):
    """Unit directions in the hemisphere of the geometric normal facing the
    ray, with mean cosine 2/3 (cosine-weighted hemisphere)."""
    sys, _, scenes, rays, meta, _ = batch_case
    _say("batched spawn distribution")
    h, sp, _ = _run(sys, rays, meta, device, BATCH_PERSISTENT)
    cos_all = []
    for r in range(BOUNCES):
        inp = rays if r == 0 else sp[r - 1]
        alive = np.nonzero(h[r][:, 0] != MISS)[0]
        scene = meta[alive, 2]
        prim = h[r][alive, 0]
        tri = np.stack([scenes[k][3 * p:3 * p + 3] for k, p in zip(scene, prim)]).astype(np.float64)
        ng = np.cross(tri[:, 1] - tri[:, 0], tri[:, 2] - tri[:, 0])
        ng /= np.linalg.norm(ng, axis=1, keepdims=True)
        din = inp[alive, 4:7].astype(np.float64)
        ng[np.sum(ng * din, axis=1) > 0] *= -1.0
        d = sp[r][alive, 4:7].astype(np.float64)
        assert np.all(np.abs(np.linalg.norm(d, axis=1) - 1.0) < 1e-5)
        c = np.sum(d * ng, axis=1)
        assert np.all(c > -1e-6), r
        cos_all.append(c)
        pos = inp[alive, 0:3].astype(np.float64) + din * h[r][alive, 1].view(np.float32)[:, None]
        assert np.all(np.abs(sp[r][alive, 0:3] - pos) < 1e-3)
    c = np.concatenate(cos_all)
    se = np.sqrt(1.0 / 18.0 / c.size)
    assert abs(c.mean() - 2.0 / 3.0) < 5 * se, (c.mean(), c.size)


def test_persistent_equals_sequential(batch_case, device):
    sys, _, _, rays, meta, offs = batch_case
    _say("batched persistent vs sequential")
    hp, spp_, stp = _run(sys, rays, meta, device, BATCH_PERSISTENT)
    hs, sps, sts = _run(sys, rays, meta, device, BATCH_SEQUENTIAL, offs=offs)
    assert np.array_equal(hp, hs)
    assert stp.rays == sts.rays
    for r in range(BOUNCES):
        alive = hp[r][:, 0] != MISS
        assert np.array_equal(spp_[r][alive], sps[r][alive]), r
    assert stp.launches == 3 * (BOUNCES + 1)
    assert sts.launches == 1 + 6 * (2 * BOUNCES + 1)


def test_graph_replay_and_recapture(batch_case, device):
    sys, _, _, rays, meta, _ = batch_case
    _say("batched graph replay")
    h1, s1, _ = _run(sys, rays, meta, device, BATCH_PERSISTENT)
    h2, s2, _ = _run(sys, rays, meta, device, BATCH_PERSISTENT)
    h3, s3, _ = _run(sys, rays, meta, device, BATCH_PERSISTENT, seed=SEED + 1)
    h4, s4, _ = _run(sys, rays, meta, device, BATCH_PERSISTENT)
    assert np.array_equal(h1, h2) and np.array_equal(h1, h4)
    assert np.array_equal(h1[0], h3[0])  # round 0 does not depend on the seed
    alive = h1[0][:, 0] != MISS
    assert not np.array_equal(s1[0][alive], s3[0][alive])
    assert np.array_equal(s1[0][alive], s4[0][alive])


def test_empty_scene_and_no_bounce(batch_case, device):
    sys, orc, _, rays, meta, offs = batch_case
    _say("batched: empty scene range, zero bounces")
    # scene 0 given no rays: drop its range
    k0 = int(offs[1])
    sub_rays, sub_meta = rays[k0:], meta[k0:]
    sub_offs = np.concatenate([[0, 0], offs[1:] - k0])
    hs, _, sts = _run(sys, sub_rays, sub_meta, device, BATCH_SEQUENTIAL, offs=sub_offs)
    hp, _, _ = _run(sys, sub_rays, sub_meta, device, BATCH_PERSISTENT)
    assert np.array_equal(hs, hp)
    h0, _, st0 = _run(sys, rays, meta, device, BATCH_PERSISTENT, bounces=0)
    oh, _ = orc.trace(rays, meta)[:2]
    d = compare(h0[0], oh)
    assert d.ok() and d.value_mismatch_bits == 0, d
    assert st0.rounds == 1 and st0.rays == [rays.shape[0]]


def test_invalid_batch_arguments(batch_case, device):
    sys, _, _, rays, meta, offs = batch_case
    _say("batched: argument errors")
    with pytest.raises(RdrtError):
        _run(sys, rays, meta, device, BATCH_SEQUENTIAL, offs=None)
    bad = offs.copy()
    bad[-1] -= 1
    with pytest.raises(RdrtError):
        _run(sys, rays, meta, device, BATCH_SEQUENTIAL, offs=bad)
    with pytest.raises(RdrtError):
        _run(sys, rays, meta, device, BATCH_PERSISTENT, bounces=17)


def test_cull_pass_equals_plain_rounds(batch_case, device):
    """With the cull pass (rays missing their scene's root box finish before
    the traversal, the rest are queued densely), every round is bit-identical
    to the plain identity queue."""
    sys, _, _, rays, meta, _ = batch_case
    _say("batched: cull pass vs identity queue")
    hp, sp, stp = _run(sys, rays, meta, device, BATCH_PERSISTENT)
    sys.set_tune(cull=1)
    try:
        hc, sc, stc = _run(sys, rays, meta, device, BATCH_PERSISTENT)
    finally:
        sys.set_tune()
    assert np.array_equal(hp, hc)
    for r in range(BOUNCES):
        alive = hp[r][:, 0] != MISS
        assert np.array_equal(sp[r][alive], sc[r][alive]), r
    assert stc.rays[0] < stp.rays[0]  # round 0 traverses only the rays that enter a scene box
    assert stc.rays[1:] == stp.rays[1:]


def test_batch_autotune(batch_case, device):
    """The batched tuner keeps one shape; persistent and sequential calls
    with it stay bit-identical, and equal the oracle except near-ties."""
    import torch

    sys, orc, _, rays, meta, offs = batch_case
    _say("batched autotune")
    n = rays.shape[0]
    d_rays = torch.from_numpy(np.ascontiguousarray(rays)).to(device)
    d_meta = torch.from_numpy(np.ascontiguousarray(meta).view(np.int32)).to(device)
    hits = torch.zeros((BOUNCES + 1) * n * 4, dtype=torch.int32, device=device)
    try:
        choice = sys.batch_autotune(d_rays, d_meta, hits, bounces=BOUNCES, seed=SEED, reps=1)
        print("batch autotune:", choice)
        assert choice["block_size"] in (64, 128, 256) and choice["ms"] > 0
        hp, _, _ = _run(sys, rays, meta, device, BATCH_PERSISTENT)
        hs, _, _ = _run(sys, rays, meta, device, BATCH_SEQUENTIAL, offs=offs)
    finally:
        sys.set_tune()
    assert np.array_equal(hp, hs)
    oh, _ = orc.trace(rays, meta)[:2]
    assert compare(hp[0], oh).ok()
