"""GPU parity of the device BLAS build (rdrt_set_device_build).

The device builds both trees of every geometry level by level with the host
builders' split arithmetic, so the built layouts equal the host build's: the
reference forest node for node (boxes, links, ranges) and the fast tree's
child quads exactly; a reference median split may choose differently among
triangles whose centres tie, so leaves are compared by their triangles'
centres.  Traversal after a device build
equals the oracle bit for bit in the reference order.
"""

import numpy as np
import pytest

from rendiation_amd import synth
from rendiation_amd.acceleration import (
    TRAVERSAL_REFERENCE,
    AccelerationStructureSystem,
    BuildOptions,
    TopLevelAccelerationStructureSourceInstance,
    TriangleGeometry,
)
from tests.scenes import build_pair

pytestmark = pytest.mark.gpu


def _say(msg):
    print(f"[gpu-case] {msg}", flush=True)


def _geometries():
    rng = np.random.default_rng(11)
    yield "soup", synth.random_triangles(0, 20_000), None
    hp, hi, _ = synth.height_field(80, seed=1)
    yield "height field", hp, hi
    # coincident centres: duplicated triangles (object medians in both trees)
    t = synth.random_triangles(2, 50)
    yield "duplicates", np.tile(t.reshape(-1, 9), (40, 1)).reshape(-1, 3), None
    # one large triangle over a tight cluster: every centre in one reference bin
    big = np.array([[-50, -50, 0], [50, -50, 0], [0, 50, 0]], dtype=np.float32)
    small = (synth.random_triangles(3, 300) * np.float32(0.01)).astype(np.float32)
    yield "cluster under a large triangle", np.concatenate([big, small]), None
    # flat: zero extent along one axis
    xy = rng.uniform(-1, 1, size=(3000, 3, 2)).astype(np.float32)
    yield "flat", np.concatenate([xy, np.zeros((3000, 3, 1), np.float32)], axis=2).reshape(-1, 3), None
    yield "one triangle", synth.random_triangles(4, 1), None


def _system(pos, idx, device, opts):
    s = AccelerationStructureSystem(device=device, options=opts)
    if device is not None and device is not False:
        s.set_device_build(True)
    b = s.create_bottom_level_acceleration_structure([TriangleGeometry(pos, idx)])
    t = s.create_top_level_acceleration_structure([TopLevelAccelerationStructureSourceInstance(np.eye(4), b)])
    s.bind_tlas([t])
    s.build()
    return s


def _centres(tris):
    """Box centres of triangles (n, 3, 3), as the builders compute them."""
    return (tris.min(axis=1) + tris.max(axis=1)) * np.float32(0.5)


def _leaf_sets(forest, redirect, cen):
    """{(begin, end): sorted box centres of its triangles} of the forest's
    leaves: a median split chooses freely among triangles whose centres tie
    (the two triangles of a height-field cell share one), so leaves are
    compared by centres, not triangle ids."""
    out = {}
    for n in forest:
        if n[3] == n[7]:  # hit == miss: a leaf
            b, e = int(n[8]), int(n[9])
            out[(b, e)] = tuple(sorted(tuple(cen[t].tolist()) for t in redirect[b:e]))
    return out


@pytest.mark.parametrize("name,pos,idx", list(_geometries()), ids=[g[0] for g in _geometries()])
def test_device_build_equals_host_build(name, pos, idx, device):
    opts = BuildOptions()
    _say(f"device build vs host build: {name}")
    host = _system(pos, idx, None, opts)
    dev = _system(pos, idx, 0, opts)
    hf, df = host.copy_built("blas_forest"), dev.copy_built("blas_forest")
    assert hf.shape == df.shape, name
    # boxes (as floats: +0 and -0 compare equal), links and ranges
    assert np.array_equal(hf[:, [0, 1, 2, 4, 5, 6]].view(np.float32), df[:, [0, 1, 2, 4, 5, 6]].view(np.float32)), name
    assert np.array_equal(hf[:, [3, 7, 8, 9]], df[:, [3, 7, 8, 9]]), name
    hr, dr = host.copy_built("redirect"), dev.copy_built("redirect")
    tris = pos.reshape(-1, 3)[idx.reshape(-1)].reshape(-1, 3, 3) if idx is not None else pos.reshape(-1, 3, 3)
    cen = _centres(tris)
    assert _leaf_sets(hf, hr, cen) == _leaf_sets(df, dr, cen), name
    hq, dq = host.copy_built("blas_quads"), dev.copy_built("blas_quads")
    assert hq.shape == dq.shape, name
    assert np.array_equal(hq[:, 24:28], dq[:, 24:28]), name  # refs: same tree, same leaves
    assert np.array_equal(hq[:, :24].view(np.float32), dq[:, :24].view(np.float32)), name
    assert host.scene_info().max_blas_depth == dev.scene_info().max_blas_depth
    host.close()
    dev.close()


def test_device_built_scene_traces_like_the_oracle(device):
    """100k random triangles, 1M rays: after a device build the reference-order
    kernel equals the oracle bit for bit (hits, contexts, visit counters)."""
    import torch

    pos = synth.random_triangles(0, synth.CLOSEST_HIT_TRIS)
    rays_np = synth.random_rays(0, synth.CLOSEST_HIT_RAYS)
    opts = BuildOptions(blas_bin_size=synth.CLOSEST_HIT_LEAF)
    _say("device-built closest-hit scene vs the oracle")
    _, orc = build_pair([[(pos, None, 1)]], [[{"transform": np.eye(4), "blas": 0}]], [0], opts)
    rh, rm, rc = orc.trace(rays_np, with_meta=True, with_counters=True)
    s = _system(pos, None, 0, opts)
    s.create_comp_instance()
    rays = torch.from_numpy(rays_np).to(device)
    n = rays_np.shape[0]
    hits = torch.empty((n, 4), dtype=torch.int32, device=device)
    hm = torch.empty((n, 4), dtype=torch.int32, device=device)
    cnt = torch.empty((n, 3), dtype=torch.int32, device=device)
    s.trace(rays, hits, hit_meta=hm, counters=cnt, traversal=TRAVERSAL_REFERENCE)
    s.sync()
    assert np.array_equal(hits.cpu().numpy().view(np.uint32), rh)
    assert np.array_equal(hm.cpu().numpy().view(np.uint32), rm)
    assert np.array_equal(cnt.cpu().numpy().view(np.uint32), rc)
    s.close()


def test_device_build_of_many_geometries(device):
    """Every geometry of a scene in one device build: several BLAS (one
    with two geometries), each equal to its host build."""
    opts = BuildOptions()
    hp, hi, _ = synth.height_field(40, seed=6)
    blas = [[(synth.random_triangles(7, 5000), None)], [(hp, hi), (synth.random_triangles(8, 1), None)],
            [(synth.random_triangles(9, 300, spread=0.2), None)]]
    _say("device build of several BLAS at once")

    def make(device_build):
        s = AccelerationStructureSystem(device=0 if device_build else None, options=opts)
        if device_build:
            s.set_device_build(True)
        hs = [s.create_bottom_level_acceleration_structure([TriangleGeometry(p, i) for p, i in geos]) for geos in blas]
        t = s.create_top_level_acceleration_structure(
            [TopLevelAccelerationStructureSourceInstance(np.eye(4), h) for h in hs])
        s.bind_tlas([t])
        s.build()
        return s

    host, dev = make(False), make(True)
    for what in ("blas_forest", "blas_quads"):
        a, b = host.copy_built(what), dev.copy_built(what)
        assert a.shape == b.shape, what
        nf = 24 if what == "blas_quads" else None
        if nf:
            assert np.array_equal(a[:, :nf].view(np.float32), b[:, :nf].view(np.float32))
            assert np.array_equal(a[:, nf:], b[:, nf:])
        else:
            cols = [0, 1, 2, 4, 5, 6]
            assert np.array_equal(a[:, cols].view(np.float32), b[:, cols].view(np.float32))
            assert np.array_equal(a[:, [3, 7, 8, 9]], b[:, [3, 7, 8, 9]])
    forest, hr, dr = host.copy_built("blas_forest"), host.copy_built("redirect"), dev.copy_built("redirect")
    tris = np.concatenate([p.reshape(-1, 3)[i.reshape(-1)].reshape(-1, 3, 3) if i is not None else p.reshape(-1, 3, 3)
                           for geos in blas for p, i in geos])
    cen = _centres(tris)
    assert _leaf_sets(forest, hr, cen) == _leaf_sets(forest, dr, cen)
    host.close()
    dev.close()
