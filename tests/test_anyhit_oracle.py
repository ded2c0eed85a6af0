"""CPU pins of the oracle's any-hit behaviour (traverse_cpu.rs:158-185):
on one non-opaque triangle, ignore and end-without-accept miss while accept
hits; a table that ignores everything equals culling non-opaque geometry;
accepting everything equals no table; opaque geometry and forced-opaque
rays ignore the table, forced-non-opaque rays use it on opaque geometry;
and with accept / ignore tables the BVH walk equals the brute-force
nearest accepted hit."""

import numpy as np

from oracle.oracle import OracleScene
from rendiation_amd import synth
from rendiation_amd.acceleration import (
    ANYHIT_ACCEPT,
    ANYHIT_END_SEARCH,
    RAY_FLAG_CULL_NON_OPAQUE,
    RAY_FLAG_FORCE_NON_OPAQUE,
    RAY_FLAG_FORCE_OPAQUE,
)

MISS = 0xFFFFFFFF


def _one_triangle(behavior, geo_flags=0):
    o = OracleScene()
    tri = np.array([[-1, -1, -5], [1, -1, -5], [0, 1, -5]], dtype=np.float32)
    o.add_blas([(tri, None, geo_flags)])
    o.add_tlas([{"transform": np.eye(4), "blas": 0}])
    if behavior is not None:
        o.set_anyhit(0, 0, [behavior])
    o.bind([0])
    o.build()
    ray = np.array([[0, 0, 0, 0, 0, 0, -1, 100]], dtype=np.float32)
    return o, ray


def test_single_triangle_behaviours():
    for beh, hit in ((None, True), (0, False), (ANYHIT_ACCEPT, True), (ANYHIT_END_SEARCH, False),
                     (ANYHIT_ACCEPT | ANYHIT_END_SEARCH, True)):
        o, ray = _one_triangle(beh)
        h, _, _ = o.trace(ray)
        assert (h[0, 0] != MISS) == hit, beh
        if hit:
            assert abs(h[0].view(np.float32)[1] - 5.0) < 1e-6
    # opaque geometry and forced-opaque rays ignore the table; forced non-opaque rays use it
    o, ray = _one_triangle(0, geo_flags=1)
    assert o.trace(ray)[0][0, 0] != MISS
    assert o.trace(ray, flags=RAY_FLAG_FORCE_NON_OPAQUE)[0][0, 0] == MISS
    o, ray = _one_triangle(0)
    assert o.trace(ray, flags=RAY_FLAG_FORCE_OPAQUE)[0][0, 0] != MISS


def _soup_scene(tables=None, flags=0):
    pos = synth.random_triangles(9, 3000, spread=0.05)
    o = OracleScene()
    o.add_blas([(pos, None, flags)])
    o.add_tlas([{"transform": np.eye(4), "blas": 0}])
    if tables is not None:
        o.set_anyhit(0, 0, tables)
    o.bind([0])
    o.build()
    return o, pos.shape[0] // 3


def test_tables_against_culling_and_brute_force():
    rays = synth.random_rays(9, 20000)
    plain, n = _soup_scene()
    ignore_all, _ = _soup_scene(np.zeros(n, np.uint8))
    accept_all, _ = _soup_scene(np.full(n, ANYHIT_ACCEPT, np.uint8))
    ref, _, _ = plain.trace(rays)
    assert np.array_equal(accept_all.trace(rays)[0], ref)
    culled, _, _ = plain.trace(rays, flags=RAY_FLAG_CULL_NON_OPAQUE)
    assert np.array_equal(ignore_all.trace(rays)[0], culled)
    assert (culled[:, 0] == MISS).all() and (ref[:, 0] != MISS).sum() > 1000
    # half the triangles ignored: the walk finds the nearest accepted hit
    table = (synth.splitmix64(9, 77, n) & np.uint64(1)).astype(np.uint8)
    half, _ = _soup_scene(table)
    walk, _, _ = half.trace(rays)
    brute = half.trace_brute(rays)
    assert np.array_equal(walk[:, 0], brute[:, 0])
    hit = walk[:, 0] != MISS
    assert hit.sum() > 500 and (table[walk[hit, 0]] == ANYHIT_ACCEPT).all()
    assert not np.array_equal(walk, ref)  # the table changed some results
