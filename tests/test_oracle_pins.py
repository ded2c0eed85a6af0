"""Pins of the CPU oracle itself, by means that do not restate its formulas:
closed forms, hand-computed SAH splits, structural invariants, brute force,
an independent FP64 solver, the Sobol stratification property and the
reference's own compaction test vector."""

import numpy as np
import pytest

from oracle import oracle as O
from rendiation_amd import synth
from tests.parity import compare

MISS = 0xFFFFFFFF
INV = 0xFFFFFFFF


# ---------------------------------------------------------------- triangle --
def test_triangle_closed_form():
    # unit right triangle in z = 0, ray straight down: t = 1, (u, v) = (x, y)
    r = O.intersect_triangle([0.25, 0.125, 1.0], [0, 0, -1], 0.0, 10.0, [0, 0, 0], [1, 0, 0], [0, 1, 0])
    assert list(r) == [-1.0, 1.0, 0.25, 0.125]  # b < 0: back-facing sign; u weights v1, v weights v2
    r = O.intersect_triangle([0.5, 0.25, -2.0], [0, 0, 1], 0.0, 10.0, [0, 0, 0], [1, 0, 0], [0, 1, 0])
    assert list(r) == [1.0, 2.0, 0.5, 0.25]
    # outside the triangle (u + v > 1), behind the ray, beyond far
    assert not O.intersect_triangle([0.75, 0.5, 1], [0, 0, -1], 0, 10, [0, 0, 0], [1, 0, 0], [0, 1, 0]).any()
    assert not O.intersect_triangle([0.25, 0.25, 1], [0, 0, 1], 0, 10, [0, 0, 0], [1, 0, 0], [0, 1, 0]).any()
    assert not O.intersect_triangle([0.25, 0.25, 1], [0, 0, -1], 0, 0.5, [0, 0, 0], [1, 0, 0], [0, 1, 0]).any()
    # t == far is accepted, t == near is accepted (range test is t < near || t > far)
    assert O.intersect_triangle([0.25, 0.25, 1], [0, 0, -1], 0, 1.0, [0, 0, 0], [1, 0, 0], [0, 1, 0])[1] == 1.0
    assert O.intersect_triangle([0.25, 0.25, 1], [0, 0, -1], 1.0, 2.0, [0, 0, 0], [1, 0, 0], [0, 1, 0])[1] == 1.0
    # ray parallel to the plane (b == 0) and a zero-area triangle are misses
    assert not O.intersect_triangle([0, 0, 0], [1, 0, 0], 0, np.inf, [0, 0, 0], [1, 0, 0], [0, 1, 0]).any()
    assert not O.intersect_triangle([0, 0, 1], [0, 0, -1], 0, np.inf, [0, 0, 0], [1, 0, 0], [1, 0, 0]).any()


def test_triangle_culling():
    args = ([0.25, 0.25, 1.0], [0, 0, -1], 0.0, 10.0, [0, 0, 0], [1, 0, 0], [0, 1, 0])  # b = -1
    assert O.intersect_triangle(*args, cull_enable=True, cull_back=False)[0] == -1.0
    assert not O.intersect_triangle(*args, cull_enable=True, cull_back=True).any()
    args2 = ([0.25, 0.25, -1.0], [0, 0, 1], 0.0, 10.0, [0, 0, 0], [1, 0, 0], [0, 1, 0])  # b = +1
    assert not O.intersect_triangle(*args2, cull_enable=True, cull_back=False).any()
    assert O.intersect_triangle(*args2, cull_enable=True, cull_back=True)[0] == 1.0


def test_triangle_against_fp64_solver():
    rng = np.random.default_rng(0)
    checked = 0
    for _ in range(4000):
        v = rng.uniform(-1, 1, (3, 3)).astype(np.float32)
        o = rng.uniform(-2, 2, 3).astype(np.float32)
        target = (v[0] + rng.uniform(0, 1) * (v[1] - v[0]) * 0.5 + rng.uniform(0, 1) * (v[2] - v[0]) * 0.5)
        d = (target - o).astype(np.float64)
        d = (d / np.linalg.norm(d)).astype(np.float32)
        r = O.intersect_triangle(o, d, 0.0, np.inf, v[0], v[1], v[2])
        # independent solve: o + t d = v0 + u e1 + v e2
        A = np.stack([d.astype(np.float64), -(v[1] - v[0]).astype(np.float64), -(v[2] - v[0]).astype(np.float64)], 1)
        if abs(np.linalg.det(A)) < 1e-2:
            continue
        t, u, w = np.linalg.solve(A, (v[0] - o).astype(np.float64))
        if not (t > 1e-3 and u > 1e-3 and w > 1e-3 and u + w < 1 - 1e-3):
            continue
        assert r[0] != 0.0
        assert abs(r[1] - t) <= 1e-4 * abs(t)
        # FP32 dot-product barycentrics lose a few digits to cancellation
        assert abs(r[2] - u) <= 1e-3 and abs(r[3] - w) <= 1e-3
        checked += 1
    assert checked > 1000


# -------------------------------------------------------------------- aabb --
def test_aabb_closed_form():
    b0, b1 = [0, 0, 0], [1, 1, 1]
    assert O.intersect_aabb([-1, 0.5, 0.5], [1, 0, 0], 0, 10, b0, b1)  # enters at 1, leaves at 2
    assert not O.intersect_aabb([-1, 0.5, 0.5], [1, 0, 0], 0, 1.0, b0, b1)  # t_near < t_max fails at t_max = 1
    assert O.intersect_aabb([-1, 0.5, 0.5], [1, 0, 0], 0, 1.0001, b0, b1)
    assert not O.intersect_aabb([-1, 0.5, 0.5], [1, 0, 0], 2.0, 10, b0, b1)  # t_min < t_far fails at 2
    assert not O.intersect_aabb([-1, 2.0, 0.5], [1, 0, 0], 0, 10, b0, b1)
    assert not O.intersect_aabb([-1, 0.5, 0.5], [-1, 0, 0], 0, 10, b0, b1)
    assert O.intersect_aabb([0.5, 0.5, 0.5], [0, 0, 1], 0, 10, b0, b1)  # origin inside, axis-parallel (inf slopes)


# ---------------------------------------------------------------- BVH build --
def _scene_one(pos, idx=None, **opts):
    s = O.OracleScene(**opts)
    b = s.add_blas([(pos, idx, 1)])
    s.bind([s.add_tlas([{"transform": np.eye(4), "blas": b}])])
    s.build()
    return s


def test_sah_split_by_hand():
    # 8 unit boxes along x: the 4-bin SAH picks the middle split at every
    # level (costs 176 / 144 / 176 at the root, 48 / 40 / 48 below).
    tris = []
    for i in range(8):
        tris += [[i, 0, 0], [i + 1, 1, 0], [i, 1, 1]]
    s = _scene_one(np.array(tris, np.float32), blas_bin_size=2)
    f = s.tri_forest()
    ranges = [tuple(r) for r in f[:, 8:10]]
    assert ranges == [(0, 8), (0, 4), (0, 2), (2, 4), (4, 8), (4, 6), (6, 8)]
    links = [tuple(r) for r in f[:, [3, 7]]]
    assert links == [(1, INV), (2, 4), (3, 3), (4, 4), (5, INV), (6, 6), (INV, INV)]
    assert list(s.redirect()) == list(range(8))


def test_bvh_invariants():
    pos = synth.random_triangles(3, 5000)
    for leaf in (1, 2, 4, 7):
        s = _scene_one(pos, blas_bin_size=leaf)
        f = s.tri_forest()
        ff = f.view(np.float32)
        red = s.redirect()
        assert np.array_equal(np.sort(red), np.arange(5000))  # a permutation
        tri = pos.reshape(-1, 3, 3)
        tmin = tri.min(axis=1)
        tmax = tri.max(axis=1)
        n = f.shape[0]
        size = np.ones(n, dtype=np.int64)
        for i in range(n - 1, -1, -1):
            b, e = f[i, 8], f[i, 9]
            hit, miss = f[i, 3], f[i, 7]
            leafnode = hit == miss
            # node box = union of its primitive boxes
            assert np.array_equal(ff[i, 0:3], tmin[red[b:e]].min(axis=0))
            assert np.array_equal(ff[i, 4:7], tmax[red[b:e]].max(axis=0))
            if leafnode:
                assert e - b <= leaf
            else:
                assert hit == i + 1
                l = i + 1
                r = l + size[l]
                assert f[l, 8] == b and f[l, 9] == f[r, 8] and f[r, 9] == e  # children partition the range
                size[i] = 1 + size[l] + size[r]
            nxt = i + size[i]
            assert miss == (nxt if nxt < n else INV)
        assert size[0] == n and f[0, 8] == 0 and f[0, 9] == 5000


# --------------------------------------------------------------- traversal --
def test_traversal_equals_brute_force():
    pos = synth.random_triangles(4, 2000, spread=0.05)
    rays = synth.random_rays(4, 20_000)
    s = _scene_one(pos, blas_bin_size=4)
    bvh, _, _ = s.trace(rays)
    brute = s.trace_brute(rays)
    d = compare(bvh, brute)
    assert (bvh[:, 0] != MISS).mean() > 0.1
    assert d.ok() and d.value_mismatch_bits == 0, d


def test_traversal_with_instances_equals_brute_force():
    from tests.scenes import instanced_scene

    blas_list, tlas_list, binding, rays, _ = instanced_scene(seed=8, nrays=20_000)
    s = O.OracleScene()
    for g in blas_list:
        s.add_blas(g)
    for t in tlas_list:
        s.add_tlas(t)
    s.bind(binding)
    s.build()
    for tl in (0, 1):
        bvh, _, _ = s.trace(rays, tlas_idx=tl)
        brute = s.trace_brute(rays, tlas_idx=tl)
        d = compare(bvh, brute)
        assert (bvh[:, 0] != MISS).mean() > 0.05
        assert d.ok(), d


def test_hits_against_fp64():
    pos = synth.random_triangles(0, 20_000)
    rays = synth.random_rays(0, 50_000)
    s = _scene_one(pos, blas_bin_size=4)
    h, _, _ = s.trace(rays)
    hit = h[:, 0] != MISS
    hf = h.view(np.float32)
    tri = pos.reshape(-1, 3, 3).astype(np.float64)[h[hit, 0]]
    o = rays[hit, 0:3].astype(np.float64)
    d = rays[hit, 4:7].astype(np.float64)
    # world t of the FP32 result vs the FP64 plane intersection of the same triangle
    n = np.cross(tri[:, 1] - tri[:, 0], tri[:, 2] - tri[:, 0])
    t64 = np.einsum("ij,ij->i", n, tri[:, 0] - o) / np.einsum("ij,ij->i", n, d)
    p = o + t64[:, None] * d
    e1, e2, w = tri[:, 1] - tri[:, 0], tri[:, 2] - tri[:, 0], p - tri[:, 0]
    uu, uv, vv = (e1 * e1).sum(1), (e1 * e2).sum(1), (e2 * e2).sum(1)
    wu, wv = (w * e1).sum(1), (w * e2).sum(1)
    den = uv * uv - uu * vv
    u64 = (uv * wv - vv * wu) / den
    v64 = (uv * wu - uu * wv) / den
    rel_t = np.abs(hf[hit, 1] - t64) / t64
    assert np.median(rel_t) < 1e-6 and np.quantile(rel_t, 0.999) < 1e-4
    assert np.quantile(np.abs(hf[hit, 2] - u64), 0.999) < 1e-3
    assert np.quantile(np.abs(hf[hit, 3] - v64), 0.999) < 1e-3


# ------------------------------------------------------------------ Sobol ---
def test_sobol_known_answers():
    assert [O.van_der_corput(n, 0) for n in range(8)] == [0, .5, .25, .75, .125, .625, .375, .875]
    assert [O.sobol(n, 0) for n in range(8)] == [0, .5, .75, .25, .625, .125, .375, .875]
    # clamp to 1 - 2^-23 when all 24 output bits are set
    assert O.van_der_corput(0, 0xFFFFFFFF) == np.float32(1.0 - 2.0 ** -23)
    assert O.sobol(0, 0xFFFFFFFF) == np.float32(1.0 - 2.0 ** -23)


def test_sobol_scramble_is_xor():
    rng = np.random.default_rng(1)
    for _ in range(200):
        n, s = int(rng.integers(0, 2 ** 32)), int(rng.integers(0, 2 ** 32))
        rev = int(format(n, "032b")[::-1], 2)
        want = min(((rev ^ s) >> 8) / 2 ** 24, 1 - 2 ** -23)
        assert O.van_der_corput(n, s) == np.float32(want)
        base = O.sobol(n, 0)
        bits = int(base * 2 ** 24) << 8
        assert O.sobol(n, s) == np.float32(min((((bits ^ s) >> 8) & 0xFFFFFF) / 2 ** 24, 1 - 2 ** -23))


@pytest.mark.parametrize("m", [4, 8, 10])
def test_sobol_is_stratified(m):
    # (0, m, 1)-net: the first 2^m points fall one per cell of width 2^-m
    for fn in (O.van_der_corput, O.sobol):
        cells = sorted(int(fn(n, 0) * 2 ** m) for n in range(2 ** m))
        assert cells == list(range(2 ** m))
    # jointly (dims 0 and 1): a (0, m, 2)-net in base 2 for elementary intervals 2^-a x 2^-(m-a)
    pts = [(O.van_der_corput(n, 0), O.sobol(n, 0)) for n in range(2 ** m)]
    for a in range(m + 1):
        cells = sorted(int(x * 2 ** a) * 2 ** (m - a) + int(y * 2 ** (m - a)) for x, y in pts)
        assert cells == list(range(2 ** m))


# ------------------------------------------------------------- compaction --
def test_compaction_reference_vector():
    # shader/parallel-compute/src/stream_compaction.rs:101-113
    vals = [1, 0, 1, 0, 1, 1, 0]
    out, k = O.compact([v == 1 for v in vals], vals)
    assert k == 4 and list(out) == [1, 1, 1, 1, 0, 0, 0]


def test_compaction_is_stable_copy_if():
    rng = np.random.default_rng(2)
    for n in (1, 2, 63, 64, 65, 1000, 4097):
        mask = rng.integers(0, 2, n).astype(np.uint32)
        out, k = O.compact(mask)
        assert k == int(mask.sum())
        assert np.array_equal(out[:k], np.nonzero(mask)[0])


# ------------------------------------------------------------ batch spawn --
def test_batch_round_spawn_closed_forms():
    """Continuation rays of the oracle's batch round, on a plane x + y = 0
    instanced with scale (2, 1, 1) (world plane x/2 + y = 0, normal along
    (1/2, 1, 0) by the inverse transpose): hit point on the plane, origin
    offset to the side the ray came from, unit directions in that
    hemisphere, cos ~ the cosine-weighted law (E cos = 2/3, cos^2 ~ U[0,1])."""
    from tests.scenes import affine

    big = np.array([[-50, 50, -50], [50, -50, -50], [0, 0, 80]], dtype=np.float32)  # in the plane x + y = 0
    s = O.OracleScene()
    s.add_blas([(big, None, 1)])
    s.add_tlas([{"transform": affine(None, (2.0, 1.0, 1.0)), "blas": 0}])
    s.bind([0])
    s.build()
    n = 40_000
    nw = np.array([0.5, 1.0, 0.0]) / np.linalg.norm([0.5, 1.0, 0.0])
    side = np.where(np.arange(n) % 2 == 0, 1.0, -1.0)
    tang = np.array([2.0, -1.0, 0.0]) / np.sqrt(5.0)
    u = synth.uniform_f32(3, 40, 2 * n).reshape(n, 2).astype(np.float64)
    o = side[:, None] * 3.0 * nw + (u[:, :1] - 0.5) * 4.0 * tang + (u[:, 1:] - 0.5) * 4.0 * np.array([0, 0, 1.0])
    rays = np.zeros((n, 8), dtype=np.float32)
    rays[:, 0:3] = o
    rays[:, 3] = 1e-4
    rays[:, 4:7] = -side[:, None] * nw
    rays[:, 7] = np.inf
    meta = np.zeros((n, 4), dtype=np.uint32)
    meta[:, 1] = 0xFFFFFFFF
    hits, sp, alive = s.batch_round(rays, meta, 1, 7)
    assert alive.all() and np.all(hits[:, 0] != MISS)
    t = hits[:, 1].view(np.float32).astype(np.float64)
    assert np.allclose(t, 3.0, rtol=1e-5)
    facing = side[:, None] * nw  # normal on the side of the ray origin
    pos = rays[:, 0:3].astype(np.float64) + rays[:, 4:7].astype(np.float64) * t[:, None]
    off = sp[:, 0:3].astype(np.float64) - pos
    assert np.all(np.sum(off * facing, axis=1) > 0) and np.all(np.abs(off) < 1e-3)
    d = sp[:, 4:7].astype(np.float64)
    assert np.allclose(np.linalg.norm(d, axis=1), 1.0, atol=1e-6)
    c = np.sum(d * facing, axis=1)
    assert np.all(c >= 0)
    assert abs(c.mean() - 2.0 / 3.0) < 5 * np.sqrt(1.0 / 18.0 / n)
    # cos^2 = 1 - r^2 with r^2 uniform on [0, 1]: Kolmogorov-Smirnov distance
    c2 = np.sort(c * c)
    ks = np.max(np.abs(c2 - (np.arange(1, n + 1) - 0.5) / n))
    assert ks < 1.63 / np.sqrt(n)
    assert np.all(sp[:, 3] == np.float32(np.finfo(np.float32).eps)) and np.all(sp[:, 7] == np.finfo(np.float32).max)
    # the sample depends on (ray, round, seed) only
    _, sp2, _ = s.batch_round(rays, meta, 1, 7)
    _, sp3, _ = s.batch_round(rays, meta, 2, 7)
    assert np.array_equal(sp, sp2) and not np.array_equal(sp[:, 4:7], sp3[:, 4:7])


def test_batch_round_per_ray_scene_selection():
    """tlas_idx picks the scene per ray: the same rays against two bound
    scenes equal plain traces of each scene; misses spawn nothing."""
    a = synth.random_triangles(21, 3000, spread=0.05)
    b = synth.random_triangles(22, 1500, spread=0.05)
    s = O.OracleScene(4)
    s.add_blas([(a, None, 1)])
    s.add_blas([(b, None, 1)])
    s.add_tlas([{"transform": np.eye(4), "blas": 0}])
    s.add_tlas([{"transform": np.eye(4), "blas": 1}])
    s.bind([0, 1])
    s.build()
    rays = synth.random_rays(23, 10_000)
    meta = np.zeros((10_000, 4), dtype=np.uint32)
    meta[:, 1] = 0xFFFFFFFF
    meta[5000:, 2] = 1
    hits, _, alive = s.batch_round(rays, meta, 1, 0)
    h0, _, _ = s.trace(rays[:5000], tlas_idx=0)
    h1, _, _ = s.trace(rays[5000:], tlas_idx=1)
    assert np.array_equal(hits, np.concatenate([h0, h1]))
    assert np.array_equal(alive, (hits[:, 0] != MISS).astype(np.uint32))
    assert 0.05 < alive.mean() < 0.95
