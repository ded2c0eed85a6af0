"""GPU parity of the refit after a vertex update (rdrt_refit).

* Refit to new positions, then traverse: every node layout equals the oracle
  built from scratch on the new positions except near-ties, with
  bit-identical t, u, v wherever the primitive ids agree (the triangle
  records are recomputed with the host pre-pass arithmetic).
* Refit away and back to the committed positions: the reference-order kernel
  equals the oracle on those positions bit for bit, hit records and per-ray
  visit counters (box unions are exact, so the refit boxes equal the built
  ones).
* Instanced scenes (TLAS rebuilt from the refit BLAS boxes) and the path
  tracer after a refit match the oracle on the new geometry.
"""

import numpy as np
import pytest

from oracle.oracle import OracleScene
from rendiation_amd import synth
from rendiation_amd.acceleration import TRAVERSAL_FAST, TRAVERSAL_REFERENCE, BuildOptions
from tests.parity import compare
from tests.scenes import build_pair

pytestmark = pytest.mark.gpu


def _say(msg):
    print(f"[gpu-case] {msg}", flush=True)


def _trace(sys, rays_np, device, traversal=TRAVERSAL_FAST):
    import torch

    rays = torch.from_numpy(np.ascontiguousarray(rays_np)).to(device)
    n = rays_np.shape[0]
    hits = torch.empty((n, 4), dtype=torch.int32, device=device)
    cnt = torch.empty((n, 3), dtype=torch.int32, device=device)
    sys.trace(rays, hits, counters=cnt, traversal=traversal)
    sys.sync()
    return hits.cpu().numpy().view(np.uint32), cnt.cpu().numpy().view(np.uint32)


def _field_rays(seed, n):
    """Rays from above the height field, downward with some spread."""
    rng = np.random.default_rng(seed)
    o = np.stack([rng.uniform(-0.8, 0.8, n), np.full(n, 1.5), rng.uniform(-0.8, 0.8, n)], axis=1)
    d = np.stack([rng.normal(0, 0.15, n), -np.ones(n), rng.normal(0, 0.15, n)], axis=1)
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    r = np.zeros((n, 8), dtype=np.float32)
    r[:, 0:3], r[:, 3], r[:, 4:7], r[:, 7] = o, 1e-4, d, np.float32(np.finfo(np.float32).max)
    return r


def _moved(pos, seed, scale=0.05):
    rng = np.random.default_rng(seed)
    p = pos.copy()
    p[:, 1] += (scale * rng.standard_normal(p.shape[0])).astype(np.float32)
    return p


def _to_dev(a, device):
    import torch

    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.float32).reshape(-1)).to(device)


def test_refit_matches_fresh_build(device):
    pos, idx, _ = synth.height_field(60, seed=3, amplitude=0.05)
    p1 = _moved(pos, 7)
    rays = _field_rays(1, 40_000)
    opts = BuildOptions()
    _say("refit 60x60 height field, three layouts vs the oracle on the new positions")
    sys, _ = build_pair([[(pos, idx, 1)]], [[{"transform": np.eye(4), "blas": 0}]], [0], opts)
    sys.refit(_to_dev(p1, device))
    orc = OracleScene(opts.blas_bin_size, opts.blas_max_depth, opts.tlas_bin_size, opts.tlas_max_depth, opts.sah_bins)
    orc.add_blas([(p1, idx, 1)])
    orc.add_tlas([{"transform": np.eye(4), "blas": 0}])
    orc.bind([0])
    orc.build()
    rh, _, _ = orc.trace(rays, with_counters=True)
    assert (rh[:, 0] != 0xFFFFFFFF).mean() > 0.8
    for wide in (0, 1, 2):
        sys.set_tune(wide=wide)
        gh, _ = _trace(sys, rays, device)
        d = compare(gh, rh)
        assert d.ok() and d.value_mismatch_bits == 0, (wide, d)
    sys.set_tune()
    gh, _ = _trace(sys, rays, device, TRAVERSAL_REFERENCE)
    d = compare(gh, rh)
    assert d.ok() and d.value_mismatch_bits == 0, d


def test_refit_round_trip_is_bit_exact(device):
    pos, idx, _ = synth.height_field(50, seed=4, amplitude=0.08)
    rays = _field_rays(2, 30_000)
    _say("refit away and back: reference-order kernel bit-exact")
    sys, orc = build_pair([[(pos, idx, 1)]], [[{"transform": np.eye(4), "blas": 0}]], [0], BuildOptions())
    rh, _, rc = orc.trace(rays, with_counters=True)
    sys.refit(_to_dev(_moved(pos, 9, 0.2), device))
    sys.refit(_to_dev(pos, device))
    gh, gc = _trace(sys, rays, device, TRAVERSAL_REFERENCE)
    assert np.array_equal(gh, rh)
    assert np.array_equal(gc, rc)
    for wide in (0, 1, 2):
        sys.set_tune(wide=wide)
        gh, _ = _trace(sys, rays, device)
        d = compare(gh, rh)
        assert d.ok() and d.value_mismatch_bits == 0, (wide, d)
    sys.set_tune()


def test_refit_instanced_scene(device):
    """Two BLAS, five instances: moving one BLAS's vertices moves its instance
    boxes; the rebuilt TLAS and the refit BLAS match a fresh oracle build."""
    a_pos, a_idx, _ = synth.height_field(30, seed=5, amplitude=0.05)
    b_pos = synth.random_triangles(6, 3000, spread=0.05)
    insts = [{"transform": np.eye(4), "blas": 0}]
    rng = np.random.default_rng(8)
    for k in range(4):
        t = np.eye(4)
        t[:3, 3] = rng.uniform(-1.0, 1.0, 3)
        t[1, 3] = -0.5 + 0.2 * k
        insts.append({"transform": t, "blas": 1 if k % 2 else 0})
    _say("refit instanced scene")
    sys, _ = build_pair([[(a_pos, a_idx, 1)], [(b_pos, None, 1)]], [insts], [0], BuildOptions())
    a1 = _moved(a_pos, 10, 0.3)
    b1 = b_pos.copy()
    b1[:, 1] += np.float32(0.25)
    sys.refit(_to_dev(np.concatenate([a1, b1]), device))
    opts = BuildOptions()
    orc = OracleScene(opts.blas_bin_size, opts.blas_max_depth, opts.tlas_bin_size, opts.tlas_max_depth, opts.sah_bins)
    orc.add_blas([(a1, a_idx, 1)])
    orc.add_blas([(b1, None, 1)])
    orc.add_tlas(insts)
    orc.bind([0])
    orc.build()
    rays = _field_rays(3, 30_000)
    rh, _, _ = orc.trace(rays, with_counters=True)
    for wide in (0, 1, 2):
        sys.set_tune(wide=wide)
        gh, _ = _trace(sys, rays, device)
        d = compare(gh, rh)
        assert d.ok() and d.value_mismatch_bits == 0, (wide, d)
    sys.set_tune()
    with pytest.raises(Exception):
        sys.refit(_to_dev(a1, device))  # wrong vertex count


def test_path_tracer_after_refit(device):
    from rendiation_amd.pathtracer import Camera, PathTracer, PathTracingConfig
    from tests.pt_util import oracle_params

    pos, idx, nrm = synth.height_field(40, seed=2, amplitude=0.06)
    p1 = _moved(pos, 12, 0.03)
    _say("path tracer after a refit vs the oracle on the new positions")
    sys, _ = build_pair([[(pos, idx, 1)]], [[{"transform": np.eye(4), "blas": 0}]], [0], BuildOptions())
    albedo, emission = np.full((1, 3), 0.6, dtype=np.float32), np.zeros((1, 3), dtype=np.float32)
    sys.set_materials(albedo, emission)
    cam = Camera(position=(0.0, 2.0, 2.5), target=(0.0, 0.0, 0.0), fov_y_deg=50.0, aspect=1.0)
    cfg = PathTracingConfig(width=64, height=64, spp=2, max_bounces=2, rr_start=99, seed=3)
    pt = PathTracer(sys)
    pt.render(cam, cfg)  # a frame before the refit (captured graph dropped by it)
    sys.refit(_to_dev(p1, device))
    img = pt.render(cam, cfg).cpu().numpy().astype(np.float64)
    opts = BuildOptions()
    orc = OracleScene(opts.blas_bin_size, opts.blas_max_depth, opts.tlas_bin_size, opts.tlas_max_depth, opts.sah_bins)
    orc.add_blas([(p1, idx, 1)])
    orc.add_tlas([{"transform": np.eye(4), "blas": 0}])
    orc.bind([0])
    orc.build()
    orc.set_materials(albedo, emission)
    ref, _, _ = orc.pt_render(oracle_params(cam, cfg))
    rel = np.sqrt(np.mean((img - ref) ** 2)) / ref.mean()
    assert rel <= 1e-2, rel
