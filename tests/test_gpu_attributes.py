"""GPU checks of the hit attributes (the bindless mesh fetch and the
shader-binding-table hit group), against facts that do not restate the
kernel's formulas:

* the position lies on the hit triangle in world space (transformed
  vertices) and equals origin + t * direction;
* the geometric normal is a unit vector orthogonal to the world-space
  triangle edges and faces the ray origin;
* with vertex normals equal to the analytic sphere normals, the shading
  normal approximates the normal of the transformed sphere at the hit point;
  without vertex normals it is the geometric normal;
* with uv = an affine function of the object-space position, the uv equals
  that function of the object-space hit point;
* the hit group is sbt_offset + sbt_stride * geometry + instance offset;
* 16-bit indices give the same hits and attributes as 32-bit ones.
"""

import numpy as np
import pytest

from rendiation_amd import synth
from rendiation_amd.acceleration import (
    AccelerationStructureSystem,
    BuildOptions,
    TopLevelAccelerationStructureSourceInstance,
    TriangleGeometry,
)

pytestmark = pytest.mark.gpu


def _say(msg):
    print(f"[gpu-case] {msg}", flush=True)


def _transforms():
    rng = np.random.default_rng(3)
    out = []
    for k, (scale, mirror) in enumerate([(1.0, False), (1.7, False), (0.6, True)]):
        q, _ = np.linalg.qr(rng.standard_normal((3, 3)))
        m = np.eye(4)
        m[:3, :3] = q * scale
        if mirror:
            m[:3, 0] *= -1.0
        m[:3, 3] = [-3.0 + 3.0 * k, 0.2 * k, -0.5 * k]
        out.append(m)
    return out


def _scene(device, index_dtype=np.uint32):
    pos, idx = synth.uv_sphere(48, 24)
    nrm = synth.sphere_normals(48, 24)
    uv = np.stack([0.25 * pos[:, 0] + 0.5 * pos[:, 2] + 0.1, -0.3 * pos[:, 1] + 0.2 * pos[:, 0]], axis=1)
    plane = np.array([[-1, 0, -1], [1, 0, -1], [1, 0, 1], [-1, 0, 1]], dtype=np.float32)
    pidx = np.array([[0, 2, 1], [0, 3, 2]])
    s = AccelerationStructureSystem(device=0, options=BuildOptions())
    b0 = s.create_bottom_level_acceleration_structure([TriangleGeometry(pos, idx.astype(index_dtype))])
    b1 = s.create_bottom_level_acceleration_structure([TriangleGeometry(plane, pidx.astype(index_dtype)),
                                                       TriangleGeometry(pos * np.float32(0.3), idx.astype(index_dtype))])
    s.set_vertex_normals(b0, 0, nrm)
    s.set_vertex_uvs(b0, 0, uv.astype(np.float32))
    ts = _transforms()
    insts = [TopLevelAccelerationStructureSourceInstance(ts[0], b0, instance_shader_binding_table_record_offset=5),
             TopLevelAccelerationStructureSourceInstance(ts[1], b1, instance_shader_binding_table_record_offset=11),
             TopLevelAccelerationStructureSourceInstance(ts[2], b0, instance_shader_binding_table_record_offset=17)]
    t = s.create_top_level_acceleration_structure(insts)
    s.bind_tlas([t])
    s.create_comp_instance()
    return s, ts, (pos, idx, nrm, uv), (plane, pidx)


def _rays(n, seed=5):
    rng = np.random.default_rng(seed)
    o = rng.uniform(-6, 6, size=(n, 3))
    target = rng.uniform(-4, 4, size=(n, 3)) * np.array([1.0, 0.4, 0.6])
    d = target - o
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    r = np.zeros((n, 8), dtype=np.float32)
    r[:, 0:3], r[:, 3], r[:, 4:7], r[:, 7] = o, 1e-4, d, np.finfo(np.float32).max
    return r


def _trace(s, rays_np, device, sbt=(3, 2)):
    import torch

    n = rays_np.shape[0]
    rays = torch.from_numpy(rays_np).to(device)
    hits = torch.empty((n, 4), dtype=torch.int32, device=device)
    hm = torch.empty((n, 4), dtype=torch.int32, device=device)
    s.trace(rays, hits, hit_meta=hm)
    out = torch.empty((n, 12), dtype=torch.int32, device=device)
    s.hit_attributes(rays, hits, hm, out, sbt_offset=sbt[0], sbt_stride=sbt[1])
    s.sync()
    return (hits.cpu().numpy().view(np.uint32), hm.cpu().numpy().view(np.uint32), out.cpu().numpy().view(np.uint32))


def test_hit_attributes(device):
    _say("hit attributes on an instanced scene with normals and uvs")
    s, ts, (pos, idx, nrm, uv), (plane, pidx) = _scene(device)
    rays = _rays(50_000)
    h, hm, out = _trace(s, rays, device)
    a = out.view(np.float32)
    hit = h[:, 0] != 0xFFFFFFFF
    assert hit.mean() > 0.2
    assert np.all(out[~hit, 11] == 0xFFFFFFFF) and np.all(a[~hit, :11] == 0.0)
    rows = np.nonzero(hit)[0]
    t = h[rows, 1].view(np.float32).astype(np.float64)
    o, d = rays[rows, 0:3].astype(np.float64), rays[rows, 4:7].astype(np.float64)
    p = a[rows, 0:3].astype(np.float64)
    assert np.allclose(p, o + d * t[:, None], atol=1e-5)
    gn, sn = a[rows, 6:9].astype(np.float64), a[rows, 3:6].astype(np.float64)
    assert np.allclose(np.linalg.norm(gn, axis=1), 1.0, atol=1e-5)
    assert np.allclose(np.linalg.norm(sn, axis=1), 1.0, atol=1e-5)
    assert np.all(np.sum((o - p) * gn, axis=1) >= -1e-6)  # faces the ray origin
    assert np.all(np.sum(gn * sn, axis=1) >= -1e-6)       # shading on the geometric side
    sphere = (pos, idx)
    small = (pos * np.float32(0.3), idx)
    for j, r in enumerate(rows):
        inst, geo, prim = int(hm[r, 1]), int(hm[r, 0]), int(h[r, 0])
        m = ts[inst]
        if inst == 1:
            vp, vi = (plane, pidx) if geo == 0 else small
        else:
            vp, vi = sphere
        tri = vp[vi[prim]].astype(np.float64)
        w = np.c_[tri, np.ones(3)] @ m.T
        w = w[:, :3]
        # on the triangle: in its plane and inside (barycentric solve)
        e1, e2 = w[1] - w[0], w[2] - w[0]
        nrm_w = np.cross(e1, e2)
        nrm_w /= np.linalg.norm(nrm_w)
        assert abs(np.dot(p[j] - w[0], nrm_w)) < 1e-4 * (1 + np.abs(p[j]).max())
        assert abs(abs(np.dot(gn[j], nrm_w)) - 1.0) < 1e-5
        assert out[r, 11] == 3 + 2 * geo + [5, 11, 17][inst]
        if inst != 1:
            # object-space hit point
            q = np.linalg.solve(m, np.r_[p[j], 1.0])[:3]
            f = np.array([0.25 * q[0] + 0.5 * q[2] + 0.1, -0.3 * q[1] + 0.2 * q[0]])
            assert np.allclose(a[r, 9:11], f, atol=2e-5)
            # analytic sphere normal of the transformed sphere at the hit point
            ns = np.linalg.inv(m[:3, :3]).T @ (q / np.linalg.norm(q))
            ns /= np.linalg.norm(ns)
            if np.dot(ns, gn[j]) < 0:
                ns = -ns
            assert np.dot(ns, sn[j]) > 0.995
        else:
            assert np.allclose(sn[j], gn[j], atol=1e-6)  # no vertex normals: the geometric normal
            assert np.all(a[r, 9:11] == 0.0)


def test_16_bit_indices(device):
    _say("16-bit index format")
    s32, _, _, _ = _scene(device, np.uint32)
    s16, _, _, _ = _scene(device, np.uint16)
    rays = _rays(20_000, seed=6)
    h32, m32, o32 = _trace(s32, rays, device)
    h16, m16, o16 = _trace(s16, rays, device)
    assert np.array_equal(h32, h16) and np.array_equal(m32, m16) and np.array_equal(o32, o16)


def test_attributes_and_hit_context_match_the_oracle(device):
    """The device fetch against the oracle's twin on the same (GPU) hits:
    same FP32 operations in the same order, so the same bits -- position,
    normals, uv, hit group, both transforms and the object-space ray."""
    import torch

    from tests.test_hit_attributes_oracle import oracle_scene

    _say("hit attributes and hit context vs the oracle twin")
    s, ts, _, _ = _scene(device)
    rays_np = _rays(50_000, seed=9)
    n = rays_np.shape[0]
    rays = torch.from_numpy(rays_np).to(device)
    hits = torch.empty((n, 4), dtype=torch.int32, device=device)
    hm = torch.empty((n, 4), dtype=torch.int32, device=device)
    s.trace(rays, hits, hit_meta=hm)
    attr = torch.empty((n, 12), dtype=torch.int32, device=device)
    xf = torch.empty((n, 40), dtype=torch.int32, device=device)
    s.hit_attributes(rays, hits, hm, attr, sbt_offset=3, sbt_stride=2)
    s.hit_transforms(rays, hits, hm, xf)
    s.sync()
    h, m = hits.cpu().numpy().view(np.uint32), hm.cpu().numpy().view(np.uint32)
    o, _ = oracle_scene()
    want_a, want_x = o.hit_attributes(rays_np, h, m, sbt_offset=3, sbt_stride=2)
    got_a, got_x = attr.cpu().numpy().view(np.uint32), xf.cpu().numpy().view(np.uint32)
    assert (h[:, 0] != 0xFFFFFFFF).mean() > 0.2
    assert np.array_equal(got_a, want_a), int(np.sum(np.any(got_a != want_a, axis=1)))
    assert np.array_equal(got_x, want_x), int(np.sum(np.any(got_x != want_x, axis=1)))
    s.close()
