"""GPU adjoint of the path tracer w.r.t. material albedo and emission.

* Against the oracle's FP64 gradient of the same frame (same sample streams,
  FP32 reference visibility): relative error (2-norm) <= 1e-3.
* Against central finite differences (h = 1e-3) of the GPU's own frame.
* Emission is linear: F(e + d) - F(e) = g_e . d.
* The autograd op and the L2 helper agree with plain torch arithmetic.
"""

import dataclasses

import numpy as np
import pytest

from rendiation_amd import synth
from rendiation_amd._lib import RdrtError
from rendiation_amd.acceleration import AccelerationStructureSystem, BlasHandle
from rendiation_amd.pathtracer import Camera, PathTracer, PathTracingConfig, load_scene, render_materials
from tests.pt_util import oracle_params, oracle_scene_from

pytestmark = pytest.mark.gpu

GRAD_REL = 1e-3
H = 1e-3


def _say(msg):
    print(f"[gpu-case] {msg}", flush=True)


@pytest.fixture(scope="module")
def grad_case(device):
    import torch

    scene = synth.pt_scene(seed=4, n_meshes=16, tess=12, extent=5.0, materials_per_mesh=4)
    system = AccelerationStructureSystem(device=0, stream=torch.cuda.current_stream().cuda_stream)
    load_scene(system, scene)
    cam = Camera(position=(0.0, 0.0, 18.0), aspect=1.5)
    cfg = PathTracingConfig(width=96, height=64, spp=2, max_bounces=3, rr_start=3, seed=5, record=True)
    rng = np.random.default_rng(0)
    adj = rng.uniform(-1.0, 1.0, size=(cfg.height, cfg.width, 3))
    albedo = np.asarray(scene["albedo"], np.float32)
    emission = np.asarray(scene["emission"], np.float32)
    return scene, system, PathTracer(system), cam, cfg, adj, albedo, emission


def _gpu_grad(tracer, cam, cfg, adj, albedo, emission):
    import torch

    a = torch.from_numpy(albedo).cuda()
    e = torch.from_numpy(emission).cuda()
    tracer.set_material_params(a, e)
    img = tracer.render(cam, cfg)
    ga = torch.empty_like(a)
    ge = torch.empty_like(e)
    tracer.backward(torch.from_numpy(adj.astype(np.float32)).cuda(), ga, ge)
    torch.cuda.synchronize()
    return img.cpu().numpy().astype(np.float64), ga.cpu().numpy().astype(np.float64), ge.cpu().numpy().astype(
        np.float64)


def _gpu_f(tracer, cam, cfg, adj, albedo, emission):
    import torch

    tracer.set_material_params(torch.from_numpy(np.ascontiguousarray(albedo)).cuda(),
                               torch.from_numpy(np.ascontiguousarray(emission)).cuda())
    img = tracer.render(cam, dataclasses.replace(cfg, record=False))
    torch.cuda.synchronize()
    return float(np.sum(adj * img.cpu().numpy().astype(np.float64)))


def _rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30))


def test_gradient_matches_fp64_oracle(grad_case):
    scene, _, tracer, cam, cfg, adj, albedo, emission = grad_case
    _say("adjoint vs FP64 oracle gradient, 96x64x2spp, 3 bounces, 64 materials")
    _, ga, ge = _gpu_grad(tracer, cam, cfg, adj, albedo, emission)
    o = oracle_scene_from(scene)
    oa, oe = o.pt_grad(oracle_params(cam, cfg), adj, albedo.shape[0])
    ra, re = _rel(ga, oa), _rel(ge, oe)
    print(f"rel err albedo {ra:.3e} emission {re:.3e}")
    assert np.count_nonzero(oa) > albedo.size // 2
    assert ra <= GRAD_REL and re <= GRAD_REL, (ra, re)


def test_gradient_against_central_differences(grad_case):
    _, _, tracer, cam, cfg, adj, albedo, emission = grad_case
    _say("adjoint vs central differences of the GPU frame")
    _, ga, ge = _gpu_grad(tracer, cam, cfg, adj, albedo, emission)
    scale = np.abs(ga).max()
    for idx in np.argsort(-np.abs(ga).ravel())[:4]:
        m, c = divmod(int(idx), 3)
        ap, am = albedo.copy(), albedo.copy()
        ap[m, c] = np.float32(albedo[m, c] + H)
        am[m, c] = np.float32(albedo[m, c] - H)
        fd = (_gpu_f(tracer, cam, cfg, adj, ap, emission) - _gpu_f(tracer, cam, cfg, adj, am, emission)) / (
            float(ap[m, c]) - float(am[m, c]))
        assert abs(fd - ga[m, c]) <= GRAD_REL * scale, (m, c, fd, ga[m, c])
    rng = np.random.default_rng(3)
    d = rng.uniform(0.0, 1.0, size=emission.shape).astype(np.float32)
    e1 = (emission + d).astype(np.float32)
    lhs = _gpu_f(tracer, cam, cfg, adj, albedo, e1) - _gpu_f(tracer, cam, cfg, adj, albedo, emission)
    rhs = float(np.sum(ge * (e1.astype(np.float64) - emission)))
    assert abs(lhs - rhs) <= 1e-4 * abs(rhs), (lhs, rhs)


def test_autograd_op_and_l2_helper(grad_case):
    import torch

    _, _, tracer, cam, cfg, _, albedo, emission = grad_case
    _say("autograd op, L2 helper")
    a0 = torch.from_numpy(albedo).cuda()
    e0 = torch.from_numpy(emission).cuda()
    with torch.no_grad():
        target = render_materials(tracer, cam, cfg, (a0 * 1.05).clamp(0, 1), e0).clone()
    a = a0.clone().requires_grad_(True)
    e = e0.clone().requires_grad_(True)
    img = render_materials(tracer, cam, cfg, a, e)
    loss = ((img - target) ** 2).sum()
    loss.backward()
    adj = torch.empty_like(img)
    loss_dev = torch.zeros(1, dtype=torch.float64, device="cuda")
    tracer.l2_adjoint(img.detach(), target, adj, loss_dev)
    ga = torch.empty_like(a0)
    ge = torch.empty_like(e0)
    tracer.backward(adj, ga, ge)
    torch.cuda.synchronize()
    assert abs(loss_dev.item() - loss.item()) <= 1e-4 * loss.item()
    assert torch.allclose(adj, 2 * (img.detach() - target), rtol=1e-6, atol=1e-7)
    assert torch.allclose(a.grad, ga, rtol=1e-4, atol=1e-6 * float(ga.abs().max()))
    assert torch.allclose(e.grad, ge, rtol=1e-4, atol=1e-6 * float(ge.abs().max()))
    assert float(a.grad.abs().sum()) > 0


def test_backward_preconditions(grad_case):
    import torch

    _, system, tracer, cam, cfg, _, albedo, emission = grad_case
    _say("adjoint preconditions")
    a = torch.from_numpy(albedo).cuda()
    e = torch.from_numpy(emission).cuda()
    tracer.set_material_params(a, e)
    g = torch.empty_like(a)
    adj = torch.zeros((cfg.height, cfg.width, 3), dtype=torch.float32, device="cuda")
    tracer.render(cam, dataclasses.replace(cfg, rr_start=1))
    with pytest.raises(RdrtError):
        tracer.backward(adj, g, torch.empty_like(e))
    tracer.render(cam, dataclasses.replace(cfg, record=False))
    with pytest.raises(RdrtError):
        tracer.backward(adj, g, torch.empty_like(e))
    system.sync()


# --------------------------------------------------------------- geometry --
@pytest.fixture(scope="module")
def geo_case(device):
    """48 x 48 height field with vertex normals under a sphere without
    normals; two views in one frame, 48 x 48 at 2 spp, 3 bounces."""
    import torch

    hp, hi, hn = synth.height_field(48, seed=1)
    sp, si = synth.uv_sphere(12, 8, 0.3)
    sp = (sp + np.array([0.2, 0.45, -0.1], np.float32)).astype(np.float32)
    scene = {"meshes": [{"positions": hp, "indices": hi, "normals": hn, "tri_material": None},
                        {"positions": sp, "indices": si, "normals": None, "tri_material": None}],
             "instances": [{"transform": np.eye(4), "mesh": 0, "material": 0},
                           {"transform": np.eye(4), "mesh": 1, "material": 1}],
             "albedo": np.array([[0.6, 0.5, 0.4], [0.7, 0.7, 0.7]], np.float32),
             "emission": np.zeros((2, 3), np.float32)}
    system = AccelerationStructureSystem(device=0, stream=torch.cuda.current_stream().cuda_stream)
    load_scene(system, scene)
    cams = [Camera(position=p, aspect=1.0) for p in synth.sphere_views(2, radius=2.5)]
    cfg = PathTracingConfig(width=48, height=48, spp=2, max_bounces=3, rr_start=3, seed=2, record=True)
    rng = np.random.default_rng(5)
    adj = rng.uniform(-1.0, 1.0, size=(2, cfg.height, cfg.width, 3))
    return scene, system, PathTracer(system), cams, cfg, adj


def test_geometry_gradient_matches_fp64_oracle(geo_case):
    import torch

    scene, system, tracer, cams, cfg, adj = geo_case
    _say("geometry adjoint vs FP64 oracle: height field + sphere, 2 views x 48x48 x 2 spp")
    nv = system.scene_info().verts
    tracer.render(cams, cfg)
    g_pos = torch.empty((nv, 3), dtype=torch.float32, device="cuda")
    g_nrm = torch.empty((nv, 3), dtype=torch.float32, device="cuda")
    tracer.backward_geometry(torch.from_numpy(adj.astype(np.float32)).cuda(), g_pos, g_nrm)
    torch.cuda.synchronize()
    gp, gn = g_pos.cpu().numpy().astype(np.float64), g_nrm.cpu().numpy().astype(np.float64)
    o = oracle_scene_from(scene)
    op, on = o.pt_grad_geometry(oracle_params(cams, cfg), adj.reshape(-1, cfg.width, 3))
    rp, rn = _rel(gp, op), _rel(gn, on)
    print(f"rel err positions {rp:.3e} normals {rn:.3e}")
    assert np.count_nonzero(op) > 1000 and np.count_nonzero(on) > 1000
    assert rp <= GRAD_REL and rn <= GRAD_REL, (rp, rn)
    nh = scene["meshes"][0]["positions"].shape[0]
    assert np.all(gn[nh:] == 0.0)
    assert system.geometry_vertex_range(BlasHandle(1), 0) == (nh, nv - nh)


def test_geometry_replay_leaves_material_gradient_intact(geo_case):
    """The replay overwrites the path state but not the recorded vertices:
    the material gradient after it equals the one before."""
    import torch

    _, system, tracer, cams, cfg, adj = geo_case
    _say("material gradient before / after a geometry replay")
    nv = system.scene_info().verts
    tracer.render(cams, cfg)
    a_t = torch.from_numpy(adj.astype(np.float32)).cuda()
    ga0, ge0 = torch.empty((2, 3), device="cuda"), torch.empty((2, 3), device="cuda")
    tracer.backward(a_t, ga0, ge0)
    g_pos = torch.empty((nv, 3), dtype=torch.float32, device="cuda")
    g_nrm = torch.empty((nv, 3), dtype=torch.float32, device="cuda")
    tracer.backward_geometry(a_t, g_pos, g_nrm)
    ga1, ge1 = torch.empty((2, 3), device="cuda"), torch.empty((2, 3), device="cuda")
    tracer.backward(a_t, ga1, ge1)
    torch.cuda.synchronize()
    # reproducible sums: the same bits, not merely close
    assert torch.equal(ga0, ga1) and torch.equal(ge0, ge1)


def test_adjoints_are_bit_reproducible(geo_case):
    """Two backward calls on one recorded frame give identical bits for the
    material and the geometry gradients, and two L2 calls the same loss."""
    import torch

    _, system, tracer, cams, cfg, adj = geo_case
    _say("bit-reproducible material / geometry gradients and L2 loss")
    nv = system.scene_info().verts
    img = tracer.render(cams, cfg).clone()
    a_t = torch.from_numpy(adj.astype(np.float32)).cuda()
    outs = []
    for _ in range(2):
        ga, ge = torch.empty((2, 3), device="cuda"), torch.empty((2, 3), device="cuda")
        g_pos = torch.full((nv, 3), 7.0, dtype=torch.float32, device="cuda")
        g_nrm = torch.full((nv, 3), 7.0, dtype=torch.float32, device="cuda")
        tracer.backward_geometry(a_t, g_pos, g_nrm)
        tracer.backward(a_t, ga, ge)
        loss = torch.zeros(1, dtype=torch.float64, device="cuda")
        adj2 = torch.empty_like(img)
        tracer.l2_adjoint(img, img * 0.5 + 0.01, adj2, loss)
        torch.cuda.synchronize()
        outs.append((ga, ge, g_pos, g_nrm, loss))
    for x, y in zip(outs[0], outs[1]):
        assert torch.equal(x, y)
    assert float(outs[0][2].abs().sum()) > 0 and float(outs[0][3].abs().sum()) > 0
    assert float(outs[0][4].item()) > 0
    ref = float(((img.double() - (img * 0.5 + 0.01).double()) ** 2).sum())
    assert abs(outs[0][4].item() - ref) <= 1e-9 * ref


def test_non_finite_adjoint_marks_touched_entries(geo_case):
    """A NaN adjoint makes exactly the entries that any path term touches
    NaN (as float addition would); entries no term touches stay 0."""
    import torch

    _, system, tracer, cams, cfg, adj = geo_case
    _say("NaN adjoint: touched gradient entries NaN, others 0")
    nv = system.scene_info().verts
    tracer.render(cams, cfg)
    ones = torch.ones((2, cfg.height, cfg.width, 3), dtype=torch.float32, device="cuda")
    ga, ge = torch.empty((2, 3), device="cuda"), torch.empty((2, 3), device="cuda")
    tracer.backward(ones, ga, ge)
    gp = torch.empty((nv, 3), dtype=torch.float32, device="cuda")
    gn = torch.empty((nv, 3), dtype=torch.float32, device="cuda")
    tracer.backward_geometry(ones, gp, gn)
    nan = torch.full_like(ones, float("nan"))
    ga2, ge2 = torch.empty_like(ga), torch.empty_like(ge)
    tracer.backward(nan, ga2, ge2)
    gp2, gn2 = torch.empty_like(gp), torch.empty_like(gn)
    tracer.backward_geometry(nan, gp2, gn2)
    torch.cuda.synchronize()
    for clean, bad in ((ga, ga2), (ge, ge2), (gp, gp2), (gn, gn2)):
        touched = clean != 0
        assert bool(torch.isnan(bad[touched]).all())
        assert bool((bad[~touched & ~torch.isnan(bad)] == 0).all())
