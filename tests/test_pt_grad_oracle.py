"""Pins of the oracle's FP64 material gradient (orc_pt_grad) by means that do
not restate it: central finite differences of the oracle's own FP64 frame
(exact parameter steps), and the exact linearity of the frame in emission."""

import numpy as np
import pytest

from rendiation_amd import synth
from rendiation_amd.pathtracer import Camera, PathTracingConfig
from tests.pt_util import oracle_params, oracle_scene_from

H = 1e-3


@pytest.fixture(scope="module")
def grad_case():
    scene = synth.pt_scene(seed=4, n_meshes=16, tess=12, extent=5.0, materials_per_mesh=4)
    o = oracle_scene_from(scene)
    cam = Camera(position=(0.0, 0.0, 18.0), aspect=1.5)
    cfg = PathTracingConfig(width=48, height=32, spp=2, max_bounces=3, rr_start=3, seed=5)
    prm = oracle_params(cam, cfg)
    rng = np.random.default_rng(0)
    adj = rng.uniform(-1.0, 1.0, size=(cfg.height, cfg.width, 3))
    albedo = np.asarray(scene["albedo"], np.float32)
    emission = np.asarray(scene["emission"], np.float32)
    ga, ge = o.pt_grad(prm, adj, albedo.shape[0])
    return o, prm, adj, albedo, emission, ga, ge


def _f(o, prm, adj, albedo, emission):
    o.set_materials(albedo, emission)
    img, _, _ = o.pt_render(prm)
    return float(np.sum(adj * img))


def test_gradient_is_not_trivial(grad_case):
    _, _, _, albedo, _, ga, ge = grad_case
    assert np.count_nonzero(ga) > albedo.size // 2 and np.count_nonzero(ge) > albedo.size // 2


def test_emission_gradient_is_exact_linearity(grad_case):
    """The frame is linear in emission: F(e + d) - F(e) = g_e . d exactly."""
    o, prm, adj, albedo, emission, _, ge = grad_case
    rng = np.random.default_rng(1)
    d = rng.uniform(0.0, 1.0, size=emission.shape).astype(np.float32)
    e1 = (emission + d).astype(np.float32)
    step = e1.astype(np.float64) - emission.astype(np.float64)
    lhs = _f(o, prm, adj, albedo, e1) - _f(o, prm, adj, albedo, emission)
    rhs = float(np.sum(ge * step))
    assert abs(lhs - rhs) <= 1e-9 * max(abs(rhs), 1.0), (lhs, rhs)


def test_albedo_gradient_against_central_differences(grad_case):
    """Single coordinates with the largest gradients, and one random
    direction over every albedo, each against a central difference."""
    o, prm, adj, albedo, emission, ga, _ = grad_case
    scale = np.abs(ga).max()
    flat = np.argsort(-np.abs(ga).ravel())[:6]
    for idx in flat:
        m, c = divmod(int(idx), 3)
        ap, am = albedo.copy(), albedo.copy()
        ap[m, c] = np.float32(albedo[m, c] + H)
        am[m, c] = np.float32(albedo[m, c] - H)
        step = float(ap[m, c]) - float(am[m, c])
        fd = (_f(o, prm, adj, ap, emission) - _f(o, prm, adj, am, emission)) / step
        assert abs(fd - ga[m, c]) <= 1e-5 * scale, (m, c, fd, ga[m, c])
    rng = np.random.default_rng(2)
    d = rng.uniform(-1.0, 1.0, size=albedo.shape)
    ap = (albedo + H * d).astype(np.float32)
    am = (albedo - H * d).astype(np.float32)
    fd = _f(o, prm, adj, ap, emission) - _f(o, prm, adj, am, emission)
    lin = float(np.sum(ga * (ap.astype(np.float64) - am.astype(np.float64))))
    assert abs(fd - lin) <= 1e-5 * abs(lin), (fd, lin)
    o.set_materials(albedo, emission)


def test_gradient_refuses_russian_roulette(grad_case):
    o, prm, adj, albedo, _, _, _ = grad_case
    old = prm.rr_start
    prm.rr_start = 1
    try:
        with pytest.raises(ValueError):
            o.pt_grad(prm, adj, albedo.shape[0])
    finally:
        prm.rr_start = old
