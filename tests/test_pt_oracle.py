"""Pins of the FP64 path-tracing oracle by closed forms of simple scenes."""

import numpy as np

from rendiation_amd.pathtracer import Camera, PathTracingConfig
from tests.pt_util import floor_scene, furnace_scene, oracle_params, oracle_scene_from


def test_furnace_closed_form():
    # camera inside a closed inward-facing sphere, every surface albedo a,
    # emission E, no light, no roulette: L = E * (1 + a + ... + a^N)
    a, E, N = 0.6, 0.5, 4
    o = oracle_scene_from(furnace_scene(a, E))
    cam = Camera(position=(0.0, 0.0, 0.0), target=(0.0, 0.0, -1.0), fov_y_deg=60.0, aspect=1.0)
    cfg = PathTracingConfig(width=16, height=16, spp=2, max_bounces=N, rr_start=99, sun_radiance=(0, 0, 0))
    img, ext, sh = o.pt_render(oracle_params(cam, cfg))
    a32 = float(np.float32(a))  # materials are FP32 inputs
    want = E * sum(a32 ** k for k in range(N + 1))
    assert np.allclose(img, want, rtol=1e-12, atol=0)
    assert list(ext) == [16 * 16 * 2] * (N + 1)  # every path survives every bounce
    assert list(sh) == [0] * (N + 1)


def test_floor_sky_and_sun_closed_form():
    # camera above an upward-facing plane looking down; one bounce:
    # L = E + a/pi * cos(sun) * S + a * sky_up  (the BSDF ray always escapes up)
    a, E, S = 0.5, 0.25, 3.0
    o = oracle_scene_from(floor_scene(a, E))
    cam = Camera(position=(0.0, 5.0, 0.0), target=(0.0, 0.0, 0.0), up=(0.0, 0.0, -1.0), fov_y_deg=40.0, aspect=1.0)
    sun = np.array([0.3, -1.0, 0.2])
    sun /= np.linalg.norm(sun)
    cfg = PathTracingConfig(width=12, height=12, spp=1, max_bounces=1, rr_start=99, sky_up=(0.7, 0.7, 0.7),
                            sky_down=(0.3, 0.3, 0.3), sun_direction=tuple(sun), sun_radiance=(S, S, S))
    img, ext, sh = o.pt_render(oracle_params(cam, cfg))
    cos_sun = float(np.float32(-sun.astype(np.float32)[1]))
    want = E + a / np.pi * cos_sun * S + a * 0.7
    assert np.allclose(img, want, rtol=1e-6)
    assert list(ext) == [144, 144] and list(sh) == [144, 0]


def test_emission_is_linear():
    # the image is affine in the emission: L(E1) + L(E2) - L(0) = L(E1 + E2)
    from rendiation_amd import synth

    sc = synth.pt_scene(seed=2, n_meshes=8, tess=12, n_emissive=2, extent=3.0)
    cam = Camera(position=(0.0, 0.0, 12.0), fov_y_deg=45.0, aspect=1.0)
    cfg = PathTracingConfig(width=24, height=24, spp=1, max_bounces=3)
    imgs = []
    for e in (0.0, 2.0, 3.0, 5.0):
        s = dict(sc)
        em = np.zeros_like(sc["emission"])
        em[:2] = e
        s["emission"] = em
        imgs.append(oracle_scene_from(s).pt_render(oracle_params(cam, cfg))[0])
    assert np.allclose(imgs[1] + imgs[2] - imgs[0], imgs[3], rtol=1e-9, atol=1e-12)
    assert np.abs(imgs[3] - imgs[0]).max() > 0.01
