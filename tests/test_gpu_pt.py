"""GPU parity of the wavefront path tracer and its device primitives.

* compaction and Sobol: bit-exact against the oracle (integer work);
* path tracer: closed forms (furnace, lit floor) to FP32 accuracy, and the
  parametric scene against the FP64 oracle within the stated image error:
  relative RMSE <= 1e-2 of the mean radiance, at least 99 % of pixels within
  1e-3 relative, per-bounce ray counts within 0.5 %;
* a graph replay equals the eager frame bit for bit.
"""

import numpy as np
import pytest

from oracle import oracle as O
from rendiation_amd import synth
from rendiation_amd.acceleration import AccelerationStructureSystem, BuildOptions
from rendiation_amd.pathtracer import Camera, PathTracer, PathTracingConfig, load_scene
from tests.pt_util import floor_scene, furnace_scene, oracle_params, oracle_scene_from

pytestmark = pytest.mark.gpu

IMAGE_RMSE_REL = 1e-2
PIXEL_REL = 1e-3
PIXEL_FRACTION = 0.99


def _say(msg):
    print(f"[gpu-case] {msg}", flush=True)


def _system(scene):
    s = AccelerationStructureSystem(device=0, options=BuildOptions())
    load_scene(s, scene)
    return s


@pytest.mark.parametrize("n", [1, 7, 255, 256, 257, 4096, 100_003, 1 << 20])
def test_compaction_matches_copy_if(n, device):
    import torch

    _say(f"compaction n={n}")
    rng = np.random.default_rng(n)
    mask = (rng.random(n) < 0.37).astype(np.uint32)
    s = AccelerationStructureSystem(device=0)
    m = torch.from_numpy(mask.view(np.int32)).to(device)
    out = torch.full((n,), -1, dtype=torch.int32, device=device)
    cnt = torch.zeros(1, dtype=torch.int32, device=device)
    s.compact(m, out, cnt)
    s.sync()
    want, k = O.compact(mask)
    assert int(cnt.item()) == k
    assert np.array_equal(out.cpu().numpy()[:k].view(np.uint32), want[:k])


def test_compaction_reference_vector(device):
    import torch

    vals = np.array([1, 0, 1, 0, 1, 1, 0], np.uint32)
    s = AccelerationStructureSystem(device=0)
    out = torch.zeros(7, dtype=torch.int32, device=device)
    cnt = torch.zeros(1, dtype=torch.int32, device=device)
    s.compact(torch.from_numpy((vals == 1).astype(np.int32)).to(device), out, cnt)
    s.sync()
    idx = out.cpu().numpy()[: int(cnt.item())]
    compacted = np.zeros(7, np.uint32)
    compacted[: idx.size] = vals[idx]
    assert int(cnt.item()) == 4 and list(compacted) == [1, 1, 1, 1, 0, 0, 0]


def test_sobol_bit_exact(device):
    import torch

    _say("sobol samples")
    n = 1 << 16
    idx = synth.splitmix64(9, 1, n).astype(np.uint32)
    idx[:1024] = np.arange(1024, dtype=np.uint32)
    scr = synth.splitmix64(9, 2, 2 * n).astype(np.uint32)
    scr[:2048] = 0
    s = AccelerationStructureSystem(device=0)
    out = torch.empty(2 * n, dtype=torch.float32, device=device)
    s.sample_sobol(torch.from_numpy(idx.view(np.int32)).to(device), torch.from_numpy(scr.view(np.int32)).to(device),
                   out)
    s.sync()
    got = out.cpu().numpy()
    want = np.array([[O.van_der_corput(int(i), int(scr[2 * k])), O.sobol(int(i), int(scr[2 * k + 1]))]
                     for k, i in enumerate(idx[:4096])], np.float32).reshape(-1)
    assert np.array_equal(got[:8192], want)
    # the rest against the closed form of dimension 0 (bit reversal, XOR, 24 bits)
    rev = np.array([int(format(int(i), "032b")[::-1], 2) for i in idx[4096:8192]], dtype=np.uint64)
    d0 = np.minimum(((rev ^ scr[2 * 4096:2 * 8192:2].astype(np.uint64)) >> np.uint64(8)).astype(np.float64) / 2 ** 24,
                    1 - 2 ** -23).astype(np.float32)
    assert np.array_equal(got[2 * 4096:2 * 8192:2], d0)


def _render(system, cam, cfg):
    pt = PathTracer(system)
    img = pt.render(cam, cfg)
    system.sync()
    return img.cpu().numpy().astype(np.float64), pt.stats()


def test_furnace_on_gpu(device):
    a, E, N = 0.6, 0.5, 4
    _say("furnace")
    s = _system(furnace_scene(a, E))
    cam = Camera(position=(0.0, 0.0, 0.0), target=(0.0, 0.0, -1.0), fov_y_deg=60.0, aspect=1.0)
    cfg = PathTracingConfig(width=64, height=64, spp=2, max_bounces=N, rr_start=99, sun_radiance=(0, 0, 0))
    img, st = _render(s, cam, cfg)
    a32 = float(np.float32(a))
    want = E * sum(a32 ** k for k in range(N + 1))
    # The reference triangle test is not watertight: a ray can slip between
    # two triangles of the closed sphere.  The oracle traces the same paths,
    # so it leaks at the same places; elsewhere both equal the closed form.
    ref, ext, _ = oracle_scene_from(furnace_scene(a, E)).pt_render(oracle_params(cam, cfg))
    assert np.allclose(img, ref, rtol=1e-5)
    assert [int(x) for x in st["extension_rays"]] == [int(x) for x in ext]
    closed = np.abs(img - want) <= 1e-5 * want
    assert closed.mean() >= 0.999, closed.mean()


def test_floor_on_gpu(device):
    a, E, S = 0.5, 0.25, 3.0
    _say("lit floor")
    s = _system(floor_scene(a, E))
    cam = Camera(position=(0.0, 5.0, 0.0), target=(0.0, 0.0, 0.0), up=(0.0, 0.0, -1.0), fov_y_deg=40.0, aspect=1.0)
    sun = np.array([0.3, -1.0, 0.2])
    sun /= np.linalg.norm(sun)
    cfg = PathTracingConfig(width=32, height=32, spp=1, max_bounces=1, rr_start=99, sky_up=(0.7, 0.7, 0.7),
                            sky_down=(0.3, 0.3, 0.3), sun_direction=tuple(sun), sun_radiance=(S, S, S))
    img, st = _render(s, cam, cfg)
    cos_sun = float(np.float32(-sun.astype(np.float32)[1]))
    want = E + a / np.pi * cos_sun * S + a * 0.7
    assert np.allclose(img, want, rtol=2e-6)
    assert st["shadow_rays"][:1] == [32 * 32]


@pytest.mark.parametrize("spp,wide,coop,height", [(1, 0, 0, 108), (4, 0, 0, 108), (1, 1, 0, 108), (1, 2, 0, 108),
                                                  (1, 1, 1 << 22, 108), (1, 1, 2000, 108), (1, 1, 0, 112),
                                                  (2, 1, 262144, 112)])
def test_parametric_scene_against_fp64_oracle(spp, wide, coop, height, device):
    """coop = 1 << 22: every traversal launch through the cooperative kernel;
    2000: only the small tail launches.  Height 112 (both sides multiples of
    8): paths walk the image in 8 x 8 tiles; 108: row by row."""
    scene = synth.pt_scene(seed=1, n_meshes=64, tess=48)
    cam = Camera()
    cfg = PathTracingConfig(width=192, height=height, spp=spp, max_bounces=4, rr_start=2, seed=5)
    _say(f"parametric scene 64 meshes tess 48, 192x{height} spp {spp}, wide={wide}, coop={coop}")
    s = _system(scene)
    s.set_tune(wide=wide, coop=coop)
    img, st = _render(s, cam, cfg)
    ref, ext, sh = oracle_scene_from(scene).pt_render(oracle_params(cam, cfg))
    rmse = float(np.sqrt(np.mean((img - ref) ** 2)))
    rel = rmse / float(ref.mean())
    pix_err = np.abs(img - ref).max(axis=2) / np.maximum(ref.max(axis=2), 1e-6)
    frac = float((pix_err <= PIXEL_REL).mean())
    print(f"spp={spp} rel_rmse={rel:.3e} pixels_within={frac:.5f} ext={st['extension_rays']} ref_ext={list(ext)} "
          f"shadow={st['shadow_rays'][:-1]} ref_shadow={list(sh[:-1])}")
    assert rel <= IMAGE_RMSE_REL
    assert frac >= PIXEL_FRACTION
    for g, r in zip(st["extension_rays"], ext):
        assert abs(int(g) - int(r)) <= 0.005 * max(int(r), 1) + 2
    for g, r in zip(st["shadow_rays"][:-1], sh[:-1]):
        assert abs(int(g) - int(r)) <= 0.005 * max(int(r), 1) + 2


def test_full_path_tracing_frame_against_fp64_oracle(device):
    """The benchmark frame at full size: 64 UV spheres / tori tessellated at
    256 x 256 (8,388,608 triangles, 64 BLAS, random affine instances in
    [-10, 10]^3), 1920 x 1080 at 1 spp, up to 4 bounces, Russian roulette
    from bounce 2, seed 0, through the graph-replayed frame with the default
    launch shape and the autotuned one.  Per-bounce extension and shadow ray
    counts equal the FP64 oracle's exactly (every path decision is FP32 on
    both sides); the image within the stated error."""
    scene = synth.pt_scene(seed=0, n_meshes=64, tess=256)
    cam = Camera()
    cfg = PathTracingConfig(width=1920, height=1080, spp=1, max_bounces=4, rr_start=2, seed=0, use_graph=True)
    _say("full frame: 8.39M triangles, 1920x1080, 1 spp, 4 bounces, RR from 2")
    s = _system(scene)
    assert s.scene_info().tris == 64 * 2 * 256 * 256
    pt = PathTracer(s)
    imgs = [pt.render(cam, cfg).cpu().numpy().astype(np.float64)]
    stats = [pt.stats()]
    pt.autotune(cam, cfg, reps=1)
    imgs.append(pt.render(cam, cfg).cpu().numpy().astype(np.float64))
    stats.append(pt.stats())
    s.sync()
    ref, ext, sh = oracle_scene_from(scene).pt_render(oracle_params(cam, cfg))
    for img, st in zip(imgs, stats):
        rmse = float(np.sqrt(np.mean((img - ref) ** 2)))
        rel = rmse / float(ref.mean())
        pix_err = np.abs(img - ref).max(axis=2) / np.maximum(ref.max(axis=2), 1e-6)
        frac = float((pix_err <= PIXEL_REL).mean())
        print(f"rel_rmse={rel:.3e} pixels_within={frac:.6f} ext={st['extension_rays']} shadow={st['shadow_rays']}")
        assert rel <= IMAGE_RMSE_REL
        assert frac >= PIXEL_FRACTION
        assert [int(x) for x in st["extension_rays"]] == [int(x) for x in ext]
        assert [int(x) for x in st["shadow_rays"][:-1]] == [int(x) for x in sh[:-1]]
    assert np.array_equal(imgs[0], imgs[1])  # launch shapes change the schedule, not the result
    s.close()


def test_graph_replay_equals_eager(device):
    scene = synth.pt_scene(seed=3, n_meshes=16, tess=32, extent=5.0)
    cam = Camera(position=(0.0, 0.0, 18.0))
    _say("graph replay vs eager")
    s = _system(scene)
    pt = PathTracer(s)
    cfg = PathTracingConfig(width=160, height=90, spp=2, max_bounces=4, use_graph=False)
    eager = pt.render(cam, cfg).clone()
    s.sync()
    cfg.use_graph = True
    out = None
    for _ in range(3):
        out = pt.render(cam, cfg, out=out)
        s.sync()
        assert np.array_equal(out.cpu().numpy(), eager.cpu().numpy())
    # a new launch shape recaptures the graph
    s.set_tune(block_size=128, wide=1)
    cfg.use_graph = False
    eager_wide = pt.render(cam, cfg).clone()
    s.sync()
    cfg.use_graph = True
    for _ in range(2):
        out = pt.render(cam, cfg, out=out)
        s.sync()
        assert np.array_equal(out.cpu().numpy(), eager_wide.cpu().numpy())
    s.set_tune()


def test_views_in_one_frame(device):
    """Three cameras in one frame equal three single-camera frames with seeds
    seed + v (bit for bit), and the FP64 oracle's multi-view frame."""
    import dataclasses

    scene = synth.pt_scene(seed=3, n_meshes=16, tess=24, extent=5.0)
    cams = [Camera(position=(0.0, 0.0, 18.0)), Camera(position=(14.0, 6.0, 9.0)),
            Camera(position=(-10.0, -12.0, 8.0), up=(0.0, 0.0, 1.0))]
    _say("three views in one frame")
    s = _system(scene)
    pt = PathTracer(s)
    cfg = PathTracingConfig(width=64, height=40, spp=2, max_bounces=3, seed=9)
    multi = pt.render(cams, cfg).cpu().numpy()
    assert multi.shape == (3, 40, 64, 3)
    for v, cam in enumerate(cams):
        single = pt.render(cam, dataclasses.replace(cfg, seed=cfg.seed + v)).cpu().numpy()
        assert np.array_equal(multi[v], single), v
    o = oracle_scene_from(scene)
    ref, _, _ = o.pt_render(oracle_params(cams, cfg))
    ref = ref.reshape(3, 40, 64, 3)
    rel = np.sqrt(np.mean((multi - ref) ** 2)) / ref.mean()
    assert rel <= IMAGE_RMSE_REL, rel


def test_shared_launches_equal_separate_launches(device):
    """A frame traces the shadow rays of bounce b with the extension rays of
    bounce b + 1 in one launch; a fast-order counting frame traces every ray
    set in a launch of its own.  Both give the same image bit for bit, and
    the launch counts are max_bounces + 1 and 2 * max_bounces + 1."""
    import dataclasses

    from rendiation_amd import _lib

    scene = synth.pt_scene(seed=3, n_meshes=16, tess=32, extent=5.0)
    cam = Camera(position=(0.0, 0.0, 18.0))
    _say("shared shadow + extension launches vs separate launches")
    s = _system(scene)
    pt = PathTracer(s)
    cfg = PathTracingConfig(width=160, height=90, spp=2, max_bounces=4, use_graph=False,
                            diagnostics=_lib.RDRT_PT_DIAG_TIME_TRACES)
    shared = pt.render(cam, cfg).cpu().numpy()
    st_shared = pt.stats()
    sep = pt.render(cam, dataclasses.replace(cfg, diagnostics=_lib.RDRT_PT_DIAG_COUNT_FAST)).cpu().numpy()
    st_sep = pt.stats()
    assert np.array_equal(shared, sep)
    assert st_shared["trace_launches"] == cfg.max_bounces + 1
    assert st_sep["trace_launches"] == 2 * cfg.max_bounces + 1
    assert st_shared["extension_rays"] == st_sep["extension_rays"]
    assert st_shared["shadow_rays"] == st_sep["shadow_rays"]
    assert len(st_sep["max_ray_visits"]) == 2 * cfg.max_bounces + 1
    assert st_sep["max_ray_visits"][0] > 0 and st_sep["visits"][0] > 0
    with pytest.raises(Exception):
        pt.render(cam, dataclasses.replace(cfg, diagnostics=_lib.RDRT_PT_DIAG_COUNT_FAST
                                                 | _lib.RDRT_PT_DIAG_COUNT_VISITS))


@pytest.mark.parametrize("tail,wide,block", [(1, 1, 256), (2, 0, 64), (3, 2, 128), (1, 2, 64)])
def test_tail_kernel_equals_per_bounce_launches(tail, wide, block, device):
    """The tail kernel (every path finished by one launch from bounce
    `tail` on, 8 lanes per path) renders the frame of the per-bounce
    launches: the same image and the same recorded path vertices bit for
    bit, the same extension and shadow ray counts per bounce, and
    tail + 2 traversal launches."""
    import dataclasses

    from rendiation_amd import _lib

    scene = synth.pt_scene(seed=7, n_meshes=32, tess=40, extent=6.0)
    cam = Camera(position=(0.0, 0.0, 20.0))
    _say(f"tail kernel from bounce {tail}, layout {wide}, block {block}")
    s = _system(scene)
    pt = PathTracer(s)
    cfg = PathTracingConfig(width=160, height=96, spp=2, max_bounces=4, rr_start=2, seed=3, use_graph=False,
                            diagnostics=_lib.RDRT_PT_DIAG_TIME_TRACES)
    # recorded frames (the adjoint's input) have no Russian roulette
    rec_cfg = dataclasses.replace(cfg, rr_start=99, record=True)
    import torch

    nm = len(scene["albedo"])
    adj = torch.from_numpy(synth.uniform_f32(11, 1, 96 * 160 * 3) * np.float32(2.0) - np.float32(1.0)).cuda()

    def frames():
        img = pt.render(cam, cfg).cpu().numpy()
        st = pt.stats()
        rec_img = pt.render(cam, rec_cfg).cpu().numpy()
        ga = torch.zeros((nm, 3), dtype=torch.float32, device="cuda:0")
        ge = torch.zeros_like(ga)
        pt.backward(adj, ga, ge)  # reads the recorded path vertices
        s.sync()
        return img, st, rec_img, ga.cpu().numpy(), ge.cpu().numpy()

    s.set_tune(block_size=block, wide=wide, persistent=1, coop=4096)
    base, st_base, rec0, ga0, ge0 = frames()
    s.set_tune(block_size=block, wide=wide, persistent=1, coop=4096, tail=tail)
    img, st, rec, ga, ge = frames()
    assert st["trace_launches"] == tail + 2
    assert st["extension_rays"] == st_base["extension_rays"]
    assert st["shadow_rays"] == st_base["shadow_rays"]
    assert np.array_equal(img, base)
    assert np.array_equal(rec, rec0)
    assert np.array_equal(ga, ga0) and np.array_equal(ge, ge0)
    # graph-replayed frames with the tail equal the eager one
    g = dataclasses.replace(cfg, use_graph=True, diagnostics=0)
    for _ in range(2):
        assert np.array_equal(pt.render(cam, g).cpu().numpy(), base)
    s.set_tune()


def test_frame_autotune(device):
    """The frame autotuner keeps one shape of the search space; frames
    afterwards equal an eager frame rendered with that shape set by hand."""
    import dataclasses

    scene = synth.pt_scene(seed=3, n_meshes=16, tess=24, extent=5.0)
    cam = Camera(position=(0.0, 0.0, 18.0))
    _say("frame autotune")
    s = _system(scene)
    pt = PathTracer(s)
    cfg = PathTracingConfig(width=96, height=64, spp=1, max_bounces=3, use_graph=True)
    choice = pt.autotune(cam, cfg, reps=1)
    print("pt autotune:", choice)
    assert choice["block_size"] in (64, 128, 256) and choice["wide"] in (0, 1, 2) and choice["ms"] > 0
    img = pt.render(cam, cfg).cpu().numpy()
    s.set_tune(choice["block_size"], choice["lds_entries"], choice["waves"], choice["persistent"], choice["cull"],
               choice["wide"], choice["coop"], choice["tail"])
    eager = pt.render(cam, dataclasses.replace(cfg, use_graph=False)).cpu().numpy()
    assert np.array_equal(img, eager)
    with pytest.raises(ValueError):
        pt.autotune(cam, dataclasses.replace(cfg, record=True))
