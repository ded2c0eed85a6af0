"""Benchmark of the ray-query hot path on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload all|pt|closest_hit|batched|diff|diffgeo|ao]

all (default): one JSON line for the whole metric -- the path-traced frame
  (pt) is the headline (value, roofline, cpu_baseline), and closest_hit,
  batched, diff, diffgeo and ao follow under "sub", each with its own
  timing, roofline, CPU baseline and error figures.  Every workload that
  traces carries "fractions": the HBM fraction of its traversal over the
  reference tree's visits, over the kernel's own visits and over the
  physical DRAM bytes (rocprofv3 FETCH_SIZE + WRITE_SIZE), and the L2 (TCC)
  hit rate of the same dispatches.
pt: wavefront path tracer on the parametric scene -- 64 UV spheres
  and tori tessellated at 256x256 (8,388,608 triangles, one BLAS each, 64
  random affine instances in [-10, 10]^3), pinhole camera at 1920x1080, 1 spp,
  Lambertian materials (albedo U[0.1, 0.9], 4 emissive meshes), up to 4
  bounces, Russian roulette from bounce 2, Sobol per-pixel scrambles, FP32
  radiance.  One step = one frame (a replayed hipGraph).  value = all rays
  traced in a frame (closest-hit + shadow) / frame time.
closest_hit: 100k random triangles (SAH 4 bins, leaf <= 4), 1M rays, seed 0;
  one step = one closest-hit launch.
batched: 64 random-triangle scenes (triangles log-uniform in [10k, 2M), rays
  log-uniform in [16k, 4M)), 3 bounces of cosine-weighted continuation rays
  with per-bounce compaction; one step = one graph-replayed persistent call
  over all scenes, compared with the per-scene sequential launches.
diff: differentiable path tracing -- 64 spheres/tori at 90x90 (1.04M
  triangles), 256 materials, 512x512 at 4 spp, 3 bounces; L2 loss against a
  target rendered from parameters perturbed by N(0, 0.05); value = ms of
  forward + adjoint per frame (lower is better), with the forward alone and
  the gradient error against the FP64 oracle beside it.
diffgeo: gradient w.r.t. 1M vertex positions and normals of a 1000x1000
  height field, 16 views at 512x512 in one frame, 2 spp, 3 bounces, paths
  recomputed by a replay of the frame; value = ms of forward + adjoint.  Also
  reports the BVH refit after a descent step on the positions (ms_refit).

roofline: HBM-bound.  Algorithmic bytes per traced ray = 32 (ray) + 16 (hit)
+ 48 x node visits + 52 x triangle tests + 176 x instance tests, with visit
counts of the reference traversal order on the same rays, divided by the
time of the traversal launches (HIP events on their stream).
cpu_baseline: the CPU oracle (OpenMP over rays / pixels on this host).
"""

from __future__ import annotations

import argparse
import json
import os
import subprocess
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

METRIC = "Mrays/s closest-hit + ms/frame fwd+bwd at stated RMSE, batched vs plain on 1 GPU"
HBM_PEAK_GBPS = 8000.0


def _cpu_model() -> str:
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True).stdout
        for line in out.splitlines():
            if line.startswith("Model name"):
                return line.split(":", 1)[1].strip()
    except Exception:
        pass
    return "unknown"


def _host_threads() -> int:
    try:
        return len(os.sched_getaffinity(0))
    except Exception:
        return os.cpu_count() or 1


def _bytes(n_rays: int, visits) -> int:
    return 48 * n_rays + 48 * int(visits[0]) + 52 * int(visits[1]) + 176 * int(visits[2])


NODE_BYTES = {0: 64, 1: 128, 2: 64}  # child-pair, child-quad, quantized child-quad node by the tune's `wide`


def _kernel_bytes(n_rays: int, visits, wide: int = 0) -> int:
    """Bytes the fast kernel touches: ray and hit, one node of its layout per
    node visit (64 B pair, 128 B quad, 64 B quantized quad), a 64-byte
    triangle record per test, a 128-byte instance record per instance test."""
    return 48 * n_rays + NODE_BYTES.get(int(wide), 64) * int(visits[0]) + 64 * int(visits[1]) + 128 * int(visits[2])


def _kernel_order(n_rays: int, visits, ms: float, wide: int = 0) -> dict:
    """The same rate over the visits of the fast kernel's own tree and order
    (node visits, triangle tests, instance tests): the bytes the kernel
    really asks of the memory system per launch."""
    b = _kernel_bytes(n_rays, visits, wide)
    gbps = b / (ms * 1e-3) / 1e9
    return {"visits_per_ray": [round(float(x) / n_rays, 3) for x in visits], "bytes": b,
            "achieved": round(gbps, 1), "frac": round(gbps / HBM_PEAK_GBPS, 4)}


def _pt_traversal(pt, cams, cfg, wide: int) -> dict:
    """Traversal rates of one eager frame of this configuration: the time of
    its traversal launches (HIP events on the frame's stream), the rays they
    traced, and the rate over the reference tree's visits (reference order)
    and over the fast kernel's own visits (kernel order)."""
    import dataclasses

    from rendiation_amd import _lib

    pt.render(cams, dataclasses.replace(cfg, record=False, diagnostics=_lib.RDRT_PT_DIAG_TIME_TRACES))
    d_time = pt.stats()
    pt.render(cams, dataclasses.replace(cfg, record=False, diagnostics=_lib.RDRT_PT_DIAG_COUNT_VISITS))
    d_ref = pt.stats()
    pt.render(cams, dataclasses.replace(cfg, record=False, diagnostics=_lib.RDRT_PT_DIAG_COUNT_FAST))
    d_fast = pt.stats()
    rays = int(sum(d_time["extension_rays"]) + sum(d_time["shadow_rays"]))
    ms = d_time["trace_ms"]
    b_ref = _bytes(rays, d_ref["visits"])
    ref = {"bytes": b_ref, "achieved": round(b_ref / (ms * 1e-3) / 1e9, 1),
           "frac": round(b_ref / (ms * 1e-3) / 1e9 / HBM_PEAK_GBPS, 4),
           "visits_per_ray": [round(v / rays, 3) for v in d_ref["visits"]]}
    return {"trace_ms": round(ms, 4), "trace_launches": d_time["trace_launches"], "rays": rays,
            "reference_order": ref, "kernel_order": _kernel_order(rays, d_fast["visits"], ms, wide)}


def _call_kernels(tune) -> str:
    k = "k_trace_persist" if tune and tune.get("persistent") else "k_trace_fast"
    if tune and tune.get("coop"):
        k = f"{k} / k_trace_coop (launches of at most {tune['coop']} rays)"
    if tune and tune.get("cull"):
        return f"k_trace_cull + k_block_scan + k_scatter2 + {k} (one rdrt_trace call)"
    return k


def _tune_arg(tune) -> str:
    return ",".join(str(tune.get(k, 0)) for k in ("block_size", "lds_entries", "waves", "persistent", "cull", "wide",
                                                   "coop", "tail"))


def _record(value, unit, args, ms_per_step, config, roofline, extra=None):
    rec = {
        "metric": METRIC,
        "value": round(value, 3),
        "unit": unit,
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": config,
        "roofline": roofline,
    }
    if extra:
        rec.update(extra)
    return rec


# --------------------------------------------------------------- path tracer
def path_tracing(args):
    import numpy as np
    import torch

    from rendiation_amd import _lib, synth
    from rendiation_amd.acceleration import AccelerationStructureSystem
    from rendiation_amd.pathtracer import Camera, PathTracer, PathTracingConfig, load_scene

    dev = torch.device("cuda:0")
    stream = torch.cuda.Stream(device=dev)
    t0 = time.perf_counter()
    scene = synth.pt_scene(seed=0, n_meshes=64, tess=256)
    gen_ms = (time.perf_counter() - t0) * 1e3
    system = AccelerationStructureSystem(device=0, stream=stream.cuda_stream)
    t0 = time.perf_counter()
    load_scene(system, scene)
    build_upload_ms = (time.perf_counter() - t0) * 1e3
    info = system.scene_info()
    cam = Camera()
    cfg = PathTracingConfig(width=1920, height=1080, spp=1, max_bounces=4, rr_start=2, seed=0, use_graph=True)
    pt = PathTracer(system)
    with torch.cuda.stream(stream):
        out = torch.empty((cfg.height, cfg.width, 3), dtype=torch.float32, device=dev)
    stream.synchronize()

    if args.tune:
        system.set_tune(*[int(x) for x in args.tune.split(",")])
        tune = system.tune()
    elif not args.no_autotune:
        tune = pt.autotune(cam, cfg, out=out, reps=3)
    else:
        tune = system.tune()
    if args.profile_only:  # the frames alone, for hardware counters
        for _ in range(args.steps):
            pt.render(cam, cfg, out=out)
        system.sync()
        system.close()
        return None
    for _ in range(args.warmup):
        pt.render(cam, cfg, out=out)
    system.sync()
    torch.cuda.synchronize()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    wall0 = time.perf_counter()
    e0.record(stream)
    for _ in range(args.steps):
        pt.render(cam, cfg, out=out)
    e1.record(stream)
    system.sync()
    torch.cuda.synchronize()
    wall = time.perf_counter() - wall0
    frame_ms_ev = e0.elapsed_time(e1) / args.steps
    ms_per_step = wall * 1e3 / args.steps
    st = pt.stats()
    rays = int(sum(st["extension_rays"]) + sum(st["shadow_rays"]))
    value = rays / (ms_per_step * 1e-3) / 1e6
    image = out.cpu().numpy().astype(np.float64)

    # diagnostics frames: traversal launch times, reference-order and kernel-order visits
    trav = _pt_traversal(pt, cam, cfg, tune.get("wide", 0))
    bytes_frame = trav["reference_order"]["bytes"]
    achieved = bytes_frame / (trav["trace_ms"] * 1e-3) / 1e9

    config = {
        "workload": "pt: 64 spheres/tori at 256x256 (8.39M tris), 1920x1080, 1 spp, <=4 bounces, RR from 2, seed 0",
        "tris": info.tris, "instances": info.instances, "width": cfg.width, "height": cfg.height, "spp": cfg.spp,
        "max_bounces": cfg.max_bounces, "rays_per_frame": rays, "extension_rays": st["extension_rays"],
        "shadow_rays": st["shadow_rays"][:-1], "frame_ms_events": round(frame_ms_ev, 4),
        "scene_gen_ms": round(gen_ms, 1), "build_upload_ms": round(build_upload_ms, 1),
        "device_scene_mb": round(info.device_bytes / 2 ** 20, 1), "graph_replay": True, "launch_shape": tune,
    }
    roofline = {
        "bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBPS, "unit": "GB/s",
        "frac": round(achieved / HBM_PEAK_GBPS, 4), "traffic": None,
        "kernel": f"{_call_kernels(dict(tune, cull=0))} (all traversal launches of one frame)",
        "trace_ms_per_frame": trav["trace_ms"], "trace_launches": trav["trace_launches"],
        "bytes_per_frame": bytes_frame, "visits_per_ray": trav["reference_order"]["visits_per_ray"],
        "kernel_order": trav["kernel_order"],
    }
    extra = {"ms_per_frame": round(ms_per_step, 4)}
    if not args.no_cpu_baseline:
        extra["cpu_baseline"], err = cpu_baseline_pt(scene, cam, cfg, image)
        extra["image_error"] = err
    system.close()
    return _record(value, "Mrays/s", args, ms_per_step, config, roofline, extra)


def cpu_baseline_pt(scene, cam, cfg, image):
    """FP64 oracle path tracer on a band of 24 rows (46,080 paths) of the same
    frame: its rays/s on this host, and the GPU image error on those pixels."""
    import numpy as np

    from tests.pt_util import oracle_params, oracle_scene_from

    o = oracle_scene_from(scene)
    prm = oracle_params(cam, cfg)
    threads = _host_threads()
    r0, r1 = cfg.height // 2 - 12, cfg.height // 2 + 12
    t0 = time.perf_counter()
    ref, ext, sh = o.pt_render_rows(prm, r0, r1, threads=threads)
    dt = time.perf_counter() - t0
    rays = int(ext.sum() + sh.sum())
    # one thread (the reference loop is sequential) on the middle 4 rows
    s0 = cfg.height // 2 - 2
    t1 = time.perf_counter()
    _, ext1, sh1 = o.pt_render_rows(prm, s0, s0 + 4, threads=1)
    dt1 = time.perf_counter() - t1
    rays1 = int(ext1.sum() + sh1.sum())
    gpu = image[r0:r1]
    rmse = float(np.sqrt(np.mean((gpu - ref) ** 2)))
    pix = np.abs(gpu - ref).max(axis=2) / np.maximum(ref.max(axis=2), 1e-6)
    base = {"value": round(rays / dt / 1e6, 3), "unit": "Mrays/s", "cores": threads, "kind": "port",
            "sample": f"rows {r0}-{r1 - 1} of the frame ({(r1 - r0) * cfg.width} paths, {rays} rays), FP64 shading",
            "single_thread_mrays": round(rays1 / dt1 / 1e6, 4),
            "single_thread_sample": f"rows {s0}-{s0 + 3} ({4 * cfg.width} paths, {rays1} rays), 1 thread",
            "cpu_model": _cpu_model()}
    err = {"vs": "FP64 CPU oracle on the same rows", "rmse": rmse, "rel_rmse": rmse / max(float(ref.mean()), 1e-12),
           "pixels_within_1e-3": float((pix <= 1e-3).mean()), "bound_rel_rmse": 1e-2}
    return base, err


# --------------------------------------------------------------- closest hit
def closest_hit(args):
    import numpy as np
    import torch

    from rendiation_amd import synth
    from rendiation_amd.acceleration import (
        TRAVERSAL_FAST,
        TRAVERSAL_REFERENCE,
        AccelerationStructureSystem,
        BuildOptions,
        TopLevelAccelerationStructureSourceInstance,
        TriangleGeometry,
    )

    dev = torch.device("cuda:0")
    stream = torch.cuda.Stream(device=dev)
    pos = synth.random_triangles(0, synth.CLOSEST_HIT_TRIS)
    rays_np = synth.random_rays(0, synth.CLOSEST_HIT_RAYS)
    n = rays_np.shape[0]
    system = AccelerationStructureSystem(device=0, options=BuildOptions(blas_bin_size=synth.CLOSEST_HIT_LEAF),
                                         stream=stream.cuda_stream)
    b = system.create_bottom_level_acceleration_structure([TriangleGeometry(pos)])
    t = system.create_top_level_acceleration_structure([TopLevelAccelerationStructureSourceInstance(np.eye(4), b)])
    system.bind_tlas([t])
    t0 = time.perf_counter()
    system.create_comp_instance()
    build_upload_ms = (time.perf_counter() - t0) * 1e3
    info = system.scene_info()

    with torch.cuda.stream(stream):
        rays = torch.from_numpy(rays_np).to(dev)
        hits = torch.empty((n, 4), dtype=torch.int32, device=dev)
        cnt = torch.empty((n, 3), dtype=torch.int32, device=dev)
    stream.synchronize()
    if args.tune:
        system.set_tune(*[int(x) for x in args.tune.split(",")])
    if args.profile_only:  # the launches alone, for hardware counters
        for _ in range(args.steps):
            system.trace(rays, hits, traversal=TRAVERSAL_FAST)
        system.sync()
        system.close()
        return None
    system.trace(rays, hits, counters=cnt, traversal=TRAVERSAL_REFERENCE)
    system.sync()
    c = cnt.cpu().numpy().astype(np.int64).sum(axis=0)
    tune = system.autotune(rays, hits, reps=3) if not (args.no_autotune or args.tune) else system.tune()
    bytes_per_launch = _bytes(n, c)
    # what the fast kernel itself visits (its own tree and order; culled rays visit nothing)
    system.trace(rays, hits, counters=cnt, traversal=TRAVERSAL_FAST)
    system.sync()
    c_fast = cnt.cpu().numpy().astype(np.int64).sum(axis=0)

    for _ in range(args.warmup):
        system.trace(rays, hits, traversal=TRAVERSAL_FAST)
    system.sync()
    torch.cuda.synchronize()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    wall0 = time.perf_counter()
    e0.record(stream)
    for _ in range(args.steps):
        system.trace(rays, hits, traversal=TRAVERSAL_FAST)
    e1.record(stream)
    system.sync()
    torch.cuda.synchronize()
    wall = time.perf_counter() - wall0
    kernel_ms = e0.elapsed_time(e1) / args.steps
    ms_per_step = wall * 1e3 / args.steps
    value = n / (ms_per_step * 1e-3) / 1e6
    achieved = bytes_per_launch / (kernel_ms * 1e-3) / 1e9
    hit_rate = float((hits.cpu().numpy()[:, 0] != -1).mean())
    config = {
        "workload": "closest_hit: 100k random triangles (SAH 4 bins, leaf<=4), 1M random rays, seed 0",
        "tris": synth.CLOSEST_HIT_TRIS, "rays": n, "leaf": synth.CLOSEST_HIT_LEAF,
        "bvh_nodes": info.tri_nodes, "bvh_depth": info.max_blas_depth, "traversal": "fast",
        "hit_fraction": round(hit_rate, 4), "build_upload_ms": round(build_upload_ms, 2), "launch_shape": tune,
    }
    roofline = {
        "bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBPS, "unit": "GB/s",
        "frac": round(achieved / HBM_PEAK_GBPS, 4), "traffic": None,
        "kernel": _call_kernels(tune), "kernel_ms": round(kernel_ms, 4), "bytes_per_launch": bytes_per_launch,
        "visits_per_ray": [round(float(x) / n, 3) for x in c],
        "kernel_order": _kernel_order(n, c_fast, kernel_ms, tune.get("wide", 0)),
    }
    extra = {}
    if not args.no_cpu_baseline:
        extra["cpu_baseline"] = cpu_baseline_closest_hit(pos, rays_np)
    system.close()
    return _record(value, "Mrays/s", args, ms_per_step, config, roofline, extra)


def cpu_baseline_closest_hit(pos, rays_np):
    """The oracle's reference traversal (OpenMP over rays, all host threads)
    on the same input; repeated 3 times (about 10 CPU-seconds)."""
    import numpy as np

    from oracle.oracle import OracleScene
    from rendiation_amd import synth

    orc = OracleScene(blas_bin_size=synth.CLOSEST_HIT_LEAF)
    b = orc.add_blas([(pos, None, 1)])
    orc.bind([orc.add_tlas([{"transform": np.eye(4), "blas": b}])])
    orc.build()
    threads = _host_threads()
    reps = 3
    t0 = time.perf_counter()
    for _ in range(reps):
        orc.trace(rays_np, threads=threads)
    dt = time.perf_counter() - t0
    one = rays_np[:50_000]
    t1 = time.perf_counter()
    orc.trace(one, threads=1)
    dt1 = time.perf_counter() - t1
    return {
        "value": round(reps * rays_np.shape[0] / dt / 1e6, 3), "unit": "Mrays/s", "cores": threads,
        "kind": "port", "sample": f"full workload ({rays_np.shape[0]} rays) x{reps}, OpenMP dynamic,1024 over rays",
        "single_thread_mrays": round(one.shape[0] / dt1 / 1e6, 4), "cpu_model": _cpu_model(),
    }


# ------------------------------------------------------- differentiable PT
def differentiable(args):
    import dataclasses

    import numpy as np
    import torch

    from rendiation_amd import synth
    from rendiation_amd.acceleration import AccelerationStructureSystem
    from rendiation_amd.pathtracer import Camera, PathTracer, PathTracingConfig, load_scene

    dev = torch.device("cuda:0")
    stream = torch.cuda.Stream(device=dev)
    scene = synth.pt_scene(seed=0, n_meshes=64, tess=90, materials_per_mesh=4)
    system = AccelerationStructureSystem(device=0, stream=stream.cuda_stream)
    t0 = time.perf_counter()
    load_scene(system, scene)
    build_upload_ms = (time.perf_counter() - t0) * 1e3
    info = system.scene_info()
    cam = Camera(aspect=1.0)
    cfg = PathTracingConfig(width=512, height=512, spp=4, max_bounces=3, rr_start=3, seed=0, use_graph=True)
    tracer = PathTracer(system)
    nmat = scene["albedo"].shape[0]
    rng = np.random.default_rng(0)
    a_t = np.clip(scene["albedo"] + rng.normal(0.0, 0.05, scene["albedo"].shape), 0.01, 0.99).astype(np.float32)
    e_t = np.maximum(scene["emission"] + rng.normal(0.0, 0.05, scene["emission"].shape), 0.0).astype(np.float32)
    with torch.cuda.stream(stream):
        albedo = torch.from_numpy(np.asarray(scene["albedo"], np.float32)).to(dev)
        emission = torch.from_numpy(np.asarray(scene["emission"], np.float32)).to(dev)
        target = torch.empty((cfg.height, cfg.width, 3), dtype=torch.float32, device=dev)
        img = torch.empty_like(target)
        adj = torch.empty_like(target)
        loss = torch.zeros(1, dtype=torch.float64, device=dev)
        ga = torch.empty_like(albedo)
        ge = torch.empty_like(emission)
        tracer.set_material_params(torch.from_numpy(a_t).to(dev), torch.from_numpy(e_t).to(dev))
    system.sync()
    # launch shape: autotuned on forward frames of this workload
    if args.tune:
        system.set_tune(*[int(x) for x in args.tune.split(",")])
        tune = system.tune()
    elif not args.no_autotune:
        tune = tracer.autotune(cam, cfg, out=target, reps=3)
    else:
        tune = system.tune()
    if args.profile_only:  # forward frames alone, for hardware counters
        for _ in range(args.steps):
            tracer.render(cam, cfg, out=target)
        system.sync()
        system.close()
        return None
    tracer.render(cam, cfg, out=target)
    system.sync()
    fwd_cfg = cfg
    rec_cfg = dataclasses.replace(cfg, record=True)

    def forward():
        tracer.set_material_params(albedo, emission)
        tracer.render(cam, fwd_cfg, out=img)

    def forward_backward():
        tracer.set_material_params(albedo, emission)
        tracer.render(cam, rec_cfg, out=img)
        tracer.l2_adjoint(img, target, adj, loss)
        tracer.backward(adj, ga, ge)

    def timed(fn):
        for _ in range(args.warmup):
            fn()
        system.sync()
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        w0 = time.perf_counter()
        e0.record(stream)
        for _ in range(args.steps):
            fn()
        e1.record(stream)
        system.sync()
        return (time.perf_counter() - w0) * 1e3 / args.steps, e0.elapsed_time(e1) / args.steps

    ms_f, ev_f = timed(forward)
    ms_fb, ev_fb = timed(forward_backward)
    # backward kernel alone, HIP events on the system's stream
    forward_backward()
    system.sync()
    b0 = torch.cuda.Event(enable_timing=True)
    b1 = torch.cuda.Event(enable_timing=True)
    b0.record(stream)
    for _ in range(args.steps):
        tracer.backward(adj, ga, ge)
    b1.record(stream)
    system.sync()
    bwd_ms = b0.elapsed_time(b1) / args.steps
    st = tracer.stats()
    trav = _pt_traversal(tracer, cam, cfg, tune.get("wide", 0))
    npaths = cfg.width * cfg.height * cfg.spp
    # backward bytes: vertex records (8 B x slots), pixel adjoint (12 B / path), gradient flush (24 B / material / block)
    slots = cfg.max_bounces + 2
    blocks = min((npaths + 255) // 256, 2048)
    bwd_bytes = npaths * (8 * slots + 12) + 24 * nmat * blocks
    achieved = bwd_bytes / (bwd_ms * 1e-3) / 1e9
    config = {
        "workload": "diff: 64 spheres/tori at 90x90 (1.04M tris), 256 materials, 512x512 x 4 spp, 3 bounces, "
                    "target from params + N(0, 0.05), L2 loss, grad w.r.t. 256x6 params",
        "tris": info.tris, "materials": nmat, "paths": npaths, "ms_forward": round(ms_f, 4),
        "ms_forward_events": round(ev_f, 4), "ms_forward_backward": round(ms_fb, 4),
        "ms_forward_backward_events": round(ev_fb, 4), "fwd_bwd_over_fwd": round(ms_fb / ms_f, 4),
        "ms_backward_kernel": round(bwd_ms, 4), "loss": float(loss.item()),
        "extension_rays": st["extension_rays"], "build_upload_ms": round(build_upload_ms, 1), "launch_shape": tune,
    }
    roofline = {
        "bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBPS, "unit": "GB/s",
        "frac": round(achieved / HBM_PEAK_GBPS, 4), "traffic": None, "kernel": "k_pt_backward",
        "bytes_per_launch": bwd_bytes,
    }
    extra = {"forward_traversal": trav}
    if not args.no_cpu_baseline:
        extra["cpu_baseline"], extra["gradient_error"] = cpu_baseline_diff(scene, cam, cfg, adj, ga, ge)
    system.close()
    rec = _record(ms_fb, "ms", args, ms_fb, config, roofline, extra)
    rec["higher_is_better"] = False
    return rec


def cpu_baseline_diff(scene, cam, cfg, adj, ga, ge):
    """FP64 oracle gradient of the same frame (full workload, all host
    threads): its time, and the relative error of the GPU gradient."""
    import numpy as np

    from tests.pt_util import oracle_params, oracle_scene_from

    o = oracle_scene_from(scene)
    prm = oracle_params(cam, cfg)
    threads = _host_threads()
    adj_np = adj.cpu().numpy().astype(np.float64)
    t0 = time.perf_counter()
    oa, oe = o.pt_grad(prm, adj_np, scene["albedo"].shape[0], threads=threads)
    dt = time.perf_counter() - t0
    g_a = ga.cpu().numpy().astype(np.float64)
    g_e = ge.cpu().numpy().astype(np.float64)

    def rel(a, b):
        return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30))

    npaths = cfg.width * cfg.height * cfg.spp
    base = {"value": round(dt * 1e3, 1), "unit": "ms (FP64 forward replay + adjoint)", "cores": threads,
            "kind": "port", "sample": f"full workload ({npaths} paths)", "cpu_model": _cpu_model()}
    err = {"vs": "FP64 CPU oracle gradient, same adjoint image", "rel_err_albedo": rel(g_a, oa),
           "rel_err_emission": rel(g_e, oe), "bound": 1e-3}
    return base, err


# ------------------------------------------------- differentiable geometry
def differentiable_geometry(args):
    import numpy as np
    import torch

    from rendiation_amd import synth
    from rendiation_amd.acceleration import AccelerationStructureSystem
    from rendiation_amd.pathtracer import Camera, PathTracer, PathTracingConfig, load_scene

    dev = torch.device("cuda:0")
    stream = torch.cuda.Stream(device=dev)
    free0 = torch.cuda.mem_get_info(dev)[0]  # device memory free before anything of this workload
    n = 1000
    t0 = time.perf_counter()
    hp, hi, hn = synth.height_field(n, seed=0)
    gen_ms = (time.perf_counter() - t0) * 1e3
    scene = {"meshes": [{"positions": hp, "indices": hi, "normals": hn, "tri_material": None}],
             "instances": [{"transform": np.eye(4), "mesh": 0, "material": 0}],
             "albedo": np.array([[0.6, 0.6, 0.6]], np.float32), "emission": np.zeros((1, 3), np.float32)}
    system = AccelerationStructureSystem(device=0, stream=stream.cuda_stream)
    t0 = time.perf_counter()
    load_scene(system, scene)
    build_upload_ms = (time.perf_counter() - t0) * 1e3
    info = system.scene_info()
    cams = [Camera(position=p, aspect=1.0, fov_y_deg=40.0) for p in synth.sphere_views(16, radius=3.0)]
    cfg = PathTracingConfig(width=512, height=512, spp=2, max_bounces=3, rr_start=3, seed=0, use_graph=True)
    tracer = PathTracer(system)
    nv = info.verts
    # target: the same views of a height field with other sinusoids
    tp, _, tn = synth.height_field(n, seed=1)
    tsys = AccelerationStructureSystem(device=0, stream=stream.cuda_stream)
    load_scene(tsys, {**scene, "meshes": [{"positions": tp, "indices": hi, "normals": tn, "tri_material": None}]})
    with torch.cuda.stream(stream):
        target = PathTracer(tsys).render(cams, cfg).clone()
        img = torch.empty_like(target)
        adj = torch.empty_like(target)
        loss = torch.zeros(1, dtype=torch.float64, device=dev)
        g_pos = torch.empty((nv, 3), dtype=torch.float32, device=dev)
        g_nrm = torch.empty((nv, 3), dtype=torch.float32, device=dev)
    system.sync()
    tsys.close()
    import dataclasses

    # launch shape: autotuned on forward frames of all 16 views
    if args.tune:
        system.set_tune(*[int(x) for x in args.tune.split(",")])
        tune = system.tune()
    elif not args.no_autotune:
        tune = tracer.autotune(cams, cfg, out=img, reps=1)
    else:
        tune = system.tune()
    if args.profile_only:  # forward frames alone, for hardware counters
        for _ in range(args.steps):
            tracer.render(cams, cfg, out=img)
        system.sync()
        system.close()
        return None
    rec_cfg = dataclasses.replace(cfg, record=True)

    def forward():
        tracer.render(cams, cfg, out=img)

    def forward_backward():
        tracer.render(cams, rec_cfg, out=img)
        tracer.l2_adjoint(img, target, adj, loss)
        tracer.backward_geometry(adj, g_pos, g_nrm)

    def timed(fn):
        for _ in range(args.warmup):
            fn()
        system.sync()
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        w0 = time.perf_counter()
        e0.record(stream)
        for _ in range(args.steps):
            fn()
        e1.record(stream)
        system.sync()
        return (time.perf_counter() - w0) * 1e3 / args.steps, e0.elapsed_time(e1) / args.steps

    ms_f, ev_f = timed(forward)
    ms_fb, ev_fb = timed(forward_backward)
    # every buffer of the workload is allocated by now (the library's are
    # grow-only, torch's cached): the device memory it holds at its peak
    peak_mb = (free0 - torch.cuda.mem_get_info(dev)[0]) / 2 ** 20
    st = tracer.stats()
    trav = _pt_traversal(tracer, cams, cfg, tune.get("wide", 0))
    npaths = len(cams) * cfg.width * cfg.height * cfg.spp
    config = {
        "workload": "diffgeo: 1000x1000-vertex height field (8 sinusoids, amplitude 0.05), 16 views at 512x512 in "
                    "one frame, 2 spp, 3 bounces, L2 loss, grad w.r.t. 1M vertex positions + normals",
        "tris": info.tris, "vertices": nv, "views": len(cams), "paths": npaths,
        "ms_forward": round(ms_f, 4), "ms_forward_events": round(ev_f, 4),
        "ms_forward_backward": round(ms_fb, 4), "ms_forward_backward_events": round(ev_fb, 4),
        "fwd_bwd_over_fwd": round(ms_fb / ms_f, 4), "loss": float(loss.item()),
        "extension_rays": st["extension_rays"], "scene_gen_ms": round(gen_ms, 1),
        "build_upload_ms": round(build_upload_ms, 1), "launch_shape": tune,
        "peak_device_mb": round(peak_mb, 1), "device_scene_mb": round(info.device_bytes / 2 ** 20, 1),
        "path_vertices": "recomputed per bounce by a replay of the recorded frame (8 B per path and bounce kept)",
    }
    extra = {"forward_traversal": trav}
    if not args.no_cpu_baseline:
        extra["cpu_baseline"], extra["gradient_error"] = cpu_baseline_diffgeo(scene, cams[:2], cfg, target, system,
                                                                             tracer)
    # a descent step on the positions, then the BVH refit to them (device
    # BLAS refit + host TLAS rebuild; synchronizes), alternating with the
    # committed positions
    with torch.cuda.stream(stream):
        pos0 = torch.from_numpy(hp).to(dev).reshape(-1)
        step = g_pos.reshape(-1) * (1e-3 / float(g_pos.abs().max().item() + 1e-30))
        pos1 = (pos0 - step).contiguous()
    system.sync()
    refit_ms = []
    for k in range(6):
        w0 = time.perf_counter()
        system.refit(pos1 if k % 2 == 0 else pos0)
        refit_ms.append((time.perf_counter() - w0) * 1e3)
    config["ms_refit"] = round(float(np.median(refit_ms[1:])), 3)
    config["ms_refit_first"] = round(refit_ms[0], 3)
    config["refit_speedup_vs_build"] = round(build_upload_ms / float(np.median(refit_ms[1:])), 1)
    system.close()
    rec = _record(ms_fb, "ms", args, ms_fb, config, None, extra)
    rec["higher_is_better"] = False
    return rec


def cpu_baseline_diffgeo(scene, cams, cfg, target, system, tracer):
    """The FP64 oracle geometry gradient of the first two views (all host
    threads), and the relative error of the GPU gradient of the same frame."""
    import dataclasses

    import numpy as np
    import torch

    from tests.pt_util import oracle_params, oracle_scene_from

    rec_cfg = dataclasses.replace(cfg, record=True)
    nv = system.scene_info().verts
    img = tracer.render(cams, rec_cfg)
    adj = (2.0 * (img - target[:len(cams)])).contiguous()
    g_pos = torch.empty((nv, 3), dtype=torch.float32, device=img.device)
    g_nrm = torch.empty((nv, 3), dtype=torch.float32, device=img.device)
    tracer.backward_geometry(adj, g_pos, g_nrm)
    system.sync()
    o = oracle_scene_from(scene)
    threads = _host_threads()
    t0 = time.perf_counter()
    op, on = o.pt_grad_geometry(oracle_params(cams, cfg), adj.cpu().numpy().astype(np.float64).reshape(
        -1, cfg.width, 3), threads=threads)
    dt = time.perf_counter() - t0

    def rel(a, b):
        return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30))

    gp = g_pos.cpu().numpy().astype(np.float64)
    gn = g_nrm.cpu().numpy().astype(np.float64)
    npaths = len(cams) * cfg.width * cfg.height * cfg.spp
    base = {"value": round(dt * 1e3, 1), "unit": "ms (FP64 forward + geometry adjoint)", "cores": threads,
            "kind": "port", "sample": f"views 0-1 of the frame ({npaths} paths)", "cpu_model": _cpu_model()}
    err = {"vs": "FP64 CPU oracle gradient, same adjoint image, views 0-1", "rel_err_positions": rel(gp, op),
           "rel_err_normals": rel(gn, on), "bound": 1e-3}
    return base, err


# ------------------------------------------------------------------ batched
def batched(args):
    import numpy as np
    import torch

    from rendiation_amd import synth
    from rendiation_amd.acceleration import (
        BATCH_PERSISTENT,
        BATCH_SEQUENTIAL,
        TRAVERSAL_FAST,
        TRAVERSAL_REFERENCE,
        AccelerationStructureSystem,
        BuildOptions,
        TopLevelAccelerationStructureSourceInstance,
        TriangleGeometry,
    )

    dev = torch.device("cuda:0")
    stream = torch.cuda.Stream(device=dev)
    t0 = time.perf_counter()
    scenes, rays_np, meta_np, offs = synth.batched_scenes(0, synth.BATCHED_SCENES, synth.BATCHED_TRIS,
                                                          synth.BATCHED_RAYS)
    gen_ms = (time.perf_counter() - t0) * 1e3
    n = rays_np.shape[0]
    bounces = synth.BATCHED_BOUNCES
    system = AccelerationStructureSystem(device=0, options=BuildOptions(blas_bin_size=synth.CLOSEST_HIT_LEAF),
                                         stream=stream.cuda_stream)
    tl = []
    for v in scenes:
        b = system.create_bottom_level_acceleration_structure([TriangleGeometry(v)])
        tl.append(system.create_top_level_acceleration_structure(
            [TopLevelAccelerationStructureSourceInstance(np.eye(4), b)]))
    system.bind_tlas(tl)
    t0 = time.perf_counter()
    system.create_comp_instance()
    build_upload_ms = (time.perf_counter() - t0) * 1e3
    info = system.scene_info()
    with torch.cuda.stream(stream):
        rays = torch.from_numpy(rays_np).to(dev)
        meta = torch.from_numpy(meta_np.view(np.int32)).to(dev)
        hits = torch.empty((bounces + 1) * n * 4, dtype=torch.int32, device=dev)
        spawned = torch.empty(bounces * n * 8, dtype=torch.float32, device=dev)
    stream.synchronize()

    # launch shape: autotuned on whole persistent calls of this workload
    if args.tune:
        system.set_tune(*[int(x) for x in args.tune.split(",")])
        tune = system.tune()
    elif not args.no_autotune:
        with torch.cuda.stream(stream):
            tune = system.batch_autotune(rays, meta, hits, bounces=bounces, seed=0, reps=1)
    else:
        tune = system.tune()

    def run(mode, steps, record=False):
        for _ in range(steps):
            system.trace_batched(rays, meta, hits, bounces=bounces, seed=0, spawned=spawned if record else None,
                                 scene_offsets=offs if mode == BATCH_SEQUENTIAL else None, mode=mode)

    def timed(mode):
        run(mode, args.warmup)
        system.sync()
        torch.cuda.synchronize()
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        w0 = time.perf_counter()
        e0.record(stream)
        run(mode, args.steps)
        e1.record(stream)
        system.sync()
        torch.cuda.synchronize()
        return (time.perf_counter() - w0) * 1e3 / args.steps, e0.elapsed_time(e1) / args.steps, system.batch_stats()

    if args.profile_only:  # the persistent calls alone, for hardware counters
        run(BATCH_PERSISTENT, args.steps)
        system.sync()
        system.close()
        return None
    ms_p, ev_p, st_p = timed(BATCH_PERSISTENT)
    hits_p = hits.view(torch.int64).sum().item()
    ms_s, ev_s, st_s = timed(BATCH_SEQUENTIAL)
    same = hits.view(torch.int64).sum().item() == hits_p
    # every primary ray is a query answered (the cull pass finishes the ones that miss every box)
    total = n + int(sum(st_p.rays[1:]))
    value = total / (ms_p * 1e-3) / 1e6

    # visit counts of the reference order on every round's rays (all work on the system's stream)
    run(BATCH_PERSISTENT, 1, record=True)
    system.sync()
    visits = np.zeros(3, dtype=np.int64)
    visits_fast = np.zeros(3, dtype=np.int64)
    with torch.cuda.stream(stream):
        cnt = torch.empty((n, 3), dtype=torch.int32, device=dev)
        hh = torch.empty((n, 4), dtype=torch.int32, device=dev)
        hv = hits.view(bounces + 1, n, 4)
        sv = spawned.view(bounces, n, 8)
        for r in range(bounces + 1):
            live = torch.ones(n, dtype=torch.bool, device=dev) if r == 0 else (hv[r - 1, :, 0] != -1)
            idx = torch.nonzero(live).flatten()
            if idx.numel() == 0:
                break
            src = rays if r == 0 else sv[r - 1]
            rr = src[idx].contiguous()
            mm = meta[idx].contiguous()
            system.trace(rr.view(-1), hh[:idx.numel()].view(-1), ray_meta=mm.view(-1),
                         counters=cnt[:idx.numel()].view(-1), traversal=TRAVERSAL_REFERENCE)
            visits += cnt[:idx.numel()].to(torch.int64).sum(dim=0).cpu().numpy()
            system.trace(rr.view(-1), hh[:idx.numel()].view(-1), ray_meta=mm.view(-1),
                         counters=cnt[:idx.numel()].view(-1), traversal=TRAVERSAL_FAST)
            visits_fast += cnt[:idx.numel()].to(torch.int64).sum(dim=0).cpu().numpy()
    system.sync()
    bytes_call = _bytes(total, visits)
    achieved = bytes_call / (ev_p * 1e-3) / 1e9
    config = {
        "workload": "batched: 64 random-triangle scenes (tris log-U[10k,2M), rays log-U[16k,4M)), 3 bounces, seed 0",
        "scenes": len(scenes), "tris": info.tris, "primary_rays": n, "rays_per_round": st_p.rays,
        "rays_per_call": total, "launches_persistent": st_p.launches, "launches_sequential": st_s.launches,
        "ms_persistent_events": round(ev_p, 4), "ms_sequential": round(ms_s, 4),
        "ms_sequential_events": round(ev_s, 4), "sequential_mrays": round(total / (ms_s * 1e-3) / 1e6, 3),
        "speedup_vs_sequential": round(ms_s / ms_p, 4), "sequential_equal": bool(same),
        "scene_gen_ms": round(gen_ms, 1), "build_upload_ms": round(build_upload_ms, 1),
        "device_scene_mb": round(info.device_bytes / 2 ** 20, 1), "launch_shape": tune,
    }
    roofline = {
        "bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBPS, "unit": "GB/s",
        "frac": round(achieved / HBM_PEAK_GBPS, 4), "traffic": None,
        "kernel": "k_trace_persist + k_spawn (whole batched call, graph replay)",
        "bytes_per_call": bytes_call, "visits_per_ray": [round(float(v) / total, 3) for v in visits],
        "kernel_order": _kernel_order(total, visits_fast, ev_p, tune.get("wide", 0)),
        "note": "achieved counts the reference tree's visits (4-bin longest-axis SAH) on every round's rays; the "
                "kernel walks its own full-SAH tree, which visits far fewer nodes on these large scenes, and "
                "finishes the rays that miss their scene box in the cull pass -- so the reference-order rate can "
                "pass the HBM peak; kernel_order gives the rate over the kernel's own visits",
    }
    extra = {}
    if not args.no_cpu_baseline:
        extra["cpu_baseline"] = cpu_baseline_batched(scenes, rays_np, meta_np, offs, bounces)
    system.close()
    return _record(value, "Mrays/s", args, ms_p, config, roofline, extra)


def cpu_baseline_batched(scenes, rays_np, meta_np, offs, bounces):
    """The oracle's batch rounds (trace + FP64 spawn, OpenMP over rays) on the
    first 4 scenes with at most 262,144 rays each."""
    import numpy as np

    from oracle.oracle import OracleScene
    from rendiation_amd import synth

    k = 4
    orc = OracleScene(blas_bin_size=synth.CLOSEST_HIT_LEAF)
    tl = []
    for v in scenes[:k]:
        b = orc.add_blas([(v, None, 1)])
        tl.append(orc.add_tlas([{"transform": np.eye(4), "blas": b}]))
    orc.bind(tl)
    orc.build()
    sel = np.concatenate([np.arange(offs[i], min(offs[i + 1], offs[i] + 262_144)) for i in range(k)])
    rays, meta = rays_np[sel].copy(), meta_np[sel].copy()
    threads = _host_threads()
    traced = 0
    t0 = time.perf_counter()
    live = np.ones(rays.shape[0], dtype=bool)
    for r in range(bounces + 1):
        traced += int(live.sum())
        hits, sp, alive = orc.batch_round(rays, meta, r + 1, 0, threads=threads)
        live = live & (alive != 0)
        rays = sp
        rays[~live] = np.array([0, 0, 0, 1, 0, 0, 1, 0], dtype=np.float32)
    dt = time.perf_counter() - t0
    return {"value": round(traced / dt / 1e6, 3), "unit": "Mrays/s", "cores": threads, "kind": "port",
            "sample": f"scenes 0-3, first <=262144 rays each ({sel.size} paths, {traced} rays over {bounces + 1} "
                      "rounds); ended paths still cost one empty-range query per round",
            "cpu_model": _cpu_model()}


# ------------------------------------------------------ HBM traffic (PMC)
PMC_STEPS = 3


# ------------------------------------------------------- ambient occlusion
def ambient_occlusion(args):
    """One AO sample per step of the path-tracing scene at the reference's AO
    size (the 1920x1080 frame clamped to 512 * 512 pixels, ao.rs:176,
    lib.rs:44-53), run as the reference's task-graph pipeline."""
    import numpy as np
    import torch

    from rendiation_amd import synth
    from rendiation_amd.acceleration import AccelerationStructureSystem
    from rendiation_amd.pathtracer import Camera, load_scene
    from rendiation_amd.task_graph import AORenderer

    dev = torch.device("cuda:0")
    stream = torch.cuda.Stream(device=dev)
    scene = synth.pt_scene(seed=0, n_meshes=64, tess=256)
    system = AccelerationStructureSystem(device=0, stream=stream.cuda_stream)
    load_scene(system, scene)
    info = system.scene_info()
    ratio = np.sqrt(np.float32(512 * 512) / np.float32(1920 * 1080)).astype(np.float32)
    W, H = int(np.floor(np.float32(1920) * ratio)), int(np.floor(np.float32(1080) * ratio))
    cam = Camera(aspect=W / H)
    r = AORenderer(system, W, H)
    if args.tune:
        system.set_tune(*[int(x) for x in args.tune.split(",")])
        tune = system.tune()
    elif not args.no_autotune:
        tune = r.autotune(cam, reps=3)
    else:
        tune = system.tune()
    if args.profile_only:
        for _ in range(args.steps):
            r.render(cam)
        system.sync()
        system.close()
        return None
    for _ in range(args.warmup):
        r.render(cam)
    system.sync()
    r.reset_sample()
    torch.cuda.synchronize()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    wall0 = time.perf_counter()
    e0.record(stream)
    for _ in range(args.steps):
        r.render(cam)
    e1.record(stream)
    system.sync()
    torch.cuda.synchronize()
    wall = time.perf_counter() - wall0
    ms_ev = e0.elapsed_time(e1) / args.steps
    ms_per_step = wall * 1e3 / args.steps
    prim, occ = r.ray_counts()
    value = (prim + occ) / (ms_per_step * 1e-3) / 1e6
    st = r.states()
    # parity on the first sample: a fresh buffer at sample 0 holds the values
    r.reset_sample()
    r.buffer.zero_()
    torch.cuda.synchronize()
    r.render(cam)
    system.sync()
    first = r.buffer.cpu().numpy()
    rays0 = sum(r.ray_counts())
    config = {
        "workload": f"ao: 64 spheres/tori at 256x256 (8.39M tris), {W}x{H} (1920x1080 clamped to 512^2 pixels), "
                    "1 sample per step, task-graph pipeline (trace, ray gen, 2 closest hit, miss), round hint 8 (5 run)",
        "tris": info.tris, "width": W, "height": H, "camera_rays": prim, "occlusion_rays": occ,
        "ms_events": round(ms_ev, 4), "graph_replay": True, "launch_shape": tune,
        "drained": st.is_empty() and not any(st.errors),
    }
    extra = {"ms_per_sample": round(ms_per_step, 4)}
    if not args.no_cpu_baseline:
        extra["cpu_baseline"], extra["parity"] = cpu_baseline_ao(scene, cam, W, H, first, rays0)
    system.close()
    return _record(value, "Mrays/s", args, ms_per_step, config, None, extra)


def pipeline(args):
    """The reference's ray-tracing pipeline test (shader/ray-tracing/src/
    test.rs:33-141: ray gen -> trace with back faces culled -> closest hit /
    miss -> ray-gen continuation, RGBA8 target) at 1920x1080 on the
    path-tracing scene, through rdrt_pipeline_* with the stage kernels of
    tests/stages (a caller's code object).  One call per step; the same
    camera rays traced by one plain rdrt_trace launch for comparison."""
    import ctypes as C

    import numpy as np
    import torch

    from rendiation_amd import synth
    from rendiation_amd.acceleration import AccelerationStructureSystem
    from rendiation_amd.pathtracer import load_scene
    from rendiation_amd.rt_pipeline import (GPURaytracingPipelineAndBindingSource, RayTracingPipeline,
                                            ShaderBindingTable)

    sys.path.insert(0, str(Path(__file__).resolve().parent))
    from tests import pipeline_expect as PE

    lib_path = Path(__file__).resolve().parent / "tests" / "stages" / "libpipeline_stages.so"
    stages = C.CDLL(str(lib_path))
    stages.test_stage.restype = C.c_void_p
    stages.test_stage.argtypes = [C.c_uint32]
    dev = torch.device("cuda:0")
    stream = torch.cuda.Stream(device=dev)
    scene = synth.pt_scene(seed=0, n_meshes=64, tess=256)
    system = AccelerationStructureSystem(device=0, stream=stream.cuda_stream)
    load_scene(system, scene)
    W, H = 1920, 1080
    n = W * H
    rays = torch.from_numpy(PE.camera_rays(W, H, 0.1, 100.0)).to(dev)
    hits = torch.zeros((n, 4), dtype=torch.int32, device=dev)
    torch.cuda.synchronize()
    if args.tune:
        system.set_tune(*[int(x) for x in args.tune.split(",")])
        tune = system.tune()
    elif not args.no_autotune:
        tune = system.autotune(rays, hits, flags=PE.CULL_BACK, reps=3)
    else:
        tune = system.tune()
    src = GPURaytracingPipelineAndBindingSource(payload_words=1)  # round hint 8, 2 in flight, default tile
    rg = src.register_ray_gen(stages.test_stage(0))
    ch = src.register_ray_closest_hit(stages.test_stage(1))
    ms = src.register_ray_miss(stages.test_stage(2))
    sbt = ShaderBindingTable(ray_type_count=1)
    sbt.config_ray_generation(rg)
    sbt.config_hit_group(0, 0, 0, ch)
    sbt.config_missing(0, ms)
    pipe = RayTracingPipeline(system, src, sbt)
    buf = torch.zeros((n * 20,), dtype=torch.uint8, device=dev)
    torch.cuda.synchronize()
    if args.profile_only:
        for _ in range(args.steps):
            pipe.trace_ray(W, H, buf.data_ptr())
        system.sync()
        pipe.close()
        system.close()
        return None
    for _ in range(args.warmup):
        pipe.trace_ray(W, H, buf.data_ptr())
    system.sync()
    torch.cuda.synchronize()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    wall0 = time.perf_counter()
    e0.record(stream)
    for _ in range(args.steps):
        pipe.trace_ray(W, H, buf.data_ptr())
    e1.record(stream)
    system.sync()
    torch.cuda.synchronize()
    wall = time.perf_counter() - wall0
    ms_ev = e0.elapsed_time(e1) / args.steps
    ms_per_step = wall * 1e3 / args.steps
    traced = pipe.ray_count()
    st = pipe.read_back_execution_states()
    out = buf.cpu().numpy()
    rec_hits = out[: n * 16].view(np.uint32).reshape(n, 4)
    img = out[n * 16:].reshape(n, 4)
    plain = []
    for _ in range(args.steps):
        system.trace(rays, hits, flags=PE.CULL_BACK)
        system.sync()
        plain.append(system.last_trace_ms())
    plain_ms = float(np.median(plain))
    ph = hits.cpu().numpy().view(np.uint32)
    hit = ph[:, 0] != PE.MISS
    same_hits = bool(np.array_equal(rec_hits[:, 0], ph[:, 0]) and np.array_equal(rec_hits[hit], ph[hit]))
    pipe.close()
    value = traced / (ms_per_step * 1e-3) / 1e6
    config = {
        "workload": "pipeline: the reference's pipeline test (ray gen, closest hit, miss, RGBA8 target) at "
                    "1920x1080 on 64 spheres/tori at 256x256 (8.39M tris), caller stage kernels, round hint 8",
        "rays_per_call": traced, "ms_events": round(ms_ev, 4), "launch_shape": tune,
        "plain_trace_ms": round(plain_ms, 4), "pipeline_over_plain_trace": round(ms_per_step / plain_ms, 2),
        "drained": st.is_empty() and not any(st.errors),
        "image_is_stage_function_of_hits": bool(np.array_equal(img, PE.debug_image(rec_hits))),
        "hits_equal_plain_trace": same_hits,
    }
    system.close()
    return _record(value, "Mrays/s", args, ms_per_step, config, None, {"ms_per_call": round(ms_per_step, 4)})


def cpu_baseline_ao(scene, cam, W, H, first, rays):
    """The oracle's AO (reference traversal, all host threads) on the same
    frame at sample 0: its rays/s (counting the rays the GPU traced for that
    sample), and the fraction of pixels where the GPU's first sample
    differs."""
    import numpy as np

    sys.path.insert(0, str(Path(__file__).resolve().parent))
    from tests.pt_util import oracle_scene_from

    o = oracle_scene_from(scene)
    vp = cam.view_projection_inv().T.reshape(-1)
    cp = [float(np.float32(x)) for x in cam.position]
    t0 = time.perf_counter()
    vals = o.ao_render(W, H, 0, vp, cp)
    dt = time.perf_counter() - t0
    diff = float((first.reshape(H, W) != vals).mean())
    return ({"value": round(rays / dt / 1e6, 3), "unit": "Mrays/s", "cores": _host_threads(), "kind": "port",
             "sample": f"whole {W}x{H} frame, sample 0"}, {"differing_pixels": round(diff, 6)})


def _pmc_pass(workload: str, counters, extra=()):
    """One rocprofv3 --pmc child run of this benchmark in --profile-only mode
    (after a 256 MiB device copy, --copy-calibration); returns the rows of
    its counter CSV, or None when rocprofv3 is unavailable or the pass
    fails or runs past its time limit."""
    import csv
    import shutil
    import signal
    import tempfile

    exe = shutil.which("rocprofv3") or "/opt/rocm/bin/rocprofv3"
    if not os.path.exists(exe):
        return None
    d = tempfile.mkdtemp(prefix="rdrt_pmc_", dir="/tmp")
    cmd = [exe, "--pmc", *counters, "-f", "csv", "-d", d, "-o", "p", "--", sys.executable, str(ROOT / "bench.py"),
           "--workload", workload, "--steps", str(PMC_STEPS), "--warmup", "0", "--no-cpu-baseline", "--no-traffic",
           "--copy-calibration", "--profile-only", *extra]
    env = dict(os.environ, TMPDIR="/tmp")
    try:
        proc = subprocess.Popen(cmd, cwd="/tmp", env=env, stdout=subprocess.DEVNULL, stderr=subprocess.DEVNULL,
                                start_new_session=True)
        try:
            proc.wait(timeout=150)
        except subprocess.TimeoutExpired:
            os.killpg(proc.pid, signal.SIGKILL)
            proc.wait()
            return None
        rows = []
        for root, _, files in os.walk(d):
            for f in files:
                if f.endswith("counter_collection.csv"):
                    rows += list(csv.DictReader(open(os.path.join(root, f))))
        return rows
    finally:
        shutil.rmtree(d, ignore_errors=True)


def pmc_counters(workload: str, kernels, extra=()):
    """Physical DRAM-side bytes and the L2 (TCC) hit rate of the dispatches
    whose names contain one of `kernels`, per step: three child passes
    (FETCH_SIZE; WRITE_SIZE; TCC_HIT_sum + TCC_MISS_sum).  FETCH_SIZE and
    WRITE_SIZE (KiB per dispatch) are each scaled by what the counter reads
    for the 256 MiB copy of the same run.  Returns a dict, or None."""
    out = {}
    for counter in ("FETCH_SIZE", "WRITE_SIZE"):
        rows = _pmc_pass(workload, [counter], extra)
        if not rows:
            return None
        copy = [float(r["Counter_Value"]) for r in rows if "copyBuffer" in r["Kernel_Name"]
                and float(r["Counter_Value"]) > 2 ** 16]
        vals = [float(r["Counter_Value"]) for r in rows if any(k in r["Kernel_Name"] for k in kernels)]
        if not copy or not vals:
            return None
        factor = 262144.0 / copy[0]  # KiB of the 256 MiB copy / what the counter read
        out[counter] = {"kib_per_step": sum(vals) / PMC_STEPS, "copy_factor": factor, "dispatches": len(vals)}
    total = sum(v["kib_per_step"] * v["copy_factor"] * 1024.0 for v in out.values())
    res = {"bytes_per_step": total, "fetch_kib": out["FETCH_SIZE"]["kib_per_step"],
           "fetch_copy_factor": out["FETCH_SIZE"]["copy_factor"], "write_kib": out["WRITE_SIZE"]["kib_per_step"],
           "write_copy_factor": out["WRITE_SIZE"]["copy_factor"],
           "dispatches_per_step": out["FETCH_SIZE"]["dispatches"] / PMC_STEPS}
    rows = _pmc_pass(workload, ["TCC_HIT_sum", "TCC_MISS_sum"], extra)
    if rows:
        sel = [r for r in rows if any(k in r["Kernel_Name"] for k in kernels)]
        hit = sum(float(r["Counter_Value"]) for r in sel if r.get("Counter_Name", "").startswith("TCC_HIT"))
        miss = sum(float(r["Counter_Value"]) for r in sel if r.get("Counter_Name", "").startswith("TCC_MISS"))
        if hit + miss > 0:
            res["l2_hit_rate"] = round(hit / (hit + miss), 4)
            res["l2_requests_per_step"] = (hit + miss) / PMC_STEPS
    return res


TRACE_KERNELS = ("k_trace_fast", "k_trace_persist", "k_trace_coop")
CULL_KERNELS = ("k_trace_cull", "k_block_scan", "k_scatter2")


def _counter_kernels(workload: str, shape) -> tuple:
    if workload == "batched":  # the whole call
        return TRACE_KERNELS + CULL_KERNELS + ("k_spawn", "k_batch_init", "k_zero_work")
    if workload == "closest_hit" and shape and shape.get("cull"):
        return TRACE_KERNELS + CULL_KERNELS
    return TRACE_KERNELS


def _fractions(workload: str, rec: dict) -> dict:
    """HBM fractions of the traversal of one step: over the reference tree's
    visits, over the kernel's own visits, and over the physical DRAM bytes
    (rocprof), with the L2 hit rate of the same dispatches."""
    roof, cfg = rec.get("roofline") or {}, rec.get("config", {})
    if workload in ("diff", "diffgeo"):
        t = rec.get("forward_traversal", {})
        ms, ref, ko = t.get("trace_ms"), t.get("reference_order", {}).get("frac"), t.get("kernel_order", {})
    elif workload == "pt":
        ms, ref, ko = roof.get("trace_ms_per_frame"), roof.get("frac"), roof.get("kernel_order", {})
    elif workload == "closest_hit":
        ms, ref, ko = roof.get("kernel_ms"), roof.get("frac"), roof.get("kernel_order", {})
    elif workload == "batched":
        ms, ref, ko = cfg.get("ms_persistent_events"), roof.get("frac"), roof.get("kernel_order", {})
    elif workload in ("ao", "pipeline"):
        # no visit counts on these paths: the traversal's physical bytes over
        # the whole sample / call (its traversal time is not split out)
        ms, ref, ko = rec.get("ms_per_sample", rec.get("ms_per_call")), None, {}
    else:
        return {}
    tr = rec.get("traffic_detail")
    phys = None
    if tr and ms:
        phys = round(tr["bytes_per_step"] / (ms * 1e-3) / 1e9 / HBM_PEAK_GBPS, 4)
    out = {"time_ms": ms, "reference_order": ref, "kernel_order": ko.get("frac"), "physical": phys,
           "l2_hit_rate": tr.get("l2_hit_rate") if tr else None}
    if workload in ("ao", "pipeline"):
        out["time_basis"] = "whole sample" if workload == "ao" else "whole call"
    return out


WORKLOADS = ("pt", "closest_hit", "batched", "diff", "diffgeo", "ao", "pipeline")
COUNTED = ("pt", "closest_hit", "batched", "diff", "diffgeo", "ao", "pipeline")
TRAFFIC_BUDGET_S = 420.0  # no further counter passes past this point of the run


def run_workload(name: str, args, t_start: float):
    rec = {"pt": path_tracing, "closest_hit": closest_hit, "batched": batched, "diff": differentiable,
           "diffgeo": differentiable_geometry, "ao": ambient_occlusion, "pipeline": pipeline}[name](args)
    if rec is None:
        return None
    traffic = None
    if not args.no_traffic and name in COUNTED:
        if time.perf_counter() - t_start < TRAFFIC_BUDGET_S:
            # counter passes in child processes, with this run's launch shape
            shape = rec["config"].get("launch_shape")
            extra = ("--tune", _tune_arg(shape)) if shape else ()
            traffic = pmc_counters(name, _counter_kernels(name, shape), extra)
        else:
            rec["traffic_skipped"] = "time budget of the run"
    if traffic is not None:
        rec["traffic_detail"] = traffic
        if rec.get("roofline") and name != "diff":
            rec["roofline"]["traffic"] = round(traffic["bytes_per_step"])
    if name in COUNTED:
        rec["fractions"] = _fractions(name, rec)
    return rec


_SUB_DROP = ("metric", "n_gpus", "steps", "warmup", "scaling", "vs_baseline", "data", "dtype")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="all", choices=["all", *WORKLOADS])
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="first run one 256 MiB device-to-device copy (a reference dispatch for HBM counters)")
    ap.add_argument("--no-traffic", action="store_true", help="skip the rocprofv3 --pmc traffic passes")
    ap.add_argument("--no-autotune", action="store_true", help="keep the default launch shape")
    ap.add_argument("--tune", default="",
                    help="launch shape block,lds,waves,persistent[,cull[,wide[,coop[,tail]]]] (no autotune)")
    ap.add_argument("--profile-only", action="store_true", help="run the timed launches only (counter passes)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU only: this build has no multi-GPU path (see the sharding design)",
              file=sys.stderr)
        sys.exit(2)
    if args.copy_calibration:
        import torch

        src = torch.ones(64 << 20, dtype=torch.float32, device="cuda:0")
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
    t_start = time.perf_counter()
    if args.workload != "all":
        rec = run_workload(args.workload, args, t_start)
        if rec is not None:
            print(json.dumps(rec))
        return
    # the whole metric: the path-traced frame is the headline; the closest-hit,
    # batched, differentiable and AO workloads ride along as sub-records
    rec = run_workload("pt", args, t_start)
    subs = {}
    for name in WORKLOADS[1:]:
        print(f"[bench] {name} ({time.perf_counter() - t_start:.0f} s into the run)", file=sys.stderr, flush=True)
        try:
            r = run_workload(name, args, t_start)
        except FileNotFoundError as e:  # the pipeline's stage library was not built
            subs[name] = {"skipped": str(e)}
            continue
        subs[name] = {k: v for k, v in r.items() if k not in _SUB_DROP}
    rec["sub"] = subs
    rec["run_s"] = round(time.perf_counter() - t_start, 1)
    print(json.dumps(rec))


if __name__ == "__main__":
    main()
