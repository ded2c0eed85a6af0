"""Benchmark of the ray-query hot path on one MI355X.

python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload closest_hit]

closest_hit: 100k random triangles (binned SAH, 4 bins, leaf <= 4), 1M rays
(origins U[-0.5,1.5]^3, directions uniform on the sphere, t in [1e-4, inf)),
seed 0, FP32.  One step = one closest-hit launch over all rays, inputs
resident in HBM.  Prints one JSON line.

roofline: algorithmic bytes per ray = 32 (ray) + 16 (hit) + 48 * node visits
+ 52 * triangle tests + 176 * instance tests, with the visit counts of the
reference traversal order (counted by the reference-order kernel on this
input), divided by the average duration of the timed traversal launches.
cpu_baseline: the CPU oracle's reference traversal on this host's cores.
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
    # reference-order visit counts of this input (algorithmic bytes)
    system.trace(rays, hits, counters=cnt, traversal=TRAVERSAL_REFERENCE)
    system.sync()
    c = cnt.cpu().numpy().astype(np.int64).sum(axis=0)
    bytes_per_launch = 48 * n + 48 * int(c[0]) + 52 * int(c[1]) + 176 * int(c[2])

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

    rec = {
        "metric": METRIC,
        "value": round(value, 3),
        "unit": "Mrays/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": "closest_hit: 100k random triangles (SAH 4 bins, leaf<=4), 1M random rays, seed 0",
            "tris": synth.CLOSEST_HIT_TRIS, "rays": n, "leaf": synth.CLOSEST_HIT_LEAF,
            "bvh_nodes": info.tri_nodes, "bvh_depth": info.max_blas_depth, "traversal": "fast",
            "hit_fraction": round(hit_rate, 4), "build_upload_ms": round(build_upload_ms, 2),
        },
        "roofline": {
            "bound": "hbm",
            "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBPS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBPS, 4),
            "traffic": None,
            "kernel": "k_trace_fast",
            "kernel_ms": round(kernel_ms, 4),
            "bytes_per_launch": bytes_per_launch,
            "visits_per_ray": [round(float(x) / n, 3) for x in c],
        },
    }
    if not args.no_cpu_baseline:
        rec["cpu_baseline"] = cpu_baseline_closest_hit(pos, rays_np, args)
    system.close()
    return rec


def cpu_baseline_closest_hit(pos, rays_np, args):
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
        "value": round(reps * rays_np.shape[0] / dt / 1e6, 3),
        "unit": "Mrays/s",
        "cores": threads,
        "kind": "port",
        "sample": f"full workload ({rays_np.shape[0]} rays) x{reps}, OpenMP dynamic,1024 over rays",
        "single_thread_mrays": round(one.shape[0] / dt1 / 1e6, 4),
        "cpu_model": _cpu_model(),
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="closest_hit", choices=["closest_hit"])
    ap.add_argument("--no-cpu-baseline", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU only: this build has no multi-GPU path (see the sharding design)",
              file=sys.stderr)
        sys.exit(2)
    rec = closest_hit(args)
    print(json.dumps(rec))


if __name__ == "__main__":
    main()
