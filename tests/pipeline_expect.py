"""What the test stages of tests/stages/pipeline_stages.hip produce, restated
in FP32 numpy with the kernels' operation order, from closest-hit records
(of the oracle, or the device's own).  Shared by the CPU and GPU pipeline
tests and by the fixture script."""

from __future__ import annotations

import numpy as np

MISS = 0xFFFFFFFF
CULL_BACK = 0x10
DEPTH = 3
F = np.float32


CHAIN_ORIGIN = (0.0137, -0.0219, 0.0)


def camera_rays(w: int, h: int, tmin: float, tmax: float, origin=(0.0, 0.0, 0.0)) -> np.ndarray:
    """(w * h, 8) rays of the stages' pinhole camera, row-major launch ids
    (test.rs:51-57: origin 0, direction ((x + .5) / w * 2 - 1, 1 - (y + .5) /
    h * 2, -1), normalized)."""
    x = np.tile(np.arange(w, dtype=F), h)
    y = np.repeat(np.arange(h, dtype=F), w)
    fx = (x + F(0.5)) / F(w) * F(2.0) - F(1.0)
    fy = F(1.0) - (y + F(0.5)) / F(h) * F(2.0)
    fz = np.full_like(fx, F(-1.0))
    ln = np.sqrt(fx * fx + fy * fy + fz * fz)
    r = np.zeros((w * h, 8), dtype=F)
    r[:, 0:3] = np.asarray(origin, dtype=F)
    r[:, 3] = F(tmin)
    r[:, 4] = fx / ln
    r[:, 5] = fy / ln
    r[:, 6] = fz / ln
    r[:, 7] = F(tmax)
    return r


def unorm8(x: np.ndarray) -> np.ndarray:
    """FP32 -> 8-bit UNORM: clamp, * 255, round to nearest even."""
    x = np.clip(np.asarray(x, dtype=F), F(0.0), F(1.0))
    return np.rint(x * F(255.0)).astype(np.uint8)


def debug_image(hits: np.ndarray) -> np.ndarray:
    """RGBA8 pixels of the debug pipeline from each pixel's closest-hit
    record (n, 4) u32: (u, v, 150, 255) on a hit, (0, 0, 150, 255) on a miss
    (ray gen 50/255, the closest hit's barycentrics, + 100/255)."""
    h = np.ascontiguousarray(hits).view(np.uint32).reshape(-1, 4)
    hf = h.view(F)
    hit = h[:, 0] != MISS
    img = np.zeros((h.shape[0], 4), dtype=np.uint8)
    b50 = unorm8(F(50.0) / F(255.0))
    img[:, 2] = unorm8(F(b50) / F(255.0) + F(100.0) / F(255.0))
    img[:, 3] = 255
    # the continuation rereads red and green through the format: idempotent
    img[hit, 0] = unorm8(unorm8(hf[hit, 2]).astype(F) / F(255.0))
    img[hit, 1] = unorm8(unorm8(hf[hit, 3]).astype(F) / F(255.0))
    return img


def mirror_rays(rays: np.ndarray, attrs: np.ndarray) -> np.ndarray:
    """The chain stages' next ray from a hit: origin position + 1e-3 *
    geometric normal, direction d - 2 (d . n) n, range [0, 100]."""
    r = np.ascontiguousarray(rays, dtype=F).reshape(-1, 8)
    a = np.ascontiguousarray(attrs).view(np.uint32).reshape(-1, 12).view(F)
    pos, n = a[:, 0:3], a[:, 6:9]
    d = r[:, 4:7]
    dn = d[:, 0] * n[:, 0] + d[:, 1] * n[:, 1] + d[:, 2] * n[:, 2]
    out = np.zeros_like(r)
    out[:, 0:3] = pos + F(1e-3) * n
    out[:, 4:7] = d - (F(2.0) * dn)[:, None] * n
    out[:, 7] = F(100.0)
    return out


def chain_payload(hits_by_depth: np.ndarray) -> np.ndarray:
    """Final (n, 3) payload of the chain pipeline from the (n, DEPTH, 4) hit
    records of each pixel's rays (unused depths: any value): word 0 sums
    prim_id + 1 over the hits plus 0x10000 if the chain ended in a miss,
    word 1 counts the hits, word 2 is the depth of the last ray."""
    h = np.ascontiguousarray(hits_by_depth).view(np.uint32).reshape(-1, DEPTH, 4)
    n = h.shape[0]
    out = np.zeros((n, 3), dtype=np.uint64)
    live = np.ones(n, dtype=bool)
    for d in range(DEPTH):
        prim = h[:, d, 0].astype(np.uint64)
        hit = live & (h[:, d, 0] != MISS)
        miss = live & (h[:, d, 0] == MISS)
        out[hit, 0] += prim[hit] + 1
        out[hit, 1] += 1
        out[miss, 0] += 0x10000
        out[live, 2] = d
        live = hit
    return (out & np.uint64(0xFFFFFFFF)).astype(np.uint32)


def chain_depths(hits_by_depth: np.ndarray) -> np.ndarray:
    """Rays traced per pixel by the chain pipeline (1..DEPTH)."""
    h = np.ascontiguousarray(hits_by_depth).view(np.uint32).reshape(-1, DEPTH, 4)
    cnt = np.ones(h.shape[0], dtype=np.int64)
    live = h[:, 0, 0] != MISS
    for d in range(1, DEPTH):
        cnt += live
        live = live & (h[:, d, 0] != MISS)
    return cnt
