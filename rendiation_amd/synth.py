"""Seeded synthetic inputs for the ray-query workloads.

Every random number comes from a counter-based SplitMix64 stream:
``x_i = mix64(key(seed, stream) + (i + 1) * 0x9E3779B97F4A7C15)`` with the
standard SplitMix64 finaliser, so any slice of any stream is reproducible
without generating the prefix.  A float in [0, 1) is ``(x >> 40) * 2^-24``
(exact in FP32).  The same arrays feed the GPU path and the CPU oracle.

Streams used by the closest-hit workload (100k triangles, 1M rays):
  stream 1: triangle centroids, 3 per triangle, U[0,1)
  stream 2: vertex offsets, 9 per triangle, U[-0.01, 0.01]
  stream 3: ray origins, 3 per ray, U[-0.5, 1.5)
  stream 4: ray directions, 2 per ray, uniform on the unit sphere through
            uniform_sample_sphere_dir (math/statistics/src/distribution_map.rs:61-68)
"""

from __future__ import annotations

import numpy as np

_GAMMA = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)
_STREAM_KEY = np.uint64(0xD1B54A32D192ED03)


def _mix(z: np.ndarray) -> np.ndarray:
    z = (z ^ (z >> np.uint64(30))) * _M1
    z = (z ^ (z >> np.uint64(27))) * _M2
    return z ^ (z >> np.uint64(31))


def splitmix64(seed: int, stream: int, n: int, offset: int = 0) -> np.ndarray:
    """n uint64 values of stream ``stream`` starting at index ``offset``."""
    with np.errstate(over="ignore"):
        key = _mix(np.array([seed], dtype=np.uint64) * _GAMMA + np.uint64(stream) * _STREAM_KEY)[0]
        i = np.arange(offset + 1, offset + n + 1, dtype=np.uint64)
        return _mix(key + i * _GAMMA)


def uniform_f32(seed: int, stream: int, n: int, offset: int = 0) -> np.ndarray:
    """n float32 in [0, 1), 24-bit resolution."""
    x = splitmix64(seed, stream, n, offset) >> np.uint64(40)
    return x.astype(np.float32) * np.float32(2.0 ** -24)


def sphere_dirs(u: np.ndarray, v: np.ndarray) -> np.ndarray:
    """uniform_sample_sphere_dir (distribution_map.rs:61-68) in FP32; cos/sin
    are evaluated in FP64 and rounded, so the result is platform-stable."""
    u = u.astype(np.float32)
    v = v.astype(np.float32)
    z = np.float32(1.0) - np.float32(2.0) * u
    r = np.sqrt(np.maximum(np.float32(0.0), np.float32(1.0) - z * z)).astype(np.float32)
    phi = (np.float32(np.pi) * np.float32(2.0)) * v
    c = np.cos(phi.astype(np.float64)).astype(np.float32)
    s = np.sin(phi.astype(np.float64)).astype(np.float32)
    return np.stack([c * r, s * r, z], axis=-1).astype(np.float32)


def random_triangles(seed: int, ntris: int, spread: float = 0.01) -> np.ndarray:
    """(3*ntris, 3) float32 non-indexed triangle soup: centroid c ~ U[0,1)^3,
    each vertex c + U[-spread, spread]^3."""
    c = uniform_f32(seed, 1, 3 * ntris).reshape(ntris, 1, 3)
    off = (np.float32(2.0) * uniform_f32(seed, 2, 9 * ntris) - np.float32(1.0)) * np.float32(spread)
    v = (c + off.reshape(ntris, 3, 3)).astype(np.float32)
    return v.reshape(-1, 3)


def random_rays(seed: int, nrays: int, tmin: float = 1e-4, tmax: float = np.inf, lo: float = -0.5,
                hi: float = 1.5, offset: int = 0) -> np.ndarray:
    """(nrays, 8) float32 rdrt_ray records: origin ~ U[lo, hi)^3, direction
    uniform on the sphere, range [tmin, tmax].  ``offset`` selects rays
    offset..offset+nrays of the same streams."""
    o = np.float32(lo) + np.float32(hi - lo) * uniform_f32(seed, 3, 3 * nrays, 3 * offset).reshape(nrays, 3)
    uv = uniform_f32(seed, 4, 2 * nrays, 2 * offset).reshape(nrays, 2)
    d = sphere_dirs(uv[:, 0], uv[:, 1])
    out = np.empty((nrays, 8), dtype=np.float32)
    out[:, 0:3] = o
    out[:, 3] = np.float32(tmin)
    out[:, 4:7] = d
    out[:, 7] = np.float32(tmax)
    return out


def parametric_grid_indices(nu: int, nv: int) -> np.ndarray:
    """Triangle indices of triangulate_parametric
    (content/mesh/generator/src/builder/mod.rs:73-125): vertex (u, v) has
    index v + (nv + 1) * u, each cell (a b / c d) gives (a, c, b), (b, c, d)."""
    u = np.arange(nu, dtype=np.uint32)[:, None]
    v = np.arange(nv, dtype=np.uint32)[None, :]
    a = v + (nv + 1) * u
    b = a + 1
    c = a + (nv + 1)
    d = c + 1
    tri = np.stack([np.stack([a, c, b], -1), np.stack([b, c, d], -1)], axis=2)
    return tri.reshape(-1, 3).astype(np.uint32)


def _grid_uv(nu: int, nv: int):
    uu = (np.arange(nu + 1, dtype=np.float32) * np.float32(1.0 / nu))[:, None]
    vv = (np.arange(nv + 1, dtype=np.float32) * np.float32(1.0 / nv))[None, :]
    return np.broadcast_to(uu, (nu + 1, nv + 1)), np.broadcast_to(vv, (nu + 1, nv + 1))


def uv_sphere(nu: int, nv: int, radius: float = 1.0):
    """UV sphere (content/mesh/generator/src/primitive.rs UVSphere):
    (cos 2pi u sin pi v, cos pi v, sin 2pi u sin pi v) * radius; returns
    (positions (nverts, 3) f32, indices (ntris, 3) u32).  The pole rows give
    zero-area triangles, as in the reference generator."""
    u, v = _grid_uv(nu, nv)
    a = 2.0 * np.pi * u.astype(np.float64)
    b = np.pi * v.astype(np.float64)
    p = np.stack([np.cos(a) * np.sin(b), np.cos(b), np.sin(a) * np.sin(b)], -1) * radius
    return p.reshape(-1, 3).astype(np.float32), parametric_grid_indices(nu, nv)


def torus(nu: int, nv: int, radius: float = 1.0, tube_radius: float = 0.2):
    """Torus around the z axis (TorusMeshParameter, builtin.rs:4-21): ring
    radius ``radius`` in the xy plane, tube radius ``tube_radius``."""
    u, v = _grid_uv(nu, nv)
    a = 2.0 * np.pi * u.astype(np.float64)
    b = 2.0 * np.pi * v.astype(np.float64)
    ring = radius + tube_radius * np.cos(b)
    p = np.stack([ring * np.cos(a), ring * np.sin(a), tube_radius * np.sin(b)], -1)
    return p.reshape(-1, 3).astype(np.float32), parametric_grid_indices(nu, nv)


def random_rotation(seed: int, stream: int, index: int) -> np.ndarray:
    """Uniform random 3x3 rotation (quaternion from 4 normals via Box-Muller)."""
    x = uniform_f32(seed, stream, 8, 8 * index).astype(np.float64)
    x = np.clip(x, 1e-12, 1.0)
    r = np.sqrt(-2.0 * np.log(x[0::2]))
    g = r * np.cos(2.0 * np.pi * x[1::2])
    q = g / np.linalg.norm(g)
    w, i, j, k = q
    return np.array([[1 - 2 * (j * j + k * k), 2 * (i * j - k * w), 2 * (i * k + j * w)],
                     [2 * (i * j + k * w), 1 - 2 * (i * i + k * k), 2 * (j * k - i * w)],
                     [2 * (i * k - j * w), 2 * (j * k + i * w), 1 - 2 * (i * i + j * j)]])


def sphere_normals(nu: int, nv: int) -> np.ndarray:
    p, _ = uv_sphere(nu, nv)
    n = p.astype(np.float64)
    n /= np.maximum(np.linalg.norm(n, axis=1, keepdims=True), 1e-30)
    return n.astype(np.float32)


def torus_normals(nu: int, nv: int, radius: float = 1.0, tube_radius: float = 0.2) -> np.ndarray:
    u, v = _grid_uv(nu, nv)
    a = 2.0 * np.pi * u.astype(np.float64)
    b = 2.0 * np.pi * v.astype(np.float64)
    n = np.stack([np.cos(b) * np.cos(a), np.cos(b) * np.sin(a), np.sin(b)], -1)
    return n.reshape(-1, 3).astype(np.float32)


def pt_scene(seed: int = 0, n_meshes: int = 64, tess: int = 256, n_emissive: int = 4, emission: float = 5.0,
             materials_per_mesh: int = 1, extent: float = 10.0) -> dict:
    """Parametric path-tracing scene: ``n_meshes`` tessellated UV spheres
    (even i) and tori (odd i, R = 1, r = 0.2) at tess x tess, each its own
    BLAS, placed by a uniform random rotation, uniform scale U[0.5, 2] and a
    translation U[-extent, extent]^3.  Materials: albedo U[0.1, 0.9]^3,
    meshes 0..n_emissive-1 emit ``emission``.  With materials_per_mesh = m,
    triangles of a mesh get m bands of material offsets along u.
      streams 20 rotation, 21 scale, 22 translation, 23 albedo"""
    sph_p, sph_i = uv_sphere(tess, tess)
    tor_p, tor_i = torus(tess, tess)
    sph_n = sphere_normals(tess, tess)
    tor_n = torus_normals(tess, tess)
    ntri = sph_i.shape[0]
    band = (np.arange(ntri, dtype=np.int64) * materials_per_mesh // ntri).astype(np.uint32)
    meshes, instances = [], []
    nmat = n_meshes * materials_per_mesh
    albedo = (np.float32(0.1) + np.float32(0.8) * uniform_f32(seed, 23, 3 * nmat)).reshape(nmat, 3)
    emit = np.zeros((nmat, 3), dtype=np.float32)
    scale = np.float32(0.5) + np.float32(1.5) * uniform_f32(seed, 21, n_meshes)
    trans = (np.float32(2.0 * extent) * uniform_f32(seed, 22, 3 * n_meshes) - np.float32(extent)).reshape(-1, 3)
    for i in range(n_meshes):
        kind = i % 2
        meshes.append({"positions": sph_p if kind == 0 else tor_p, "indices": sph_i if kind == 0 else tor_i,
                       "normals": sph_n if kind == 0 else tor_n,
                       "tri_material": band if materials_per_mesh > 1 else None})
        m = np.eye(4)
        m[:3, :3] = random_rotation(seed, 20, i) * float(scale[i])
        m[:3, 3] = trans[i]
        instances.append({"transform": m, "mesh": i, "material": i * materials_per_mesh})
        if i < n_emissive:
            emit[i * materials_per_mesh:(i + 1) * materials_per_mesh] = emission
    return {"meshes": meshes, "instances": instances, "albedo": albedo, "emission": emit}


# Closest-hit workload sizes (100k triangles, 1M rays, seed 0).
CLOSEST_HIT_TRIS = 100_000
CLOSEST_HIT_RAYS = 1_000_000
CLOSEST_HIT_LEAF = 4


def log_uniform_counts(seed: int, stream: int, n: int, lo: int, hi: int) -> np.ndarray:
    """n integer counts floor(exp(U[ln lo, ln hi))) (log-uniform in [lo, hi))."""
    u = uniform_f32(seed, stream, n).astype(np.float64)
    c = np.floor(np.exp(np.log(lo) + u * (np.log(hi) - np.log(lo)))).astype(np.int64)
    return np.clip(c, lo, hi - 1)


def batched_scenes(seed: int = 0, nscenes: int = 64, tris=(10_000, 2_000_000), rays=(16_384, 4_194_304),
                   spread: float = 0.01):
    """Independent random-triangle scenes with uneven ray sets: scene k has a
    log-uniform triangle count (stream 30) and ray count (stream 31); its
    triangles and rays come from the closest-hit distributions under seed
    ``seed * 1009 + k + 1``.  Returns (scenes, rays (N, 8) f32, meta (N, 4)
    u32 with tlas_idx = k, offsets (nscenes + 1))."""
    ntri = log_uniform_counts(seed, 30, nscenes, *tris)
    nray = log_uniform_counts(seed, 31, nscenes, *rays)
    offsets = np.zeros(nscenes + 1, dtype=np.int64)
    offsets[1:] = np.cumsum(nray)
    all_rays = np.empty((int(offsets[-1]), 8), dtype=np.float32)
    meta = np.zeros((int(offsets[-1]), 4), dtype=np.uint32)
    meta[:, 1] = 0xFFFFFFFF
    scenes = []
    for k in range(nscenes):
        sk = seed * 1009 + k + 1
        scenes.append(random_triangles(sk, int(ntri[k]), spread))
        all_rays[offsets[k]:offsets[k + 1]] = random_rays(sk, int(nray[k]))
        meta[offsets[k]:offsets[k + 1], 2] = k
    return scenes, all_rays, meta, offsets


# Batched workload: 64 scenes, triangles log-U[10k, 2M), rays log-U[16k, 4M), 3 bounces.
BATCHED_SCENES = 64
BATCHED_TRIS = (10_000, 2_000_000)
BATCHED_RAYS = (16_384, 4_194_304)
BATCHED_BOUNCES = 3


def height_field(n: int = 1000, seed: int = 0, amplitude: float = 0.05, waves: int = 8, extent: float = 1.0):
    """n x n vertex grid over [-extent, extent]^2 in the xz-plane displaced
    along +y by sum_k amplitude * sin(w_k . (x, z) + phi_k), k < waves, with
    w_k = 2 pi f_k (cos t_k, sin t_k), f_k ~ U[0.5, 3], t_k ~ U[0, 2 pi),
    phi_k ~ U[0, 2 pi) (stream 50).  Returns (positions (n*n, 3) f32,
    indices (2 (n-1)^2, 3) u32 facing +y, analytic normals (n*n, 3) f32)."""
    u = uniform_f32(seed, 50, 3 * waves).astype(np.float64).reshape(waves, 3)
    f = 0.5 + 2.5 * u[:, 0]
    th = 2.0 * np.pi * u[:, 1]
    phi = 2.0 * np.pi * u[:, 2]
    wx, wz = 2.0 * np.pi * f * np.cos(th), 2.0 * np.pi * f * np.sin(th)
    g = np.linspace(-extent, extent, n)
    x, z = np.meshgrid(g, g, indexing="xy")
    x, z = x.reshape(-1), z.reshape(-1)
    arg = x[:, None] * wx[None, :] + z[:, None] * wz[None, :] + phi[None, :]
    y = amplitude * np.sin(arg).sum(axis=1)
    dy_dx = amplitude * (np.cos(arg) * wx[None, :]).sum(axis=1)
    dy_dz = amplitude * (np.cos(arg) * wz[None, :]).sum(axis=1)
    pos = np.stack([x, y, z], axis=1).astype(np.float32)
    nrm = np.stack([-dy_dx, np.ones_like(x), -dy_dz], axis=1)
    nrm = (nrm / np.linalg.norm(nrm, axis=1, keepdims=True)).astype(np.float32)
    i = np.arange(n - 1, dtype=np.uint32)
    r, c = np.meshgrid(i, i, indexing="ij")
    v00 = (r * n + c).reshape(-1)
    v01, v10, v11 = v00 + 1, v00 + n, v00 + n + 1
    # (v00, v10, v01): e1 = +z, e2 = +x, e1 x e2 = +y
    idx = np.concatenate([np.stack([v00, v10, v01], axis=1), np.stack([v01, v10, v11], axis=1)]).astype(np.uint32)
    return pos, idx, nrm


def sphere_views(count: int = 16, radius: float = 3.0, min_elevation_deg: float = 20.0):
    """Camera positions on the upper part of a sphere (Fibonacci spiral over
    elevations [min_elevation, 80] degrees), all looking at the origin."""
    k = np.arange(count, dtype=np.float64)
    el = np.radians(min_elevation_deg + (80.0 - min_elevation_deg) * (k + 0.5) / count)
    az = k * np.pi * (3.0 - np.sqrt(5.0))
    return [(float(radius * np.cos(e) * np.cos(a)), float(radius * np.sin(e)), float(radius * np.cos(e) * np.sin(a)))
            for e, a in zip(el, az)]
