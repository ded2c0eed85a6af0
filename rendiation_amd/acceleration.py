"""Host mirror of the reference's acceleration-structure provider.

Names, argument meaning and error behaviour follow
``GPUAccelerationStructureSystemProvider`` and the naive SAH-BVH system
(shader/ray-tracing/src/api/backend.rs:104-177,
shader/ray-tracing/src/backend/wavefront_compute/geometry/naive/mod.rs:496-609):

* ``create_bottom_level_acceleration_structure(sources) -> BlasHandle``
* ``create_top_level_acceleration_structure(instances) -> TlasHandle``
* ``delete_bottom_level_acceleration_structure`` / ``delete_top_level_acceleration_structure``
* ``bind_tlas(handles)``, ``bind_tlas_max_len()``
* ``create_comp_instance()`` builds and uploads (the reference builds lazily
  there, naive/mod.rs:521-567).

Tracing takes device tensors (torch is only the allocator here) and calls
the HIP kernels through the C-ABI; nothing runs on the CPU.
"""

from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field
from typing import Optional, Sequence

import numpy as np

from . import _lib as L

# ray flags (shader/ray-tracing/src/api/ty.rs:102-114)
RAY_FLAG_NONE = 0x00
RAY_FLAG_FORCE_OPAQUE = 0x01
RAY_FLAG_FORCE_NON_OPAQUE = 0x02
RAY_FLAG_ACCEPT_FIRST_HIT_AND_END_SEARCH = 0x04
RAY_FLAG_SKIP_CLOSEST_HIT_SHADER = 0x08
RAY_FLAG_CULL_BACK_FACING_TRIANGLES = 0x10
RAY_FLAG_CULL_FRONT_FACING_TRIANGLES = 0x20
RAY_FLAG_CULL_OPAQUE = 0x40
RAY_FLAG_CULL_NON_OPAQUE = 0x80
RAY_FLAG_SKIP_TRIANGLES = 0x100
# ty.rs:139-159
HIT_KIND_FRONT_FACING_TRIANGLE = 0xFE
HIT_KIND_BACK_FACING_TRIANGLE = 0xFF
GEOMETRY_INSTANCE_TRIANGLE_FACING_CULL_DISABLE = 0x1
GEOMETRY_INSTANCE_TRIANGLE_FLIP_FACING = 0x2
GEOMETRY_INSTANCE_FORCE_OPAQUE = 0x4
GEOMETRY_INSTANCE_FORCE_NO_OPAQUE = 0x8
GEOMETRY_FLAG_OPAQUE = 0x1
GEOMETRY_FLAG_NO_DUPLICATE_ANYHIT_INVOCATION = 0x2
ANYHIT_ACCEPT = 0x1  # any-hit behaviours (rdrt_blas_set_anyhit)
ANYHIT_END_SEARCH = 0x2

TRAVERSAL_FAST = 0
TRAVERSAL_REFERENCE = 1
BATCH_PERSISTENT = 0
BATCH_SEQUENTIAL = 1
MISS = 0xFFFFFFFF

RAY_DTYPE = np.dtype([("o", "<f4", 3), ("tmin", "<f4"), ("d", "<f4", 3), ("tmax", "<f4")])
HIT_DTYPE = np.dtype([("prim_id", "<u4"), ("t", "<f4"), ("u", "<f4"), ("v", "<f4")])


@dataclass(frozen=True)
class BlasHandle:
    index: int


@dataclass(frozen=True)
class TlasHandle:
    index: int


@dataclass
class TriangleGeometry:
    """``BottomLevelAccelerationStructureBuildBuffer::Triangles`` plus its flags."""

    positions: np.ndarray  # (n, 3) float32
    indices: Optional[np.ndarray] = None  # (m, 3) or (3m,) uint32 or uint16; None = non-indexed
    flags: int = GEOMETRY_FLAG_OPAQUE


# The reference name of one BLAS source entry.
BottomLevelAccelerationStructureBuildSource = TriangleGeometry


@dataclass
class TopLevelAccelerationStructureSourceInstance:
    """One TLAS instance; ``transform`` is a 4x4 matrix in math order
    (``transform[row, col]``), stored column-major as the reference's Mat4."""

    transform: np.ndarray
    acceleration_structure_handle: BlasHandle
    instance_custom_index: int = 0
    mask: int = 0xFFFFFFFF
    instance_shader_binding_table_record_offset: int = 0
    flags: int = 0


@dataclass
class BuildOptions:
    """TreeBuildOption for BLAS / TLAS and the SAH bin count
    (content/space/src/utils.rs:192-210); defaults are the reference's."""

    blas_bin_size: int = 2
    blas_max_depth: int = 50
    tlas_bin_size: int = 10
    tlas_max_depth: int = 50
    sah_bins: int = 4


@dataclass
class SceneInfo:
    tri_nodes: int
    tlas_nodes: int
    tris: int
    verts: int
    instances: int
    geometries: int
    blas: int
    tlas: int
    max_blas_depth: int
    max_tlas_depth: int
    stack_entries: int
    quad_stack_entries: int
    device_bytes: int
    build_ms: float
    tri_quads: int = 0
    tlas_quads: int = 0
    build_phase_ms: tuple = ()


@dataclass
class BatchStats:
    rounds: int
    rays: list  # rays traced per round
    launches: int
    ms: float


def _ptr(t) -> Optional[int]:
    if t is None:
        return None
    return int(t.data_ptr())


class AccelerationStructureSystem:
    """MI355X implementation of the naive SAH-BVH acceleration-structure
    provider.  ``device=None`` makes a host-only system (BVH build and
    inspection, no tracing)."""

    def __init__(self, device: Optional[int] = 0, options: BuildOptions = BuildOptions(),
                 traversal: int = TRAVERSAL_FAST, stream: Optional[int] = None, block_size: int = 0):
        self._L = L.lib()
        cfg = L.RdrtConfig()
        L.check(self._L.rdrt_config_default(C.byref(cfg)))
        cfg.device = -1 if device is None else int(device)
        cfg.stream = stream
        cfg.build = L.RdrtBuildOpts(options.blas_bin_size, options.blas_max_depth, options.tlas_bin_size,
                                    options.tlas_max_depth, options.sah_bins)
        cfg.traversal = traversal
        cfg.block_size = block_size
        h = C.c_void_p()
        L.check(self._L.rdrt_ctx_create(C.byref(cfg), C.byref(h)))
        self._h = h
        self.device = device
        self.stream = stream  # HIP stream handle the work is enqueued on (None: a stream of the context)

    def close(self) -> None:
        if getattr(self, "_h", None):
            self._L.rdrt_ctx_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ---- GPUAccelerationStructureSystemProvider ----
    def create_bottom_level_acceleration_structure(self, sources: Sequence[TriangleGeometry]) -> BlasHandle:
        L.check(self._L.rdrt_blas_begin(self._h))
        keep = []
        for s in sources:
            pos = np.ascontiguousarray(s.positions, dtype=np.float32).reshape(-1, 3)
            keep.append(pos)
            if s.indices is None:
                L.check(self._L.rdrt_blas_add_triangles(self._h, pos.ctypes.data, pos.shape[0], None,
                                                         pos.shape[0] // 3, s.flags))
            elif np.asarray(s.indices).dtype == np.uint16:  # the 16-bit index format
                idx = np.ascontiguousarray(s.indices, dtype=np.uint16).reshape(-1)
                keep.append(idx)
                L.check(self._L.rdrt_blas_add_triangles_u16(self._h, pos.ctypes.data, pos.shape[0], idx.ctypes.data,
                                                             idx.size // 3, s.flags))
            else:
                idx = np.ascontiguousarray(s.indices, dtype=np.uint32).reshape(-1)
                keep.append(idx)
                L.check(self._L.rdrt_blas_add_triangles(self._h, pos.ctypes.data, pos.shape[0], idx.ctypes.data,
                                                         idx.size // 3, s.flags))
        out = L.u32()
        L.check(self._L.rdrt_blas_end(self._h, C.byref(out)))
        return BlasHandle(out.value)

    def delete_bottom_level_acceleration_structure(self, handle: BlasHandle) -> None:
        L.check(self._L.rdrt_blas_destroy(self._h, handle.index))

    def create_top_level_acceleration_structure(
            self, instances: Sequence[TopLevelAccelerationStructureSourceInstance]) -> TlasHandle:
        arr = (L.RdrtInstance * max(1, len(instances)))()
        for i, s in enumerate(instances):
            m = np.asarray(s.transform, dtype=np.float32).reshape(4, 4)
            col_major = m.T.reshape(-1)  # column-major storage
            arr[i].transform[:] = [float(x) for x in col_major]
            arr[i].instance_custom_index = s.instance_custom_index
            arr[i].mask = s.mask & 0xFFFFFFFF
            arr[i].sbt_offset = s.instance_shader_binding_table_record_offset
            arr[i].flags = s.flags
            arr[i].blas = s.acceleration_structure_handle.index
        out = L.u32()
        L.check(self._L.rdrt_tlas_create(self._h, arr, len(instances), C.byref(out)))
        return TlasHandle(out.value)

    def delete_top_level_acceleration_structure(self, handle: TlasHandle) -> None:
        L.check(self._L.rdrt_tlas_destroy(self._h, handle.index))

    def bind_tlas(self, tlas: Sequence[TlasHandle]) -> None:
        arr = (L.u32 * max(1, len(tlas)))(*[t.index for t in tlas])
        L.check(self._L.rdrt_bind_tlas(self._h, arr, len(tlas)))

    def bind_tlas_max_len(self) -> int:
        return int(self._L.rdrt_bind_tlas_max_len(self._h))

    def build(self) -> None:
        """Host build of every BVH (no device work)."""
        L.check(self._L.rdrt_build(self._h))

    def create_comp_instance(self) -> None:
        """Build if needed and upload to the device."""
        L.check(self._L.rdrt_commit(self._h))

    commit = create_comp_instance

    def scene_info(self) -> SceneInfo:
        info = L.RdrtSceneInfo()
        L.check(self._L.rdrt_get_scene_info(self._h, C.byref(info)))
        return SceneInfo(info.tri_nodes, info.tlas_nodes, info.tris, info.verts, info.instances, info.geometries,
                         info.blas, info.tlas, info.max_blas_depth, info.max_tlas_depth, info.stack_entries,
                         info.quad_stack_entries, info.device_bytes, info.build_ms, info.tri_quads, info.tlas_quads,
                         tuple(round(float(x), 1) for x in info.build_phase_ms))

    def copy_built(self, which: str) -> np.ndarray:
        """Built arrays: 'blas_forest' / 'tlas_forest' (48 B nodes as (n, 12)
        uint32), 'redirect', 'instance_order' (uint32), 'blas_quads' (128 B
        child-quad nodes as (n, 32) uint32), 'blas_qquads' (64 B quantized
        twins as (n, 16) uint32)."""
        info = self.scene_info()
        table = {"blas_forest": (0, info.tri_nodes * 12, 12), "tlas_forest": (1, info.tlas_nodes * 12, 12),
                 "redirect": (2, info.tris, 0), "instance_order": (3, info.instances, 0),
                 "blas_quads": (4, info.tri_quads * 32, 32), "blas_qquads": (5, info.tri_quads * 16, 16)}
        code, n, width = table[which]
        out = np.zeros(max(1, n), dtype=np.uint32)
        L.check(self._L.rdrt_copy_built(self._h, code, out.ctypes.data, out.nbytes))
        out = out[:n]
        return out.reshape(-1, width) if width else out

    # ---- shading attributes ----
    def set_vertex_normals(self, blas: BlasHandle, geometry: int, normals: np.ndarray) -> None:
        n = np.ascontiguousarray(normals, dtype=np.float32).reshape(-1, 3)
        L.check(self._L.rdrt_blas_set_vertex_normals(self._h, blas.index, geometry, n.ctypes.data, n.shape[0]))

    def pick(self, blas: BlasHandle, geometry: int, rays: np.ndarray, face_side: int = 2):
        """Nearest hits of host rays (n, 6: origin, direction) against one
        geometry on the host (the reference's picking tests).  Returns
        (primitive uint32 (MISS when none), distance f32, position (n, 3))."""
        r = np.ascontiguousarray(rays, dtype=np.float32).reshape(-1, 6)
        out = np.zeros((max(1, r.shape[0]), 6), dtype=np.uint32)
        L.check(self._L.rdrt_pick(self._h, blas.index, geometry, r.ctypes.data, r.shape[0], face_side,
                                  out.ctypes.data))
        out = out[:r.shape[0]]
        return out[:, 0].copy(), out[:, 1].view(np.float32).copy(), out[:, 2:5].view(np.float32).copy()

    def pick_all(self, blas: BlasHandle, geometry: int, ray, face_side: int = 2):
        """Every hit of one host ray (6 floats) in the reference's visit
        order: (primitive uint32, distance f32, position (k, 3))."""
        r = np.ascontiguousarray(ray, dtype=np.float32).reshape(6)
        cap = 64
        while True:
            out = np.zeros((cap, 6), dtype=np.uint32)
            cnt = C.c_uint32(0)
            L.check(self._L.rdrt_pick_all(self._h, blas.index, geometry, r.ctypes.data, face_side, out.ctypes.data,
                                          cap, C.byref(cnt)))
            if cnt.value <= cap:
                break
            cap = cnt.value
        out = out[:cnt.value]
        return out[:, 0].copy(), out[:, 1].view(np.float32).copy(), out[:, 2:5].view(np.float32).copy()

    def set_vertex_uvs(self, blas: BlasHandle, geometry: int, uvs: np.ndarray) -> None:
        """Per-vertex texture coordinates (nverts, 2) of one geometry."""
        u = np.ascontiguousarray(uvs, dtype=np.float32).reshape(-1)
        L.check(self._L.rdrt_blas_set_vertex_uvs(self._h, blas.index, geometry, u.ctypes.data, u.size // 2))

    def hit_attributes(self, rays, hits, hit_meta, out, *, ray_meta=None, tlas_idx: int = 0, sbt_offset: int = 0,
                       sbt_stride: int = 1) -> None:
        """Shading attributes of traced closest hits into ``out`` (device
        tensor of n * 12 32-bit words: position, shading normal, geometric
        normal, uv, hit-group record)."""
        n = rays.numel() // 8
        _check_dev(rays, n * 8, "rays")
        _check_dev(hits, n * 4, "hits")
        _check_dev(hit_meta, n * 4, "hit_meta")
        _check_dev(out, n * 12, "out")
        if ray_meta is not None:
            _check_dev(ray_meta, n * 4, "ray_meta")
        L.check(self._L.rdrt_fetch_hit_attributes(self._h, _ptr(rays), _ptr(ray_meta), _ptr(hits), _ptr(hit_meta), n,
                                                  tlas_idx, sbt_offset, sbt_stride, _ptr(out)))

    def hit_transforms(self, rays, hits, hit_meta, out, *, ray_meta=None, tlas_idx: int = 0) -> None:
        """Hit context of traced closest hits into ``out`` (device tensor of
        n * 40 32-bit words: object-to-world 16, world-to-object 16,
        object-space ray origin 3 and direction 3, instance hit-group offset,
        pad)."""
        n = rays.numel() // 8
        _check_dev(rays, n * 8, "rays")
        _check_dev(hits, n * 4, "hits")
        _check_dev(hit_meta, n * 4, "hit_meta")
        _check_dev(out, n * 40, "out")
        if ray_meta is not None:
            _check_dev(ray_meta, n * 4, "ray_meta")
        L.check(self._L.rdrt_fetch_hit_transforms(self._h, _ptr(rays), _ptr(ray_meta), _ptr(hits), _ptr(hit_meta), n,
                                                  tlas_idx, _ptr(out)))

    def set_anyhit(self, blas: BlasHandle, geometry: int, behavior: np.ndarray) -> None:
        """Per-triangle any-hit behaviour (ANYHIT_ACCEPT | ANYHIT_END_SEARCH,
        0 = ignore) of hits on the geometry while it is not opaque for the
        ray; takes effect at the next create_comp_instance."""
        b = np.ascontiguousarray(behavior, dtype=np.uint8).reshape(-1)
        L.check(self._L.rdrt_blas_set_anyhit(self._h, blas.index, geometry, b.ctypes.data, b.size))

    def set_triangle_materials(self, blas: BlasHandle, geometry: int, offsets: np.ndarray) -> None:
        o = np.ascontiguousarray(offsets, dtype=np.uint32).reshape(-1)
        L.check(self._L.rdrt_blas_set_triangle_materials(self._h, blas.index, geometry, o.ctypes.data, o.size))

    def set_materials(self, albedo: np.ndarray, emission: np.ndarray) -> None:
        """Lambertian materials: albedo (n, 3), emission (n, 3)."""
        a = np.asarray(albedo, dtype=np.float32).reshape(-1, 3)
        e = np.asarray(emission, dtype=np.float32).reshape(-1, 3)
        arr = (L.RdrtMaterial * max(1, a.shape[0]))()
        for i in range(a.shape[0]):
            arr[i].albedo[:] = [float(x) for x in a[i]]
            arr[i].emission[:] = [float(x) for x in e[i]]
        L.check(self._L.rdrt_set_materials(self._h, arr, a.shape[0]))

    # ---- device primitives ----
    def compact(self, mask, out_idx, count) -> None:
        """Stable compaction of the indices where ``mask`` (device u32) is
        non-zero into ``out_idx``; ``count`` (device, 1 word) gets their number."""
        n = mask.numel()
        L.check(self._L.rdrt_compact(self._h, _ptr(mask), n, _ptr(out_idx), _ptr(count)))

    def sample_sobol(self, index, scramble, out) -> None:
        """out[2i] = van_der_corput(index[i], scramble[2i]); out[2i+1] = sobol(index[i], scramble[2i+1])."""
        L.check(self._L.rdrt_sample_sobol(self._h, _ptr(index), _ptr(scramble), index.numel(), _ptr(out)))

    # ---- tracing ----
    def trace(self, rays, hits, *, ray_meta=None, hit_meta=None, counters=None, flags: int = 0,
              cull_mask: int = 0xFFFFFFFF, tlas_idx: int = 0, traversal: int = TRAVERSAL_FAST) -> None:
        """Enqueue one closest-hit launch.  ``rays`` is a contiguous device
        tensor of n*8 float32 (rdrt_ray records), ``hits`` n*4 32-bit words;
        optional ``ray_meta`` / ``hit_meta`` n*4 words, ``counters`` n*3."""
        n = rays.numel() // 8
        _check_dev(rays, n * 8, "rays")
        _check_dev(hits, n * 4, "hits")
        if ray_meta is not None:
            _check_dev(ray_meta, n * 4, "ray_meta")
        if hit_meta is not None:
            _check_dev(hit_meta, n * 4, "hit_meta")
        if counters is not None:
            _check_dev(counters, n * 3, "counters")
        a = L.RdrtTraceArgs()
        a.rays, a.ray_meta, a.hits = _ptr(rays), _ptr(ray_meta), _ptr(hits)
        a.hit_meta, a.counters = _ptr(hit_meta), _ptr(counters)
        a.n, a.flags, a.cull_mask, a.tlas_idx, a.traversal = n, flags, cull_mask & 0xFFFFFFFF, tlas_idx, traversal
        L.check(self._L.rdrt_trace(self._h, C.byref(a)))

    def trace_batched(self, rays, ray_meta, hits, *, bounces: int, seed: int = 0, spawned=None,
                      scene_offsets=None, mode: int = BATCH_PERSISTENT) -> None:
        """Enqueue a batched multi-scene trace of ``bounces + 1`` rounds.

        ``rays`` (n*8 f32) and ``ray_meta`` (n*4 words, tlas_idx = scene)
        hold the primary rays of all scenes, grouped by scene; ``hits`` gets
        (bounces+1)*n records, round r at [r*n, (r+1)*n); ``spawned``
        (optional, bounces*n*8 f32) the rays of rounds 1.. .  Sequential mode
        needs ``scene_offsets`` (nscenes + 1 host ints)."""
        n = rays.numel() // 8
        _check_dev(rays, n * 8, "rays")
        _check_dev(ray_meta, n * 4, "ray_meta")
        _check_dev(hits, (bounces + 1) * n * 4, "hits")
        if spawned is not None:
            _check_dev(spawned, bounces * n * 8, "spawned")
        a = L.RdrtBatchParams()
        a.rays, a.ray_meta, a.hits, a.spawned = _ptr(rays), _ptr(ray_meta), _ptr(hits), _ptr(spawned)
        offs = None
        if scene_offsets is not None:
            offs = (L.u32 * len(scene_offsets))(*[int(x) for x in scene_offsets])
            a.scene_offsets = C.cast(offs, C.POINTER(L.u32))
            a.nscenes = len(scene_offsets) - 1
        a.n, a.bounces, a.seed, a.mode = n, bounces, seed & 0xFFFFFFFF, mode
        L.check(self._L.rdrt_trace_batched(self._h, C.byref(a)))

    def batch_autotune(self, rays, ray_meta, hits, *, bounces: int, seed: int = 0, spawned=None,
                       reps: int = 2) -> dict:
        """Time every launch shape on whole persistent batched calls of these
        arguments (outputs overwritten) and keep the fastest; returns the
        choice, its ``ms`` the call time."""
        n = rays.numel() // 8
        _check_dev(rays, n * 8, "rays")
        _check_dev(ray_meta, n * 4, "ray_meta")
        _check_dev(hits, (bounces + 1) * n * 4, "hits")
        if spawned is not None:
            _check_dev(spawned, bounces * n * 8, "spawned")
        a = L.RdrtBatchParams()
        a.rays, a.ray_meta, a.hits, a.spawned = _ptr(rays), _ptr(ray_meta), _ptr(hits), _ptr(spawned)
        a.n, a.bounces, a.seed, a.mode = n, bounces, seed & 0xFFFFFFFF, BATCH_PERSISTENT
        t = L.RdrtTune()
        L.check(self._L.rdrt_batch_autotune(self._h, C.byref(a), reps, C.byref(t)))
        return self.tune()

    def batch_stats(self) -> BatchStats:
        """Counts and timing of the last batched trace (synchronizes)."""
        o = L.RdrtBatchStats()
        L.check(self._L.rdrt_batch_get_stats(self._h, C.byref(o)))
        return BatchStats(rounds=o.rounds, rays=list(o.rays)[:o.rounds], launches=o.launches, ms=float(o.ms))

    def set_device_build(self, enable: bool = True) -> None:
        """Build the BLAS trees on the device at the next build / commit."""
        L.check(self._L.rdrt_set_device_build(self._h, 1 if enable else 0))

    def refit(self, positions, normals=None) -> None:
        """Refit every BVH to new vertex positions (device tensor, every vertex
        of the committed scene in the scene vertex order, 3 f32 each) and
        optionally new vertex normals: same BLAS trees with recomputed boxes
        and triangle records, TLAS rebuilt.  Synchronizes."""
        info = self.scene_info()
        _check_dev(positions, info.verts * 3, "positions")
        if normals is not None:
            _check_dev(normals, info.verts * 3, "normals")
        L.check(self._L.rdrt_refit(self._h, _ptr(positions), _ptr(normals)))

    def geometry_vertex_range(self, blas: BlasHandle, geometry: int = 0):
        """(first, count) of a geometry's vertices in the scene vertex order."""
        first, count = L.u32(), L.u32()
        L.check(self._L.rdrt_geometry_vertex_range(self._h, blas.index, geometry, C.byref(first), C.byref(count)))
        return int(first.value), int(count.value)

    def autotune(self, rays, hits, *, ray_meta=None, flags: int = 0, cull_mask: int = 0xFFFFFFFF, tlas_idx: int = 0,
                 reps: int = 3) -> dict:
        """Time every traversal launch shape on a sample launch and keep the
        fastest for this system; returns the choice."""
        n = rays.numel() // 8
        _check_dev(rays, n * 8, "rays")
        _check_dev(hits, n * 4, "hits")
        a = L.RdrtTraceArgs()
        a.rays, a.ray_meta, a.hits = _ptr(rays), _ptr(ray_meta), _ptr(hits)
        a.n, a.flags, a.cull_mask, a.tlas_idx, a.traversal = n, flags, cull_mask & 0xFFFFFFFF, tlas_idx, TRAVERSAL_FAST
        t = L.RdrtTune()
        L.check(self._L.rdrt_autotune(self._h, C.byref(a), reps, C.byref(t)))
        return self.tune()

    def tune(self) -> dict:
        t = L.RdrtTune()
        L.check(self._L.rdrt_get_tune(self._h, C.byref(t)))
        return {"block_size": t.block_size, "lds_entries": t.lds_entries, "waves": t.waves,
                "persistent": t.persistent, "cull": t.cull, "wide": t.wide, "ms": float(t.ms), "coop": t.coop,
                "tail": t.tail}

    def set_tune(self, block_size: int = 256, lds_entries: int = 16, waves: int = 0, persistent: int = 0,
                 cull: int = 0, wide: int = 0, coop: int = 0, tail: int = 0) -> None:
        t = L.RdrtTune(block_size, lds_entries, waves, persistent, cull, wide, 0.0, coop, tail)
        L.check(self._L.rdrt_set_tune(self._h, C.byref(t)))

    def sync(self) -> None:
        L.check(self._L.rdrt_sync(self._h))

    def last_trace_ms(self) -> float:
        ms = L.f32()
        L.check(self._L.rdrt_last_trace_ms(self._h, C.byref(ms)))
        return float(ms.value)


def _check_dev(t, numel: int, name: str) -> None:
    if not getattr(t, "is_cuda", False):
        raise ValueError(f"{name} must be a device tensor")
    if not t.is_contiguous() or t.numel() != numel or t.element_size() != 4:
        raise ValueError(f"{name} must be contiguous with {numel} 32-bit elements")


def trace_closest(system: AccelerationStructureSystem, rays, *, with_meta: bool = False,
                  with_counters: bool = False, **kw):
    """Allocate outputs (torch, same device as ``rays``), trace, and return
    ``(hits[n,4] int32 view, hit_meta or None, counters or None)``."""
    import torch

    n = rays.numel() // 8
    hits = torch.empty((n, 4), dtype=torch.int32, device=rays.device)
    hm = torch.empty((n, 4), dtype=torch.int32, device=rays.device) if with_meta else None
    cnt = torch.empty((n, 3), dtype=torch.int32, device=rays.device) if with_counters else None
    system.trace(rays, hits, hit_meta=hm, counters=cnt, **kw)
    return hits, hm, cnt
