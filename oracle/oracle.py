"""ctypes wrapper of the CPU oracle (liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, by the smoke check and by the
cpu_baseline leg of bench.py, never by the product package.
"""

from __future__ import annotations

import ctypes as C
import subprocess
from pathlib import Path
from typing import Optional

import numpy as np

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "liboracle.so"

u32 = C.c_uint32
f32 = C.c_float
P = C.c_void_p


class OrcBuildOpts(C.Structure):
    _fields_ = [("blas_bin_size", u32), ("blas_max_depth", u32), ("tlas_bin_size", u32),
                ("tlas_max_depth", u32), ("sah_bins", u32)]


class OrcSizes(C.Structure):
    _fields_ = [(n, u32) for n in ("tri_nodes", "tlas_nodes", "tris", "verts", "instances", "geometries", "blas",
                                   "tlas", "max_blas_depth", "max_tlas_depth")]


class OrcAoParams(C.Structure):
    _fields_ = [("width", u32), ("height", u32), ("sample_index", u32), ("pad", u32),
                ("view_projection_inv", f32 * 16), ("camera_position", f32 * 3), ("pad2", f32)]


class OrcInstance(C.Structure):
    _fields_ = [("transform", f32 * 16), ("instance_custom_index", u32), ("mask", u32), ("sbt_offset", u32),
                ("flags", u32), ("blas", u32), ("pad", u32 * 3)]


class OrcPtParams(C.Structure):
    _fields_ = [("view_projection_inv", f32 * 16), ("camera_position", f32 * 3), ("width", u32), ("height", u32),
                ("spp", u32), ("max_bounces", u32), ("rr_start", u32), ("seed", u32), ("sample_index_base", u32),
                ("sky_up", f32 * 3), ("sky_down", f32 * 3), ("sun_direction", f32 * 3), ("sun_radiance", f32 * 3),
                ("use_graph", u32), ("diagnostics", u32), ("record", u32), ("pad", u32),
                ("views", u32), ("pad2", u32), ("view_params", C.c_void_p)]


_SIG = {
    "orc_scene_new": (P, [C.POINTER(OrcBuildOpts)]),
    "orc_scene_free": (None, [P]),
    "orc_blas_begin": (None, [P]),
    "orc_blas_add_triangles": (None, [P, P, u32, P, u32, u32]),
    "orc_blas_end": (u32, [P]),
    "orc_blas_delete": (None, [P, u32]),
    "orc_tlas_create": (u32, [P, C.POINTER(OrcInstance), u32]),
    "orc_tlas_delete": (None, [P, u32]),
    "orc_bind_tlas": (None, [P, P, u32]),
    "orc_build": (C.c_int, [P]),
    "orc_get_sizes": (None, [P, C.POINTER(OrcSizes)]),
    "orc_copy_tri_forest": (None, [P, P]),
    "orc_copy_tlas_forest": (None, [P, P]),
    "orc_copy_indices_redirect": (None, [P, P]),
    "orc_copy_instance_order": (None, [P, P]),
    "orc_copy_geometry_meta": (None, [P, P]),
    "orc_copy_tlas_bounding": (None, [P, P]),
    "orc_copy_transform_inv": (None, [P, P]),
    "orc_trace": (C.c_int, [P, P, P, u32, u32, u32, u32, P, P, P, C.c_int]),
    "orc_trace_brute": (C.c_int, [P, P, u32, u32, u32, u32, P, C.c_int]),
    "orc_intersect_triangle": (None, [P, P, f32, f32, P, P, P, C.c_int, C.c_int, P]),
    "orc_intersect_aabb": (C.c_int, [P, P, f32, f32, P, P]),
    "orc_pick_triangle": (f32, [P, P, P, P, P, u32]),
    "orc_pick_box": (C.c_int, [P, P, P, P]),
    "orc_pick": (C.c_int, [P, u32, u32, P, u32, u32, C.c_int, P]),
    "orc_ao_hemisphere_sample": (None, [u32, P]),
    "orc_ao_sampler": (u32, [u32, u32, u32, u32, P, P, P]),
    "orc_ao_render": (C.c_int, [P, P, P, P, P, C.c_int]),
    "orc_pick_all": (C.c_int, [P, u32, u32, P, u32, P, u32, P]),
    "orc_xxhash32": (u32, [u32, u32, u32]),
    "orc_set_vertex_uvs": (C.c_int, [P, u32, u32, P, u32]),
    "orc_hit_attributes": (C.c_int, [P, P, P, P, P, u32, u32, u32, u32, P, P]),
    "orc_camera_ray": (None, [P, u32, u32, u32, u32, P, P]),
    "orc_van_der_corput": (f32, [u32, u32]),
    "orc_sobol": (f32, [u32, u32]),
    "orc_compact": (u32, [P, P, u32, P]),
    "orc_set_vertex_normals": (C.c_int, [P, u32, u32, P, u32]),
    "orc_set_triangle_materials": (C.c_int, [P, u32, u32, P, u32]),
    "orc_set_anyhit": (C.c_int, [P, u32, u32, P, u32]),
    "orc_set_materials": (None, [P, P, u32]),
    "orc_pt_render": (C.c_int, [P, C.POINTER(OrcPtParams), P, P, C.c_int]),
    "orc_pt_render_rows": (C.c_int, [P, C.POINTER(OrcPtParams), u32, u32, P, P, C.c_int]),
    "orc_batch_round": (C.c_int, [P, P, P, u32, u32, u32, P, P, P, C.c_int]),
    "orc_pt_grad": (C.c_int, [P, C.POINTER(OrcPtParams), P, P, P, C.c_int]),
    "orc_vertex_count": (u32, [P]),
    "orc_pt_grad_geometry": (C.c_int, [P, C.POINTER(OrcPtParams), P, P, P, C.c_int]),
    "orc_pt_replay_geometry": (C.c_int, [P, C.POINTER(OrcPtParams), P, P, P, P, C.c_int]),
}

_lib = None


def build_oracle() -> Path:
    subprocess.run(["make", "-s", "-C", str(HERE)], check=True)
    return LIB_PATH


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            build_oracle()
        L = C.CDLL(str(LIB_PATH))
        for name, (res, args) in _SIG.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def _p(a: Optional[np.ndarray]):
    return None if a is None else a.ctypes.data


class OracleScene:
    """The reference NaiveSahBvhSource + NaiveSahBvhCpu, restated in C++."""

    def __init__(self, blas_bin_size=2, blas_max_depth=50, tlas_bin_size=10, tlas_max_depth=50, sah_bins=4):
        self.L = lib()
        opts = OrcBuildOpts(blas_bin_size, blas_max_depth, tlas_bin_size, tlas_max_depth, sah_bins)
        self.h = self.L.orc_scene_new(C.byref(opts))
        self._keep = []

    def __del__(self):
        try:
            if self.h:
                self.L.orc_scene_free(self.h)
        except Exception:
            pass

    def add_blas(self, geometries) -> int:
        """geometries: list of (positions (n,3) f32, indices (m,3) u32 or None, flags)."""
        self.L.orc_blas_begin(self.h)
        for pos, idx, flags in geometries:
            pos = np.ascontiguousarray(pos, dtype=np.float32).reshape(-1, 3)
            if idx is None:
                self.L.orc_blas_add_triangles(self.h, pos.ctypes.data, pos.shape[0], None, pos.shape[0] // 3, flags)
            else:
                idx = np.ascontiguousarray(idx, dtype=np.uint32).reshape(-1)
                self.L.orc_blas_add_triangles(self.h, pos.ctypes.data, pos.shape[0], idx.ctypes.data, idx.size // 3,
                                              flags)
        return int(self.L.orc_blas_end(self.h))

    def delete_blas(self, h: int) -> None:
        self.L.orc_blas_delete(self.h, h)

    def add_tlas(self, instances) -> int:
        """instances: list of dicts with transform (4x4 math order), blas, and
        optional custom, mask, sbt, flags."""
        arr = (OrcInstance * max(1, len(instances)))()
        for i, d in enumerate(instances):
            m = np.asarray(d["transform"], dtype=np.float32).reshape(4, 4).T.reshape(-1)
            arr[i].transform[:] = [float(x) for x in m]
            arr[i].instance_custom_index = d.get("custom", 0)
            arr[i].mask = d.get("mask", 0xFFFFFFFF)
            arr[i].sbt_offset = d.get("sbt", 0)
            arr[i].flags = d.get("flags", 0)
            arr[i].blas = d["blas"]
        return int(self.L.orc_tlas_create(self.h, arr, len(instances)))

    def delete_tlas(self, h: int) -> None:
        self.L.orc_tlas_delete(self.h, h)

    def bind(self, handles) -> None:
        a = np.asarray(handles, dtype=np.uint32)
        self.L.orc_bind_tlas(self.h, a.ctypes.data, a.size)

    def build(self) -> None:
        rc = self.L.orc_build(self.h)
        if rc != 0:
            raise RuntimeError(f"oracle build failed ({rc})")

    def sizes(self) -> OrcSizes:
        s = OrcSizes()
        self.L.orc_get_sizes(self.h, C.byref(s))
        return s

    def tri_forest(self) -> np.ndarray:
        n = self.sizes().tri_nodes
        out = np.zeros((max(1, n), 12), dtype=np.uint32)
        self.L.orc_copy_tri_forest(self.h, out.ctypes.data)
        return out[:n]

    def tlas_forest(self) -> np.ndarray:
        n = self.sizes().tlas_nodes
        out = np.zeros((max(1, n), 12), dtype=np.uint32)
        self.L.orc_copy_tlas_forest(self.h, out.ctypes.data)
        return out[:n]

    def redirect(self) -> np.ndarray:
        n = self.sizes().tris
        out = np.zeros(max(1, n), dtype=np.uint32)
        self.L.orc_copy_indices_redirect(self.h, out.ctypes.data)
        return out[:n]

    def instance_order(self) -> np.ndarray:
        n = self.sizes().instances
        out = np.zeros(max(1, n), dtype=np.uint32)
        self.L.orc_copy_instance_order(self.h, out.ctypes.data)
        return out[:n]

    def geometry_meta(self) -> np.ndarray:
        n = self.sizes().geometries
        out = np.zeros((max(1, n), 4), dtype=np.uint32)
        self.L.orc_copy_geometry_meta(self.h, out.ctypes.data)
        return out[:n]

    def trace(self, rays: np.ndarray, meta: Optional[np.ndarray] = None, flags=0, cull_mask=0xFFFFFFFF, tlas_idx=0,
              threads: int = 0, with_meta=False, with_counters=False):
        rays = np.ascontiguousarray(rays, dtype=np.float32).reshape(-1, 8)
        n = rays.shape[0]
        hits = np.zeros((n, 4), dtype=np.uint32)
        hm = np.zeros((n, 4), dtype=np.uint32) if with_meta else None
        cnt = np.zeros((n, 3), dtype=np.uint32) if with_counters else None
        if meta is not None:
            meta = np.ascontiguousarray(meta, dtype=np.uint32).reshape(-1, 4)
        rc = self.L.orc_trace(self.h, rays.ctypes.data, _p(meta), n, flags, cull_mask & 0xFFFFFFFF, tlas_idx,
                              hits.ctypes.data, _p(hm), _p(cnt), threads)
        if rc != 0:
            raise RuntimeError("oracle trace failed (scene not built)")
        return hits, hm, cnt

    def batch_round(self, rays: np.ndarray, meta: np.ndarray, round_: int, seed: int, threads: int = 0):
        """One round of batched tracing: (hits (n,4) u32, spawned (n,8) f32,
        alive (n,) u32) -- closest hits, then continuation rays of the hits."""
        rays = np.ascontiguousarray(rays, dtype=np.float32).reshape(-1, 8)
        meta = np.ascontiguousarray(meta, dtype=np.uint32).reshape(-1, 4)
        n = rays.shape[0]
        if meta.shape[0] != n:
            raise ValueError("one metadata record per ray")
        hits = np.zeros((n, 4), dtype=np.uint32)
        spawned = np.zeros((n, 8), dtype=np.float32)
        alive = np.zeros(n, dtype=np.uint32)
        rc = self.L.orc_batch_round(self.h, rays.ctypes.data, meta.ctypes.data, n, round_, seed & 0xFFFFFFFF,
                                    hits.ctypes.data, spawned.ctypes.data, alive.ctypes.data, threads)
        if rc != 0:
            raise RuntimeError("oracle batch round failed (scene not built)")
        return hits, spawned, alive

    def set_vertex_uvs(self, blas: int, geometry: int, uvs) -> None:
        u = np.ascontiguousarray(uvs, dtype=np.float32).reshape(-1, 2)
        if self.L.orc_set_vertex_uvs(self.h, blas, geometry, u.ctypes.data, u.shape[0]) != 0:
            raise ValueError("bad uvs")

    def hit_attributes(self, rays: np.ndarray, hits: np.ndarray, hit_meta: np.ndarray, meta=None, tlas_idx=0,
                       sbt_offset=0, sbt_stride=0):
        """(attributes (n, 12) u32 words, hit context (n, 40) u32 words) of
        closest hits, the bindless-mesh fetch and the hit context of the
        reference's closest-hit stage (float words viewed as u32)."""
        rays = np.ascontiguousarray(rays, dtype=np.float32).reshape(-1, 8)
        n = rays.shape[0]
        hits = np.ascontiguousarray(hits).view(np.uint32).reshape(-1, 4)
        hm = np.ascontiguousarray(hit_meta).view(np.uint32).reshape(-1, 4)
        if meta is not None:
            meta = np.ascontiguousarray(meta).view(np.uint32).reshape(-1, 4)
        attr = np.zeros((n, 12), dtype=np.uint32)
        xf = np.zeros((n, 40), dtype=np.uint32)
        rc = self.L.orc_hit_attributes(self.h, rays.ctypes.data, _p(meta), hits.ctypes.data, hm.ctypes.data, n,
                                       tlas_idx, sbt_offset, sbt_stride, attr.ctypes.data, xf.ctypes.data)
        if rc != 0:
            raise RuntimeError("oracle hit attributes failed (scene not built)")
        return attr, xf

    # ---- path tracer ----
    def set_vertex_normals(self, blas: int, geometry: int, normals) -> None:
        n = np.ascontiguousarray(normals, dtype=np.float32).reshape(-1, 3)
        if self.L.orc_set_vertex_normals(self.h, blas, geometry, n.ctypes.data, n.shape[0]) != 0:
            raise ValueError("bad normals")

    def set_anyhit(self, blas: int, geometry: int, behavior) -> None:
        """Per-triangle any-hit behaviour (bit 0 accept, bit 1 end search) of
        hits on the geometry when it is not opaque for the ray."""
        b = np.ascontiguousarray(behavior, dtype=np.uint8).reshape(-1)
        if self.L.orc_set_anyhit(self.h, blas, geometry, b.ctypes.data, b.size) != 0:
            raise ValueError("bad any-hit table")

    def set_triangle_materials(self, blas: int, geometry: int, offsets) -> None:
        o = np.ascontiguousarray(offsets, dtype=np.uint32).reshape(-1)
        if self.L.orc_set_triangle_materials(self.h, blas, geometry, o.ctypes.data, o.size) != 0:
            raise ValueError("bad material offsets")

    def set_materials(self, albedo, emission) -> None:
        a = np.asarray(albedo, dtype=np.float32).reshape(-1, 3)
        e = np.asarray(emission, dtype=np.float32).reshape(-1, 3)
        m = np.zeros((a.shape[0], 8), dtype=np.float32)
        m[:, 0:3] = a
        m[:, 4:7] = e
        self._mats = m
        self.L.orc_set_materials(self.h, m.ctypes.data, m.shape[0])

    def pt_render(self, params: "OrcPtParams", threads: int = 0):
        """Returns (height, width, 3) float64 radiance and the per-bounce
        (extension, shadow) ray counts."""
        out = np.zeros((max(params.views, 1) * params.height, params.width, 3), dtype=np.float64)
        nb = params.max_bounces + 1
        counts = np.zeros(2 * nb, dtype=np.uint64)
        rc = self.L.orc_pt_render(self.h, C.byref(params), out.ctypes.data, counts.ctypes.data, threads)
        if rc != 0:
            raise RuntimeError("oracle pt_render failed")
        return out, counts[:nb], counts[nb:]

    def pt_render_rows(self, params: "OrcPtParams", row_begin: int, row_end: int, threads: int = 0):
        """Rows [row_begin, row_end) of the frame: ((rows, width, 3) float64, ext counts, shadow counts)."""
        out = np.zeros((row_end - row_begin, params.width, 3), dtype=np.float64)
        nb = params.max_bounces + 1
        counts = np.zeros(2 * nb, dtype=np.uint64)
        rc = self.L.orc_pt_render_rows(self.h, C.byref(params), row_begin, row_end, out.ctypes.data,
                                       counts.ctypes.data, threads)
        if rc != 0:
            raise RuntimeError("oracle pt_render_rows failed")
        return out, counts[:nb], counts[nb:]

    def pt_grad(self, params: "OrcPtParams", adj: np.ndarray, nmaterials: int, threads: int = 0):
        """Gradient of sum(adj * frame) w.r.t. (albedo, emission), each
        (nmaterials, 3) float64, in FP64 with visibility fixed."""
        adj = np.ascontiguousarray(adj, dtype=np.float64).reshape(max(params.views, 1) * params.height,
                                                                   params.width, 3)
        ga = np.zeros((nmaterials, 3), dtype=np.float64)
        ge = np.zeros((nmaterials, 3), dtype=np.float64)
        rc = self.L.orc_pt_grad(self.h, C.byref(params), adj.ctypes.data, ga.ctypes.data, ge.ctypes.data, threads)
        if rc == -2:
            raise ValueError("the gradient needs rr_start >= max_bounces (no Russian roulette)")
        if rc != 0:
            raise RuntimeError("oracle pt_grad failed")
        return ga, ge

    def vertex_count(self) -> int:
        return int(self.L.orc_vertex_count(self.h))

    def _adj(self, params, adj):
        return np.ascontiguousarray(adj, dtype=np.float64).reshape(max(params.views, 1) * params.height,
                                                                   params.width, 3)

    def pt_grad_geometry(self, params: "OrcPtParams", adj: np.ndarray, threads: int = 0):
        """Gradient of sum(adj * frame) w.r.t. (positions, normals), each
        (vertex_count, 3) float64, paths held fixed."""
        adj = self._adj(params, adj)
        nv = self.vertex_count()
        gp = np.zeros((nv, 3), dtype=np.float64)
        gn = np.zeros((nv, 3), dtype=np.float64)
        rc = self.L.orc_pt_grad_geometry(self.h, C.byref(params), adj.ctypes.data, gp.ctypes.data, gn.ctypes.data,
                                         threads)
        if rc == -2:
            raise ValueError("the gradient needs rr_start >= max_bounces (no Russian roulette)")
        if rc != 0:
            raise RuntimeError("oracle pt_grad_geometry failed")
        return gp, gn

    def pt_replay_geometry(self, params: "OrcPtParams", adj: np.ndarray, pos=None, nrm=None, threads: int = 0):
        """sum(adj * frame) of the fixed paths re-shaded with vertex positions
        / normals ((vertex_count, 3) float64, None = the scene's own)."""
        adj = self._adj(params, adj)
        pos = None if pos is None else np.ascontiguousarray(pos, dtype=np.float64)
        nrm = None if nrm is None else np.ascontiguousarray(nrm, dtype=np.float64)
        out = C.c_double()
        rc = self.L.orc_pt_replay_geometry(self.h, C.byref(params), adj.ctypes.data, _p(pos), _p(nrm),
                                           C.byref(out), threads)
        if rc != 0:
            raise RuntimeError("oracle pt_replay_geometry failed")
        return float(out.value)

    def pick(self, blas: int, geometry: int, rays: np.ndarray, side: int = 2, bvh: bool = True):
        """Host picking: (prim uint32, distance f32, position (n, 3) f32)."""
        r = np.ascontiguousarray(rays, dtype=np.float32).reshape(-1, 6)
        out = np.zeros((max(1, r.shape[0]), 6), dtype=np.uint32)
        rc = self.L.orc_pick(self.h, blas, geometry, r.ctypes.data, r.shape[0], side, int(bvh), out.ctypes.data)
        if rc != 0:
            raise RuntimeError("oracle pick failed")
        out = out[:r.shape[0]]
        return out[:, 0].copy(), out[:, 1].view(np.float32).copy(), out[:, 2:5].view(np.float32).copy()

    def pick_all(self, blas: int, geometry: int, ray, side: int = 2):
        r = np.ascontiguousarray(ray, dtype=np.float32).reshape(6)
        cnt = C.c_uint32(0)
        self.L.orc_pick_all(self.h, blas, geometry, r.ctypes.data, side, None, 0, C.byref(cnt))
        out = np.zeros((max(1, cnt.value), 6), dtype=np.uint32)
        rc = self.L.orc_pick_all(self.h, blas, geometry, r.ctypes.data, side, out.ctypes.data, cnt.value, C.byref(cnt))
        if rc != 0:
            raise RuntimeError("oracle pick failed")
        out = out[:cnt.value]
        return out[:, 0].copy(), out[:, 1].view(np.float32).copy(), out[:, 2:5].view(np.float32).copy()

    def ao_render(self, width: int, height: int, sample_index: int, vp_inv_colmajor, cam_pos, accum=None,
                  accum_rgba8=None, threads: int = 0):
        """One AO sample; returns the per-pixel values (0 / 1) and updates the
        running means in place when given."""
        p = OrcAoParams()
        p.width, p.height, p.sample_index = width, height, sample_index
        p.view_projection_inv[:] = [float(x) for x in vp_inv_colmajor]
        p.camera_position[:] = [float(x) for x in cam_pos]
        vals = np.zeros(width * height, dtype=np.float32)
        rc = self.L.orc_ao_render(self.h, C.byref(p), _p(accum), _p(accum_rgba8), vals.ctypes.data, threads)
        if rc != 0:
            raise RuntimeError("oracle AO failed")
        return vals.reshape(height, width)

    def trace_brute(self, rays: np.ndarray, flags=0, cull_mask=0xFFFFFFFF, tlas_idx=0, threads: int = 0):
        rays = np.ascontiguousarray(rays, dtype=np.float32).reshape(-1, 8)
        hits = np.zeros((rays.shape[0], 4), dtype=np.uint32)
        rc = self.L.orc_trace_brute(self.h, rays.ctypes.data, rays.shape[0], flags, cull_mask & 0xFFFFFFFF,
                                    tlas_idx, hits.ctypes.data, threads)
        if rc != 0:
            raise RuntimeError("oracle trace failed (scene not built)")
        return hits


def intersect_triangle(o, d, near, far, v0, v1, v2, cull_enable=False, cull_back=False) -> np.ndarray:
    L = lib()
    a = [np.ascontiguousarray(x, dtype=np.float32) for x in (o, d, v0, v1, v2)]
    out = np.zeros(4, dtype=np.float32)
    L.orc_intersect_triangle(a[0].ctypes.data, a[1].ctypes.data, near, far, a[2].ctypes.data, a[3].ctypes.data,
                             a[4].ctypes.data, int(cull_enable), int(cull_back), out.ctypes.data)
    return out


def intersect_aabb(o, d, near, far, bmin, bmax) -> bool:
    L = lib()
    a = [np.ascontiguousarray(x, dtype=np.float32) for x in (o, d, bmin, bmax)]
    return bool(L.orc_intersect_aabb(a[0].ctypes.data, a[1].ctypes.data, near, far, a[2].ctypes.data,
                                     a[3].ctypes.data))


def pick_triangle(o, d, a, b, c, side=2) -> float:
    """Host GTE ray/triangle test with a face side (0 Front, 1 Back, 2
    Double); NaN when there is no hit."""
    v = [np.ascontiguousarray(x, dtype=np.float32) for x in (o, d, a, b, c)]
    return float(lib().orc_pick_triangle(*[x.ctypes.data for x in v], side))


def pick_box(o, d, bmin, bmax) -> bool:
    v = [np.ascontiguousarray(x, dtype=np.float32) for x in (o, d, bmin, bmax)]
    return bool(lib().orc_pick_box(*[x.ctypes.data for x in v]))


def ao_sampler(x: int, y: int, sample_index: int, n: int):
    """(seed, states, hashes, floats) of n steps of the AO pixel sampler."""
    states = np.zeros(n, dtype=np.uint32)
    bits = np.zeros(n, dtype=np.uint32)
    vals = np.zeros(n, dtype=np.float32)
    seed = lib().orc_ao_sampler(x, y, sample_index, n, states.ctypes.data, bits.ctypes.data, vals.ctypes.data)
    return int(seed), states, bits, vals


def ao_hemisphere_sample(s: int) -> np.ndarray:
    out = np.zeros(3, dtype=np.float32)
    lib().orc_ao_hemisphere_sample(s, out.ctypes.data)
    return out


def xxhash32(x: int, y: int, z: int) -> int:
    return int(lib().orc_xxhash32(x, y, z))


def camera_ray(params: "OrcPtParams", px: int, py: int, sample: int, view: int = 0):
    """(origin, unit direction) float32 of the path tracer's camera ray."""
    o = np.zeros(3, dtype=np.float32)
    d = np.zeros(3, dtype=np.float32)
    lib().orc_camera_ray(C.byref(params), px, py, sample, view, o.ctypes.data, d.ctypes.data)
    return o, d


def van_der_corput(n: int, scramble: int) -> float:
    return float(lib().orc_van_der_corput(n, scramble))


def sobol(n: int, scramble: int) -> float:
    return float(lib().orc_sobol(n, scramble))


def compact(mask, values=None):
    mask = np.ascontiguousarray(mask, dtype=np.uint32)
    vals = None if values is None else np.ascontiguousarray(values, dtype=np.uint32)
    out = np.zeros(max(1, mask.size), dtype=np.uint32)
    k = lib().orc_compact(_p(vals), mask.ctypes.data, mask.size, out.ctypes.data)
    return out[:mask.size], int(k)
