"""ctypes binding of librdrt.so (include/rdrt.h).

The library is loaded from this package directory only.  There is no
fallback: if it is missing, importing the bindings raises, telling the user
to build it.
"""

from __future__ import annotations

import ctypes as C
from pathlib import Path

LIB_PATH = Path(__file__).resolve().parent / "librdrt.so"

u32 = C.c_uint32
i32 = C.c_int32
f32 = C.c_float


class RdrtRay(C.Structure):
    _fields_ = [("ox", f32), ("oy", f32), ("oz", f32), ("tmin", f32),
                ("dx", f32), ("dy", f32), ("dz", f32), ("tmax", f32)]


class RdrtTaskGroupDesc(C.Structure):
    _fields_ = [("kind", u32), ("child_group", u32), ("max_in_flight", u32), ("pad", u32)]


class RdrtAoParams(C.Structure):
    _fields_ = [("width", u32), ("height", u32), ("sample_index", u32), ("format", u32),
                ("view_projection_inv", f32 * 16), ("camera_position", f32 * 3), ("rounds", u32),
                ("tile_pixels", u32), ("use_graph", u32), ("pad", u32 * 2)]


class RdrtPipelineDesc(C.Structure):
    _fields_ = [("ray_gen", C.c_void_p), ("closest_hit", C.c_void_p * 4), ("miss", C.c_void_p * 4),
                ("nclosest", u32), ("nmiss", u32), ("payload_words", u32), ("max_in_flight", u32),
                ("round_hint", u32), ("tile_pixels", u32), ("hit_context", u32), ("use_graph", u32)]


class RdrtTraceCall(C.Structure):
    _fields_ = [("origin", f32 * 3), ("tmin", f32), ("direction", f32 * 3), ("tmax", f32),
                ("flags", u32), ("cull_mask", u32), ("tlas_idx", u32), ("miss_index", u32),
                ("sbt_offset", u32), ("sbt_stride", u32), ("pad", u32 * 2)]


class RdrtMetrics(C.Structure):
    _fields_ = [("last_call_ms", f32), ("tune_ms", f32), ("device_bytes", C.c_uint64), ("build_ms", C.c_double),
                ("status", u32), ("pad", u32)]


class RdrtRayMeta(C.Structure):
    _fields_ = [("flags", u32), ("cull_mask", u32), ("tlas_idx", u32), ("pad", u32)]


class RdrtHit(C.Structure):
    _fields_ = [("prim_id", u32), ("t", f32), ("u", f32), ("v", f32)]


class RdrtHitMeta(C.Structure):
    _fields_ = [("geometry_id", u32), ("instance_id", u32), ("hit_kind", u32), ("instance_custom_index", u32)]


class RdrtInstance(C.Structure):
    _fields_ = [("transform", f32 * 16), ("instance_custom_index", u32), ("mask", u32),
                ("sbt_offset", u32), ("flags", u32), ("blas", u32), ("pad", u32 * 3)]


class RdrtBuildOpts(C.Structure):
    _fields_ = [("blas_bin_size", u32), ("blas_max_depth", u32), ("tlas_bin_size", u32),
                ("tlas_max_depth", u32), ("sah_bins", u32)]


class RdrtConfig(C.Structure):
    _fields_ = [("device", i32), ("stream", C.c_void_p), ("build", RdrtBuildOpts),
                ("traversal", u32), ("block_size", u32)]


class RdrtTraceArgs(C.Structure):
    _fields_ = [("rays", C.c_void_p), ("ray_meta", C.c_void_p), ("hits", C.c_void_p),
                ("hit_meta", C.c_void_p), ("counters", C.c_void_p), ("n", u32), ("flags", u32),
                ("cull_mask", u32), ("tlas_idx", u32), ("traversal", u32), ("pad", u32)]


class RdrtSceneInfo(C.Structure):
    _fields_ = [("tri_nodes", u32), ("tlas_nodes", u32), ("tris", u32), ("verts", u32),
                ("instances", u32), ("geometries", u32), ("blas", u32), ("tlas", u32),
                ("max_blas_depth", u32), ("max_tlas_depth", u32), ("stack_entries", u32), ("quad_stack_entries", u32),
                ("device_bytes", C.c_uint64), ("build_ms", C.c_double), ("tri_quads", u32), ("tlas_quads", u32),
                ("build_phase_ms", C.c_double * 4)]


class RdrtMaterial(C.Structure):
    _fields_ = [("albedo", f32 * 3), ("pad0", f32), ("emission", f32 * 3), ("pad1", f32)]


class RdrtPtParams(C.Structure):
    _fields_ = [("view_projection_inv", f32 * 16), ("camera_position", f32 * 3), ("width", u32), ("height", u32),
                ("spp", u32), ("max_bounces", u32), ("rr_start", u32), ("seed", u32), ("sample_index_base", u32),
                ("sky_up", f32 * 3), ("sky_down", f32 * 3), ("sun_direction", f32 * 3), ("sun_radiance", f32 * 3),
                ("use_graph", u32), ("diagnostics", u32), ("record", u32), ("pad", u32),
                ("views", u32), ("pad2", u32), ("view_params", C.c_void_p)]


class RdrtPtStats(C.Structure):
    _fields_ = [("bounces", u32), ("extension_rays", u32 * 17), ("shadow_rays", u32 * 17), ("frame_ms", f32),
                ("trace_ms", f32), ("trace_launches", u32), ("visits", C.c_uint64 * 3),
                ("max_ray_visits", u32 * 33), ("launch_ms", f32 * 33)]


RDRT_PT_DIAG_TIME_TRACES = 0x1
RDRT_FACE_FRONT, RDRT_FACE_BACK, RDRT_FACE_DOUBLE = 0, 1, 2
RDRT_TASK_READY, RDRT_TASK_SPAWN_THEN = 0, 1
RDRT_AO_FORMAT_F32, RDRT_AO_FORMAT_RGBA8 = 0, 1
RDRT_PT_DIAG_COUNT_FAST = 0x4
RDRT_PT_DIAG_COUNT_VISITS = 0x2


class RdrtBatchParams(C.Structure):
    _fields_ = [("rays", C.c_void_p), ("ray_meta", C.c_void_p), ("hits", C.c_void_p), ("spawned", C.c_void_p),
                ("scene_offsets", C.POINTER(u32)), ("n", u32), ("nscenes", u32), ("bounces", u32), ("seed", u32),
                ("mode", u32), ("pad", u32)]


class RdrtBatchStats(C.Structure):
    _fields_ = [("rounds", u32), ("rays", u32 * 17), ("launches", u32), ("ms", f32)]


class RdrtTune(C.Structure):
    _fields_ = [("block_size", u32), ("lds_entries", u32), ("waves", u32), ("persistent", u32), ("cull", u32),
                ("wide", u32), ("ms", f32), ("coop", u32), ("tail", u32)]


RDRT_BATCH_PERSISTENT = 0
RDRT_BATCH_SEQUENTIAL = 1


# name -> (restype, argtypes); the full exported surface of rdrt.h
SIGNATURES = {
    "rdrt_last_error": (C.c_char_p, []),
    "rdrt_abi_version": (i32, []),
    "rdrt_config_default": (i32, [C.POINTER(RdrtConfig)]),
    "rdrt_ctx_create": (i32, [C.POINTER(RdrtConfig), C.POINTER(C.c_void_p)]),
    "rdrt_ctx_destroy": (i32, [C.c_void_p]),
    "rdrt_sync": (i32, [C.c_void_p]),
    "rdrt_blas_begin": (i32, [C.c_void_p]),
    "rdrt_blas_add_triangles": (i32, [C.c_void_p, C.c_void_p, u32, C.c_void_p, u32, u32]),
    "rdrt_blas_end": (i32, [C.c_void_p, C.POINTER(u32)]),
    "rdrt_blas_destroy": (i32, [C.c_void_p, u32]),
    "rdrt_tlas_create": (i32, [C.c_void_p, C.POINTER(RdrtInstance), u32, C.POINTER(u32)]),
    "rdrt_tlas_destroy": (i32, [C.c_void_p, u32]),
    "rdrt_bind_tlas": (i32, [C.c_void_p, C.POINTER(u32), u32]),
    "rdrt_bind_tlas_max_len": (u32, [C.c_void_p]),
    "rdrt_build": (i32, [C.c_void_p]),
    "rdrt_commit": (i32, [C.c_void_p]),
    "rdrt_get_scene_info": (i32, [C.c_void_p, C.POINTER(RdrtSceneInfo)]),
    "rdrt_copy_built": (i32, [C.c_void_p, u32, C.c_void_p, C.c_size_t]),
    "rdrt_trace": (i32, [C.c_void_p, C.POINTER(RdrtTraceArgs)]),
    "rdrt_trace_closest": (i32, [C.c_void_p, C.c_void_p, u32, C.c_void_p]),
    "rdrt_last_trace_ms": (i32, [C.c_void_p, C.POINTER(f32)]),
    "rdrt_blas_set_vertex_normals": (i32, [C.c_void_p, u32, u32, C.c_void_p, u32]),
    "rdrt_blas_set_triangle_materials": (i32, [C.c_void_p, u32, u32, C.c_void_p, u32]),
    "rdrt_set_materials": (i32, [C.c_void_p, C.POINTER(RdrtMaterial), u32]),
    "rdrt_pt_render": (i32, [C.c_void_p, C.POINTER(RdrtPtParams), C.c_void_p]),
    "rdrt_pt_get_stats": (i32, [C.c_void_p, C.POINTER(RdrtPtStats)]),
    "rdrt_pt_set_material_params": (i32, [C.c_void_p, C.c_void_p, C.c_void_p, u32]),
    "rdrt_pt_backward": (i32, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]),
    "rdrt_l2_adjoint": (i32, [C.c_void_p, C.c_void_p, C.c_void_p, u32, C.c_void_p, C.c_void_p]),
    "rdrt_pt_backward_geometry": (i32, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]),
    "rdrt_geometry_vertex_range": (i32, [C.c_void_p, u32, u32, C.POINTER(u32), C.POINTER(u32)]),
    "rdrt_trace_batched": (i32, [C.c_void_p, C.POINTER(RdrtBatchParams)]),
    "rdrt_pt_autotune": (i32, [C.c_void_p, C.POINTER(RdrtPtParams), C.c_void_p, u32, C.POINTER(RdrtTune)]),
    "rdrt_batch_autotune": (i32, [C.c_void_p, C.POINTER(RdrtBatchParams), u32, C.POINTER(RdrtTune)]),
    "rdrt_refit": (i32, [C.c_void_p, C.c_void_p, C.c_void_p]),
    "rdrt_set_device_build": (i32, [C.c_void_p, u32]),
    "rdrt_fetch_hit_attributes": (i32, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, u32, u32, u32, u32,
                                       C.c_void_p]),
    "rdrt_fetch_hit_transforms": (i32, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, u32, u32,
                                       C.c_void_p]),
    "rdrt_blas_set_vertex_uvs": (i32, [C.c_void_p, u32, u32, C.c_void_p, u32]),
    "rdrt_blas_add_triangles_u16": (i32, [C.c_void_p, C.c_void_p, u32, C.c_void_p, u32, u32]),
    "rdrt_pick": (i32, [C.c_void_p, u32, u32, C.c_void_p, u32, u32, C.c_void_p]),
    "rdrt_pick_all": (i32, [C.c_void_p, u32, u32, C.c_void_p, u32, C.c_void_p, u32, C.c_void_p]),
    "rdrt_task_graph_create": (i32, [C.c_void_p, C.c_void_p, u32, u32, C.c_void_p]),
    "rdrt_task_graph_destroy": (i32, [C.c_void_p]),
    "rdrt_task_graph_spawn": (i32, [C.c_void_p, u32, u32]),
    "rdrt_task_graph_execute": (i32, [C.c_void_p, u32]),
    "rdrt_task_graph_states": (i32, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]),
    "rdrt_ao_render": (i32, [C.c_void_p, C.c_void_p, C.c_void_p]),
    "rdrt_ao_states": (i32, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]),
    "rdrt_ao_ray_counts": (i32, [C.c_void_p, C.c_void_p]),
    "rdrt_ao_autotune": (i32, [C.c_void_p, C.c_void_p, C.c_void_p, u32, C.c_void_p]),
    "rdrt_ao_hemisphere_sample": (i32, [u32, C.c_void_p]),
    "rdrt_blas_create": (i32, [C.c_void_p, C.c_void_p, u32, C.c_void_p, u32, u32, C.c_void_p, C.c_void_p]),
    "rdrt_pt_forward": (i32, [C.c_void_p, C.c_void_p, C.c_void_p]),
    "rdrt_sobol": (i32, [u32, u32, u32, C.c_void_p]),
    "rdrt_get_metrics": (i32, [C.c_void_p, C.c_void_p]),
    "rdrt_autotune": (i32, [C.c_void_p, C.POINTER(RdrtTraceArgs), u32, C.POINTER(RdrtTune)]),
    "rdrt_set_tune": (i32, [C.c_void_p, C.POINTER(RdrtTune)]),
    "rdrt_get_tune": (i32, [C.c_void_p, C.POINTER(RdrtTune)]),
    "rdrt_batch_get_stats": (i32, [C.c_void_p, C.POINTER(RdrtBatchStats)]),
    "rdrt_compact": (i32, [C.c_void_p, C.c_void_p, u32, C.c_void_p, C.c_void_p]),
    "rdrt_compact_static": (i32, [C.c_void_p, C.c_void_p, u32, C.c_void_p, C.c_void_p]),
    "rdrt_pipeline_create": (i32, [C.c_void_p, C.POINTER(RdrtPipelineDesc), C.c_void_p]),
    "rdrt_pipeline_destroy": (i32, [C.c_void_p]),
    "rdrt_pipeline_trace_rays": (i32, [C.c_void_p, u32, u32, C.c_void_p]),
    "rdrt_pipeline_states": (i32, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]),
    "rdrt_pipeline_ray_count": (i32, [C.c_void_p, C.POINTER(C.c_uint64)]),
    "rdrt_external_buffer_import": (i32, [C.c_void_p, C.c_int, C.c_uint64, C.c_uint64, C.c_uint64, C.c_void_p,
                                          C.c_void_p]),
    "rdrt_external_buffer_release": (i32, [C.c_void_p]),
    "rdrt_blas_set_anyhit": (i32, [C.c_void_p, u32, u32, C.c_void_p, u32]),
    "rdrt_sample_sobol": (i32, [C.c_void_p, C.c_void_p, C.c_void_p, u32, C.c_void_p]),
}

RDRT_OK = 0
ERROR_NAMES = {
    -1: "RDRT_ERR_INVALID_ARG", -2: "RDRT_ERR_HIP", -3: "RDRT_ERR_NOT_BUILT", -4: "RDRT_ERR_OUT_OF_MEMORY",
    -5: "RDRT_ERR_STATE", -6: "RDRT_ERR_NO_DEVICE", -7: "RDRT_ERR_CAPACITY", -8: "RDRT_ERR_UNSUPPORTED",
}


class RdrtError(RuntimeError):
    """A librdrt call returned a non-zero status."""

    def __init__(self, code: int, message: str):
        self.code = code
        super().__init__(f"{ERROR_NAMES.get(code, code)}: {message}")


_lib = None


def lib() -> C.CDLL:
    """The loaded librdrt.so; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise ImportError(
                f"{LIB_PATH} is missing: build it with `python -m rendiation_amd.build` "
                "(or __graft_entry__.build()); there is no non-native fallback")
        L = C.CDLL(str(LIB_PATH))
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def check(rc: int) -> None:
    if rc != RDRT_OK:
        msg = lib().rdrt_last_error()
        raise RdrtError(rc, msg.decode() if msg else "")
