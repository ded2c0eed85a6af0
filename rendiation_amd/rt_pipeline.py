"""Programmable ray-tracing pipelines (the C-ABI of ``rdrt_pipeline_*``).

Mirrors the reference's pipeline API:

* ``GPURaytracingPipelineAndBindingSource`` (shader/ray-tracing/src/api/
  pipeline.rs:75-130) collects the stages: ``register_ray_gen``,
  ``register_ray_closest_hit``, ``register_ray_miss`` return shader handles.
  A stage is a HIP kernel ``__global__ void f(rdrt_stage_args)`` compiled
  for gfx950 into any code object loaded in the process; it is registered
  by its host-side handle (the address of ``f``, as an int).
* ``ShaderBindingTable`` (wavefront_compute/sbt.rs:12-48) maps hit-group
  indices ``ray_ty_idx + geometry_idx * ray_type_count + tlas_offset`` to
  closest-hit stages and miss indices to miss stages.
* ``RayTracingPipeline.trace_ray`` is ``RayTracingEncoderProvider::trace_ray``
  (api/backend.rs:48-56): one ray-gen task per launch entry, run to
  completion on the device task graph.
"""

from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field
from typing import Dict, List, Optional

import numpy as np

from . import _lib as L

MAX_STAGES_PER_KIND = 4
MAX_PAYLOAD_WORDS = 16
HIT_ATTRIBUTES = 1
HIT_TRANSFORMS = 2


@dataclass
class GPURaytracingPipelineAndBindingSource:
    """Stages plus the executor's sizing.  ``payload_words``: 32-bit words of
    every payload (1..16); ``max_in_flight``: trace tasks in flight per
    launch entry; ``execution_round_hint``: task-graph rounds per tile (a
    chain of k traces needs 2k + 1); ``tile_pixels``: launch entries per
    tile (0: 2^22; the reference uses 512 * 512); ``hit_context``:
    HIT_ATTRIBUTES / HIT_TRANSFORMS per-hit data the stages read;
    ``use_graph``: capture a call into a hipGraph and replay it."""

    payload_words: int = 1
    max_in_flight: int = 2
    execution_round_hint: int = 8
    tile_pixels: int = 0
    hit_context: int = 0
    use_graph: bool = True
    ray_gen: List[int] = field(default_factory=list)
    closest_hit: List[int] = field(default_factory=list)
    miss: List[int] = field(default_factory=list)

    def register_ray_gen(self, stage: int) -> int:
        self.ray_gen.append(int(stage))
        return len(self.ray_gen) - 1

    def register_ray_closest_hit(self, stage: int) -> int:
        self.closest_hit.append(int(stage))
        return len(self.closest_hit) - 1

    def register_ray_miss(self, stage: int) -> int:
        self.miss.append(int(stage))
        return len(self.miss) - 1


class ShaderBindingTable:
    """Which registered stage runs for the ray gen, each hit group and each
    miss index (``create_sbt(mesh_count, max_geometry_count_per_mesh,
    ray_type_count)``)."""

    def __init__(self, ray_type_count: int = 1):
        self.ray_type_count = ray_type_count
        self.ray_gen: Optional[int] = None
        self.hit_groups: Dict[int, int] = {}
        self.misses: Dict[int, int] = {}

    def config_ray_generation(self, handle: int) -> None:
        self.ray_gen = handle

    def config_hit_group(self, geometry_idx: int, tlas_offset: int, ray_ty_idx: int,
                         closest_hit: Optional[int]) -> None:
        idx = ray_ty_idx + geometry_idx * self.ray_type_count + tlas_offset
        if closest_hit is None:
            self.hit_groups.pop(idx, None)
        else:
            self.hit_groups[idx] = closest_hit

    def config_missing(self, ray_ty_idx: int, handle: int) -> None:
        self.misses[ray_ty_idx] = handle


@dataclass
class PipelineStates:
    wake_counts: List[int]
    sleep_or_finished_counts: List[int]
    empty_counts: List[int]
    errors: List[int]

    def is_empty(self) -> bool:
        return not any(self.wake_counts) and not any(self.sleep_or_finished_counts)


class RayTracingPipeline:
    """A compiled pipeline on a system (its context, stream and bound TLAS)."""

    def __init__(self, system, source: GPURaytracingPipelineAndBindingSource, sbt: ShaderBindingTable):
        if sbt.ray_gen is None:
            raise ValueError("the binding table has no ray-gen stage")
        nclosest = max(sbt.hit_groups) + 1 if sbt.hit_groups else 0
        nmiss = max(sbt.misses) + 1 if sbt.misses else 0
        if nclosest > MAX_STAGES_PER_KIND or nmiss > MAX_STAGES_PER_KIND:
            raise ValueError("at most 4 hit groups and 4 miss indices")
        d = L.RdrtPipelineDesc()
        d.ray_gen = source.ray_gen[sbt.ray_gen]
        for k, h in sbt.hit_groups.items():
            d.closest_hit[k] = source.closest_hit[h]
        for k, h in sbt.misses.items():
            d.miss[k] = source.miss[h]
        d.nclosest, d.nmiss = nclosest, nmiss
        d.payload_words = source.payload_words
        d.max_in_flight = source.max_in_flight
        d.round_hint = source.execution_round_hint
        d.tile_pixels = source.tile_pixels
        d.hit_context = source.hit_context
        d.use_graph = 1 if source.use_graph else 0
        self.system = system
        self.ngroups = 2 + nclosest + nmiss
        h = C.c_void_p()
        L.check(system._L.rdrt_pipeline_create(system._h, C.byref(d), C.byref(h)))
        self._h = h

    def close(self) -> None:
        if self._h:
            self.system._L.rdrt_pipeline_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def trace_ray(self, width: int, height: int, user: int = 0) -> None:
        """Enqueue the launch on the system's stream; ``user`` is handed to
        every stage (usually a device pointer)."""
        L.check(self.system._L.rdrt_pipeline_trace_rays(self._h, width, height, C.c_void_p(int(user))))

    def read_back_execution_states(self) -> PipelineStates:
        arrs = [np.zeros(self.ngroups, dtype=np.uint32) for _ in range(4)]
        L.check(self.system._L.rdrt_pipeline_states(self._h, *[a.ctypes.data for a in arrs]))
        return PipelineStates(*[a.tolist() for a in arrs])

    def ray_count(self) -> int:
        v = C.c_uint64()
        L.check(self.system._L.rdrt_pipeline_ray_count(self._h, C.byref(v)))
        return int(v.value)
