"""Device task graph and ambient occlusion on it (the C-ABI of
``rdrt_task_graph_*`` and ``rdrt_ao_*``).

``DeviceTaskGraphBuildSource`` / ``DeviceTaskGraphExecutor`` mirror the
reference's task-graph API (shader/task-graph/src/runtime/mod.rs:55-460)
for the two generic task kinds: a task that resolves on its first poll
(``BaseShaderFuture``) and one that spawns a task of another (or its own)
group and awaits it (``then`` + ``TaskFuture``).

``AORenderer`` mirrors ``SceneRayTracingAORenderer``
(scene/rendering/gpu-ray-tracing/src/feature/ao.rs:121-315): each
``render`` adds one sample to a running mean per pixel and advances the
sample count; ``reset_sample`` starts over.
"""

from __future__ import annotations

import ctypes as C
from dataclasses import dataclass
from typing import List

import numpy as np

from . import _lib as L

READY = L.RDRT_TASK_READY
SPAWN_THEN = L.RDRT_TASK_SPAWN_THEN
MAX_SAMPLE = 256


@dataclass
class TaskGraphExecutionStates:
    wake_counts: List[int]
    sleep_or_finished_counts: List[int]
    empty_counts: List[int]
    errors: List[int]

    def is_empty(self) -> bool:
        return all(c == 0 for c in self.sleep_or_finished_counts) and all(c == 0 for c in self.wake_counts)


class DeviceTaskGraphBuildSource:
    def __init__(self):
        self.tasks = []
        self.capacity = 0

    def next_task_idx(self) -> int:
        return len(self.tasks)

    def define_task(self, kind: int, max_in_flight: int, child_group: int = 0) -> int:
        """``kind`` READY or SPAWN_THEN (spawning into ``child_group``)."""
        self.tasks.append((kind, child_group, max_in_flight))
        return len(self.tasks) - 1

    def build(self, system) -> "DeviceTaskGraphExecutor":
        return DeviceTaskGraphExecutor(system, self)


class DeviceTaskGraphExecutor:
    def __init__(self, system, source: DeviceTaskGraphBuildSource):
        self.system = system
        n = len(source.tasks)
        descs = (L.RdrtTaskGroupDesc * max(1, n))()
        for i, (kind, child, mif) in enumerate(source.tasks):
            descs[i].kind, descs[i].child_group, descs[i].max_in_flight = kind, child, mif
        h = C.c_void_p()
        L.check(system._L.rdrt_task_graph_create(system._h, C.cast(descs, C.c_void_p), n, source.capacity,
                                                 C.byref(h)))
        self._h = h
        self.n = n

    def close(self):
        if self._h:
            self.system._L.rdrt_task_graph_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def dispatch_allocate_init_task(self, count: int, task_id: int) -> None:
        L.check(self.system._L.rdrt_task_graph_spawn(self._h, task_id, count))

    def execute(self, rounds: int) -> None:
        L.check(self.system._L.rdrt_task_graph_execute(self._h, rounds))

    def read_back_execution_states(self) -> TaskGraphExecutionStates:
        arrs = [np.zeros(self.n, dtype=np.uint32) for _ in range(4)]
        L.check(self.system._L.rdrt_task_graph_states(self._h, *[a.ctypes.data for a in arrs]))
        return TaskGraphExecutionStates(*[a.tolist() for a in arrs])


def hemisphere_sample(s: int) -> np.ndarray:
    """Tangent-space occlusion direction of sample index ``s``."""
    out = np.zeros(3, dtype=np.float32)
    L.check(L.lib().rdrt_ao_hemisphere_sample(s, out.ctypes.data))
    return out


class AORenderer:
    """Progressive ambient occlusion of the TLAS bound at index 0."""

    def __init__(self, system, width: int, height: int, rgba8: bool = False, rounds: int = 0,
                 tile_pixels: int = 0, use_graph: bool = True):
        import torch

        self.system = system
        self.width, self.height = width, height
        self.format = L.RDRT_AO_FORMAT_RGBA8 if rgba8 else L.RDRT_AO_FORMAT_F32
        dev = f"cuda:{system.device}"
        if rgba8:
            self.buffer = torch.zeros((height, width, 4), dtype=torch.uint8, device=dev)
        else:
            self.buffer = torch.zeros((height, width), dtype=torch.float32, device=dev)
        torch.cuda.synchronize(dev)  # the renders run on the system's stream
        self.sample_count = 0
        self.rounds, self.tile_pixels, self.use_graph = rounds, tile_pixels, use_graph

    def reset_sample(self) -> None:
        self.sample_count = 0

    def params(self, camera) -> L.RdrtAoParams:
        p = L.RdrtAoParams()
        p.width, p.height, p.sample_index, p.format = self.width, self.height, self.sample_count, self.format
        p.view_projection_inv[:] = [float(x) for x in camera.view_projection_inv().T.reshape(-1)]
        p.camera_position[:] = [float(np.float32(x)) for x in camera.position]
        p.rounds, p.tile_pixels, p.use_graph = self.rounds, self.tile_pixels, int(self.use_graph)
        return p

    def render(self, camera):
        """One more sample into the buffer (returned); the buffer starts over
        at sample 0 (the reference clears its texture on resize)."""
        p = self.params(camera)
        L.check(self.system._L.rdrt_ao_render(self.system._h, C.byref(p), int(self.buffer.data_ptr())))
        self.sample_count += 1
        return self.buffer

    def autotune(self, camera, reps: int = 3) -> dict:
        """Time every traversal launch shape on AO calls of this camera and
        keep the fastest for the system (the buffer is overwritten and the
        sample count restarts); returns the choice."""
        p = self.params(camera)
        t = L.RdrtTune()
        L.check(self.system._L.rdrt_ao_autotune(self.system._h, C.byref(p), int(self.buffer.data_ptr()), reps,
                                                C.byref(t)))
        self.sample_count = 0
        return self.system.tune()

    def ray_counts(self):
        """(camera rays, occlusion rays) traced by the last render."""
        out = np.zeros(2, dtype=np.uint32)
        L.check(self.system._L.rdrt_ao_ray_counts(self.system._h, out.ctypes.data))
        return int(out[0]), int(out[1])

    def states(self) -> TaskGraphExecutionStates:
        arrs = [np.zeros(5, dtype=np.uint32) for _ in range(4)]
        L.check(self.system._L.rdrt_ao_states(self.system._h, *[a.ctypes.data for a in arrs]))
        return TaskGraphExecutionStates(*[a.tolist() for a in arrs])
