"""Wavefront path tracer on the MI355X ray-query path.

Host mirror of the reference's ``DeviceReferencePathTracingRenderer``
(scene/rendering/gpu-ray-tracing/src/feature/path_tracing/mod.rs:102-236):
a camera, a configuration (size, samples, depth), and ``render`` producing
an FP32 radiance image on the device.  The frame runs as HIP kernels
(librdrt ``rdrt_pt_render``), replayed from a captured hipGraph.

Estimator (per path, bounces b = 0..max_bounces):
  miss            -> L += beta * sky(dir.y > 0 ? sky_up : sky_down)
  hit             -> L += beta * emission
  b < max_bounces -> next-event estimation toward the directional light
                     (any-hit shadow ray), Lambertian BSDF sample
                     (beta *= albedo), Russian roulette from rr_start with
                     q = min(max(beta), 0.95).
Samples: pair k of pixel (px, py) uses the Sobol (0,2)-sequence at index
sample_index_base + s, XOR-scrambled by xxhash32(px, py, seed * 2^16 + k);
k = 0 pixel jitter, k = 1 + b BSDF direction, roulette uses dimension 64 + b.
"""

from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field
from typing import Optional, Sequence

import numpy as np

from . import _lib as L
from .acceleration import (
    AccelerationStructureSystem,
    TopLevelAccelerationStructureSourceInstance,
    TriangleGeometry,
)


@dataclass
class Camera:
    """Pinhole camera; ``view_projection_inv`` maps NDC (z in [0, 1]) to world."""

    position: Sequence[float] = (0.0, 0.0, 32.0)
    target: Sequence[float] = (0.0, 0.0, 0.0)
    up: Sequence[float] = (0.0, 1.0, 0.0)
    fov_y_deg: float = 45.0
    aspect: float = 16.0 / 9.0
    near: float = 0.1
    far: float = 1000.0

    def view(self) -> np.ndarray:
        eye, tgt, up = (np.asarray(x, dtype=np.float64) for x in (self.position, self.target, self.up))
        f = tgt - eye
        f /= np.linalg.norm(f)
        r = np.cross(f, up)
        r /= np.linalg.norm(r)
        u = np.cross(r, f)
        m = np.eye(4)
        m[0, :3], m[1, :3], m[2, :3] = r, u, -f
        m[:3, 3] = -m[:3, :3] @ eye
        return m

    def projection(self) -> np.ndarray:
        t = 1.0 / np.tan(np.radians(self.fov_y_deg) / 2.0)
        n, fa = self.near, self.far
        m = np.zeros((4, 4))
        m[0, 0] = t / self.aspect
        m[1, 1] = t
        m[2, 2] = fa / (n - fa)
        m[2, 3] = n * fa / (n - fa)
        m[3, 2] = -1.0
        return m

    def view_projection_inv(self) -> np.ndarray:
        """4x4 float32 in math order (row, col)."""
        return np.linalg.inv(self.projection() @ self.view()).astype(np.float32)


@dataclass
class PathTracingConfig:
    width: int = 1920
    height: int = 1080
    spp: int = 1
    max_bounces: int = 4
    rr_start: int = 2
    seed: int = 0
    sample_index_base: int = 0
    sky_up: Sequence[float] = (0.7, 0.7, 0.7)
    sky_down: Sequence[float] = (0.3, 0.3, 0.3)
    sun_direction: Sequence[float] = (-0.27, -0.92, -0.28)
    sun_radiance: Sequence[float] = (3.0, 3.0, 3.0)
    use_graph: bool = True
    diagnostics: int = 0  # L.RDRT_PT_DIAG_* bits (eager frame)
    record: bool = False  # keep the path vertices for PathTracer.backward

    def to_c(self, camera: Camera) -> L.RdrtPtParams:
        p = L.RdrtPtParams()
        p.view_projection_inv[:] = [float(x) for x in camera.view_projection_inv().T.reshape(-1)]
        p.camera_position[:] = [float(np.float32(x)) for x in camera.position]
        p.width, p.height, p.spp = self.width, self.height, self.spp
        p.max_bounces, p.rr_start, p.seed = self.max_bounces, self.rr_start, self.seed
        p.sample_index_base = self.sample_index_base
        sd = np.asarray(self.sun_direction, dtype=np.float64)
        sd = (sd / np.linalg.norm(sd)).astype(np.float32)
        p.sky_up[:] = [float(x) for x in self.sky_up]
        p.sky_down[:] = [float(x) for x in self.sky_down]
        p.sun_direction[:] = [float(x) for x in sd]
        p.sun_radiance[:] = [float(x) for x in self.sun_radiance]
        p.use_graph = 1 if self.use_graph else 0
        p.diagnostics = self.diagnostics
        p.record = 1 if self.record else 0
        return p


class PathTracer:
    """DeviceReferencePathTracingRenderer on librdrt."""

    def __init__(self, system: AccelerationStructureSystem):
        self.system = system
        self.recorded_frames = 0  # frames rendered with record = True

    def render(self, camera, cfg: PathTracingConfig, out=None):
        """Render one frame; returns the (height, width, 3) float32 device
        radiance tensor (``out`` if given).  ``camera`` may be a list of
        cameras: all views in one frame, (views, height, width, 3)."""
        import torch

        cams = list(camera) if isinstance(camera, (list, tuple)) else [camera]
        nv = len(cams)
        shape = (cfg.height, cfg.width, 3) if nv == 1 and not isinstance(camera, (list, tuple)) else (
            nv, cfg.height, cfg.width, 3)
        if out is None:
            out = torch.empty(shape, dtype=torch.float32, device=f"cuda:{self.system.device}")
        if not out.is_contiguous() or out.numel() != nv * cfg.width * cfg.height * 3:
            raise ValueError("out must be a contiguous views*height*width*3 float32 tensor")
        prm = cfg.to_c(cams[0])
        views = None
        if nv > 1:
            views = (L.f32 * (20 * nv))()
            for v, cam in enumerate(cams):
                views[20 * v:20 * v + 16] = [float(x) for x in cam.view_projection_inv().T.reshape(-1)]
                views[20 * v + 16:20 * v + 19] = [float(np.float32(x)) for x in cam.position]
            prm.views = nv
            prm.view_params = C.cast(views, C.c_void_p)
        L.check(self.system._L.rdrt_pt_render(self.system._h, C.byref(prm), int(out.data_ptr())))
        if cfg.record:
            self.recorded_frames += 1
        return out

    def autotune(self, camera, cfg: PathTracingConfig, out=None, reps: int = 3) -> dict:
        """Time every traversal launch shape on eager frames of this camera
        (or list of cameras, all views in each frame) and configuration and
        keep the fastest for the system (``out`` is overwritten); returns the
        choice, its ``ms`` the frame time."""
        import torch

        if cfg.record or cfg.diagnostics:
            raise ValueError("tune on plain frames (record and diagnostics off)")
        cams = list(camera) if isinstance(camera, (list, tuple)) else [camera]
        nv = len(cams)
        if out is None:
            out = torch.empty((nv, cfg.height, cfg.width, 3), dtype=torch.float32, device=f"cuda:{self.system.device}")
        if not out.is_contiguous() or out.numel() != nv * cfg.width * cfg.height * 3:
            raise ValueError("out must be a contiguous views*height*width*3 float32 tensor")
        prm = cfg.to_c(cams[0])
        views = None
        if nv > 1:  # every view in each timed frame, as render does
            views = (L.f32 * (20 * nv))()
            for v, cam in enumerate(cams):
                views[20 * v:20 * v + 16] = [float(x) for x in cam.view_projection_inv().T.reshape(-1)]
                views[20 * v + 16:20 * v + 19] = [float(np.float32(x)) for x in cam.position]
            prm.views = nv
            prm.view_params = C.cast(views, C.c_void_p)
        t = L.RdrtTune()
        L.check(self.system._L.rdrt_pt_autotune(self.system._h, C.byref(prm), int(out.data_ptr()), reps, C.byref(t)))
        return self.system.tune()

    def set_material_params(self, albedo, emission) -> None:
        """Replace the materials with device tensors (n, 3) float32 each,
        copied on the system's stream (no host round trip)."""
        for t, name in ((albedo, "albedo"), (emission, "emission")):
            if not getattr(t, "is_cuda", False) or not t.is_contiguous() or t.dtype != _f32():
                raise ValueError(f"{name} must be a contiguous float32 device tensor")
        n = albedo.numel() // 3
        if albedo.numel() != 3 * n or emission.numel() != 3 * n:
            raise ValueError("albedo and emission must both be (n, 3)")
        L.check(self.system._L.rdrt_pt_set_material_params(self.system._h, int(albedo.data_ptr()),
                                                            int(emission.data_ptr()), n))

    def backward(self, adj, grad_albedo, grad_emission) -> None:
        """Gradient of sum(adj * frame) of the last recorded frame w.r.t. the
        material albedo and emission (written to the (n, 3) float32 tensors)."""
        for t in (adj, grad_albedo, grad_emission):
            if not getattr(t, "is_cuda", False) or not t.is_contiguous() or t.dtype != _f32():
                raise ValueError("backward takes contiguous float32 device tensors")
        L.check(self.system._L.rdrt_pt_backward(self.system._h, int(adj.data_ptr()), int(grad_albedo.data_ptr()),
                                                int(grad_emission.data_ptr())))

    def backward_geometry(self, adj, grad_positions, grad_normals=None) -> None:
        """Gradient of sum(adj * frame) of the last recorded frame w.r.t.
        vertex positions and normals ((vertex_count, 3) float32 tensors in
        scene vertex order), replaying the frame with its paths held fixed."""
        for t in (adj, grad_positions) + (() if grad_normals is None else (grad_normals,)):
            if not getattr(t, "is_cuda", False) or not t.is_contiguous() or t.dtype != _f32():
                raise ValueError("backward_geometry takes contiguous float32 device tensors")
        L.check(self.system._L.rdrt_pt_backward_geometry(
            self.system._h, int(adj.data_ptr()), int(grad_positions.data_ptr()),
            None if grad_normals is None else int(grad_normals.data_ptr())))

    def l2_adjoint(self, image, target, adj, loss) -> None:
        """adj = 2 (image - target); loss (one float64 device element) =
        sum (image - target)^2, enqueued on the system's stream."""
        n = image.numel()
        if target.numel() != n or adj.numel() != n or loss.numel() != 1:
            raise ValueError("shape mismatch")
        L.check(self.system._L.rdrt_l2_adjoint(self.system._h, int(image.data_ptr()), int(target.data_ptr()), n,
                                               int(adj.data_ptr()), int(loss.data_ptr())))

    def stats(self) -> dict:
        s = L.RdrtPtStats()
        L.check(self.system._L.rdrt_pt_get_stats(self.system._h, C.byref(s)))
        nb = s.bounces
        return {"extension_rays": list(s.extension_rays[:nb]), "shadow_rays": list(s.shadow_rays[:nb]),
                "frame_ms": float(s.frame_ms), "trace_ms": float(s.trace_ms), "trace_launches": int(s.trace_launches),
                "visits": [int(x) for x in s.visits],
                "max_ray_visits": [int(x) for x in s.max_ray_visits[:s.trace_launches]],
                "launch_ms": [float(x) for x in s.launch_ms[:s.trace_launches]]}


def _f32():
    import torch

    return torch.float32


def render_materials(tracer: PathTracer, camera: Camera, cfg: PathTracingConfig, albedo, emission):
    """Differentiable frame: an autograd op of (albedo, emission) -> frame
    (height, width, 3), whose backward runs the adjoint kernel on the
    recorded path vertices (visibility and sampled directions fixed)."""
    import dataclasses

    import torch

    def ordered(fn):
        # torch's stream and the system's stream: wait on both sides unless they are one stream
        cur = torch.cuda.current_stream()
        same = tracer.system.stream == cur.cuda_stream
        if not same:
            cur.synchronize()
        out = fn()
        if not same:
            tracer.system.sync()
        return out

    class _Render(torch.autograd.Function):
        @staticmethod
        def forward(ctx, a, e):
            a = a.detach().contiguous()
            e = e.detach().contiguous()
            img = torch.empty((cfg.height, cfg.width, 3), dtype=torch.float32, device=a.device)

            def run():
                tracer.set_material_params(a, e)
                return tracer.render(camera, dataclasses.replace(cfg, record=True), out=img)

            ctx.shapes = (a.shape, e.shape)
            ctx.device = a.device
            out = ordered(run)
            ctx.frame = tracer.recorded_frames
            return out

        @staticmethod
        def backward(ctx, grad):
            if ctx.frame != tracer.recorded_frames:
                raise RuntimeError("the path vertices of this frame were replaced by a later recorded frame")
            ga = torch.empty(ctx.shapes[0], dtype=torch.float32, device=ctx.device)
            ge = torch.empty(ctx.shapes[1], dtype=torch.float32, device=ctx.device)
            g = grad.contiguous()
            ordered(lambda: tracer.backward(g, ga, ge))
            return ga, ge

    return _Render.apply(albedo, emission)


def load_scene(system: AccelerationStructureSystem, scene: dict) -> None:
    """Create the BLAS / TLAS / attributes / materials of a scene dict from
    ``synth.pt_scene`` (or any dict with the same keys) and commit."""
    handles = []
    for mesh in scene["meshes"]:
        h = system.create_bottom_level_acceleration_structure([TriangleGeometry(mesh["positions"], mesh["indices"])])
        if mesh.get("normals") is not None:
            system.set_vertex_normals(h, 0, mesh["normals"])
        if mesh.get("tri_material") is not None:
            system.set_triangle_materials(h, 0, mesh["tri_material"])
        handles.append(h)
    inst = [TopLevelAccelerationStructureSourceInstance(d["transform"], handles[d["mesh"]],
                                                        instance_custom_index=d["material"])
            for d in scene["instances"]]
    t = system.create_top_level_acceleration_structure(inst)
    system.bind_tlas([t])
    system.set_materials(scene["albedo"], scene["emission"])
    system.create_comp_instance()
