"""Entry points of the boundary that restate others: the one-call BLAS
(rdrt_blas_create), the host Sobol point (rdrt_sobol), the metrics record
(rdrt_get_metrics) and rdrt_pt_forward."""

import ctypes as C

import numpy as np
import pytest

from oracle import oracle as O
from rendiation_amd import _lib as L
from rendiation_amd import synth
from rendiation_amd.acceleration import (
    AccelerationStructureSystem,
    BuildOptions,
    TopLevelAccelerationStructureSourceInstance,
    TriangleGeometry,
)


def test_host_sobol_equals_oracle():
    lib = L.lib()
    out = np.zeros(2, np.float32)
    rng = np.random.default_rng(3)
    ns = list(range(0, 300)) + [int(x) for x in rng.integers(0, 2**32, 300)]
    for n in ns:
        s0, s1 = (int(x) for x in rng.integers(0, 2**32, 2))
        L.check(lib.rdrt_sobol(n, s0, s1, out.ctypes.data))
        assert out[0] == np.float32(O.van_der_corput(n, s0)), (n, s0)
        assert out[1] == np.float32(O.sobol(n, s1)), (n, s1)


def _forest(sys):
    return sys.copy_built("blas_forest"), sys.copy_built("redirect")


def test_blas_create_equals_begin_add_end():
    pos = synth.random_triangles(8, 3000)
    idx = np.arange(pos.shape[0], dtype=np.uint32).reshape(-1, 3)
    a = AccelerationStructureSystem(device=None, options=BuildOptions(blas_bin_size=2))
    ha = a.create_bottom_level_acceleration_structure([TriangleGeometry(pos, idx, 1)])
    b = AccelerationStructureSystem(device=None, options=BuildOptions(blas_bin_size=2))
    out = C.c_uint32(0)
    p = np.ascontiguousarray(pos, np.float32)
    L.check(b._L.rdrt_blas_create(b._h, p.ctypes.data, p.shape[0], idx.ctypes.data, idx.shape[0], 1, None,
                                  C.byref(out)))
    opts = L.RdrtBuildOpts(2, 50, 10, 50, 4)  # equal to the context's: accepted
    out2 = C.c_uint32(0)
    L.check(b._L.rdrt_blas_create(b._h, p.ctypes.data, p.shape[0], idx.ctypes.data, idx.shape[0], 1,
                                  C.byref(opts), C.byref(out2)))
    assert (out.value, out2.value) == (0, 1) and ha.index == 0
    ta = a.create_top_level_acceleration_structure([TopLevelAccelerationStructureSourceInstance(np.eye(4), ha)])
    a.bind_tlas([ta])
    tb = b.create_top_level_acceleration_structure(
        [TopLevelAccelerationStructureSourceInstance(np.eye(4), type(ha)(out.value))])
    b.bind_tlas([tb])
    a.build()
    b.build()
    fa, ra = _forest(a)
    fb, rb = _forest(b)
    n = fa.shape[0]
    assert np.array_equal(fb[:n], fa) and np.array_equal(rb[:ra.shape[0]], ra)
    bad = L.RdrtBuildOpts(4, 50, 10, 50, 4)
    with pytest.raises(L.RdrtError):
        L.check(b._L.rdrt_blas_create(b._h, p.ctypes.data, p.shape[0], idx.ctypes.data, idx.shape[0], 1,
                                      C.byref(bad), C.byref(out2)))
    m = L.RdrtMetrics()
    L.check(b._L.rdrt_get_metrics(b._h, C.byref(m)))
    assert m.last_call_ms == -1.0 and m.build_ms > 0.0 and m.status == 0
    a.close()
    b.close()


@pytest.mark.gpu
def test_pt_forward_equals_render_and_metrics(device):
    from rendiation_amd.pathtracer import Camera, PathTracer, PathTracingConfig, load_scene

    print("[gpu-case] rdrt_pt_forward vs rdrt_pt_render", flush=True)
    scene = synth.pt_scene(seed=2, n_meshes=8, tess=16, extent=3.0)
    s = AccelerationStructureSystem(device=0)
    load_scene(s, scene)
    pt = PathTracer(s)
    cam = Camera(position=(0.0, 0.0, 12.0))
    cfg = PathTracingConfig(width=64, height=48, spp=1, max_bounces=3, use_graph=False)
    a = pt.render(cam, cfg).clone()
    s.sync()
    import torch

    b = torch.empty_like(a)
    prm = cfg.to_c(cam)
    L.check(s._L.rdrt_pt_forward(s._h, C.byref(prm), int(b.data_ptr())))
    s.sync()
    assert np.array_equal(a.cpu().numpy(), b.cpu().numpy())
    m = L.RdrtMetrics()
    L.check(s._L.rdrt_get_metrics(s._h, C.byref(m)))
    assert m.last_call_ms > 0.0 and m.device_bytes > 0 and m.status == 0
    s.close()
