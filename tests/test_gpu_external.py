"""A buffer exported by another API as an opaque file descriptor (the form
a Vulkan device gives wgpu's buffers) is imported through the C-ABI and
used in place: rays are read from it and hits written to it by rdrt_trace,
and the other mapping of the same memory sees the hits the plain launch
gives for the same rays."""

import ctypes as C
import os
from pathlib import Path

import numpy as np
import pytest

from rendiation_amd import _lib as L
from rendiation_amd import synth
from rendiation_amd.acceleration import (
    AccelerationStructureSystem,
    BuildOptions,
    TopLevelAccelerationStructureSourceInstance,
    TriangleGeometry,
)

pytestmark = pytest.mark.gpu

STAGES = Path(__file__).resolve().parent / "stages" / "libpipeline_stages.so"


def _helpers():
    lib = C.CDLL(str(STAGES))
    lib.test_exportable_create.restype = C.c_int
    lib.test_exportable_create.argtypes = [C.c_size_t, C.POINTER(C.c_int), C.POINTER(C.c_void_p),
                                           C.POINTER(C.c_size_t), C.POINTER(C.c_ulonglong)]
    lib.test_exportable_destroy.restype = C.c_int
    lib.test_exportable_destroy.argtypes = [C.c_void_p, C.c_size_t, C.c_ulonglong]
    lib.test_copy.restype = C.c_int
    lib.test_copy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    return lib


def test_imported_buffer_is_traced_in_place(device):
    import torch

    print("[gpu-case] external buffer import", flush=True)
    h = _helpers()
    n = 1 << 14
    rays = synth.random_rays(5, n)
    nbytes = n * 32 + n * 16
    fd, ptr, mapped, handle = C.c_int(-1), C.c_void_p(), C.c_size_t(), C.c_ulonglong()
    rc = h.test_exportable_create(nbytes, C.byref(fd), C.byref(ptr), C.byref(mapped), C.byref(handle))
    if rc == 1:
        pytest.skip("no virtual memory management on this device: nothing exports a file descriptor")
    assert rc == 0
    try:
        system = AccelerationStructureSystem(device=0, options=BuildOptions(blas_bin_size=4))
        b = system.create_bottom_level_acceleration_structure([TriangleGeometry(synth.random_triangles(5, 5000))])
        t = system.create_top_level_acceleration_structure([TopLevelAccelerationStructureSourceInstance(np.eye(4), b)])
        system.bind_tlas([t])
        system.create_comp_instance()
        # the other API's mapping: rays in, hits cleared
        host = np.zeros(nbytes, dtype=np.uint8)
        host[: n * 32] = np.ascontiguousarray(rays).view(np.uint8).reshape(-1)
        host[n * 32:] = 0xEE
        assert h.test_copy(ptr, host.ctypes.data, nbytes, 0) == 0
        buf, dptr = C.c_void_p(), C.c_void_p()
        dup = os.dup(fd.value)
        rc = system._L.rdrt_external_buffer_import(system._h, dup, mapped.value, 0, mapped.value, C.byref(buf),
                                                   C.byref(dptr))
        if rc != 0:
            os.close(dup)
            L.check(rc)
        try:
            a = L.RdrtTraceArgs()
            a.rays, a.ray_meta, a.hits = dptr.value, None, dptr.value + n * 32
            a.hit_meta, a.counters = None, None
            a.n, a.flags, a.cull_mask, a.tlas_idx, a.traversal = n, 0, 0xFFFFFFFF, 0, 0
            L.check(system._L.rdrt_trace(system._h, C.byref(a)))
            system.sync()
        finally:
            L.check(system._L.rdrt_external_buffer_release(buf))
        got = np.zeros(n * 16, dtype=np.uint8)
        assert h.test_copy(got.ctypes.data, C.c_void_p(ptr.value + n * 32), n * 16, 1) == 0
        got = got.view(np.uint32).reshape(n, 4)
        rd = torch.from_numpy(np.ascontiguousarray(rays)).cuda()
        hd = torch.zeros((n, 4), dtype=torch.int32, device="cuda:0")
        system.trace(rd, hd)
        system.sync()
        want = hd.cpu().numpy().view(np.uint32)
        assert (want[:, 0] != 0xFFFFFFFF).sum() > 100
        assert np.array_equal(got, want)
        system.close()
    finally:
        os.close(fd.value)
        assert h.test_exportable_destroy(ptr, mapped.value, handle.value) == 0
