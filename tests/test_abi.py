"""The C-ABI library loads, exports every function rdrt.h declares, and its
record layouts match; host-only calls behave (no compute, no GPU needed)."""

import ctypes as C
import re
from pathlib import Path

import numpy as np
import pytest

from rendiation_amd import _lib as L
from rendiation_amd.acceleration import AccelerationStructureSystem, TriangleGeometry

ROOT = Path(__file__).resolve().parent.parent


def _declared(header: Path):
    text = header.read_text()
    return sorted(set(re.findall(r"\b(rdrt_[a-z0-9_]+)\s*\(", text)))


def test_every_declared_symbol_is_exported():
    lib = C.CDLL(str(L.LIB_PATH))
    names = _declared(ROOT / "include" / "rdrt.h")
    assert len(names) >= 20
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    # and the binding table covers the header exactly
    assert sorted(L.SIGNATURES) == names


def test_record_sizes():
    assert C.sizeof(L.RdrtRay) == 32
    assert C.sizeof(L.RdrtRayMeta) == 16
    assert C.sizeof(L.RdrtHit) == 16
    assert C.sizeof(L.RdrtHitMeta) == 16
    assert C.sizeof(L.RdrtInstance) == 96
    assert C.sizeof(L.RdrtTraceArgs) == 5 * 8 + 6 * 4


def test_version_and_defaults():
    lib = L.lib()
    assert lib.rdrt_abi_version() == 1
    cfg = L.RdrtConfig()
    L.check(lib.rdrt_config_default(C.byref(cfg)))
    b = cfg.build
    assert (b.blas_bin_size, b.blas_max_depth, b.tlas_bin_size, b.tlas_max_depth, b.sah_bins) == (2, 50, 10, 50, 4)


def test_host_only_context_refuses_device_work():
    sys = AccelerationStructureSystem(device=None)
    sys.create_bottom_level_acceleration_structure([TriangleGeometry(np.zeros((3, 3), np.float32))])
    with pytest.raises(L.RdrtError) as e:
        sys.create_comp_instance()
    assert e.value.code == -6


def test_errors_are_reported():
    lib = L.lib()
    sys = AccelerationStructureSystem(device=None)
    with pytest.raises(L.RdrtError) as e:
        sys.delete_bottom_level_acceleration_structure(type("H", (), {"index": 9})())
    assert e.value.code == -1
    assert b"unknown BLAS" in lib.rdrt_last_error()
    # index out of range of the geometry's vertices
    with pytest.raises(L.RdrtError):
        sys.create_bottom_level_acceleration_structure(
            [TriangleGeometry(np.zeros((3, 3), np.float32), np.array([[0, 1, 5]], np.uint32))])
    # blas_end without begin
    out = L.u32()
    assert lib.rdrt_blas_end(sys._h, C.byref(out)) == -5
    # scene info before a build
    info = L.RdrtSceneInfo()
    assert lib.rdrt_get_scene_info(sys._h, C.byref(info)) == -3
