"""rdrt_compact_static (no device allocation) against rdrt_compact, through
the C-ABI: identical index lists and counts on two calls in a row, the
second with a smaller input (the fixed scratch keeps the first call's
contents), and a clear capacity error above RDRT_COMPACT_STATIC_MAX."""

import ctypes as C

import numpy as np
import pytest

from rendiation_amd import _lib as L
from rendiation_amd.acceleration import AccelerationStructureSystem

pytestmark = pytest.mark.gpu

STATIC_MAX = 2097152


def _say(msg):
    print(f"[gpu-case] {msg}", flush=True)


def _run(lib, h, fn, mask, n, device):
    import torch

    out = torch.full((max(n, 1),), -1, dtype=torch.int32, device=device)
    cnt = torch.full((1,), -1, dtype=torch.int32, device=device)
    L.check(getattr(lib, fn)(h, C.c_void_p(mask.data_ptr()), n, C.c_void_p(out.data_ptr()),
                             C.c_void_p(cnt.data_ptr())))
    return out, cnt


def test_static_compaction_matches_allocating_entry_point(device):
    import torch

    s = AccelerationStructureSystem(device=0)
    lib, h = s._L, s._h
    for n, p, seed in ((STATIC_MAX, 0.41, 1), (100003, 0.77, 2)):
        _say(f"rdrt_compact_static vs rdrt_compact, n={n}")
        mask_np = (np.random.default_rng(seed).random(n) < p).astype(np.int32)
        mask = torch.from_numpy(mask_np).to(device)
        o_ref, c_ref = _run(lib, h, "rdrt_compact", mask, n, device)
        o_st, c_st = _run(lib, h, "rdrt_compact_static", mask, n, device)
        s.sync()
        k = int(c_ref.item())
        assert k == int(np.count_nonzero(mask_np))
        assert int(c_st.item()) == k
        assert torch.equal(o_st[:k], o_ref[:k])
        assert bool((o_st[k:] == -1).all())
    s.close()


def test_static_compaction_rejects_oversize(device):
    import torch

    s = AccelerationStructureSystem(device=0)
    lib, h = s._L, s._h
    _say("rdrt_compact_static above its maximum")
    mask = torch.zeros(STATIC_MAX + 1, dtype=torch.int32, device=device)
    out = torch.zeros(STATIC_MAX + 1, dtype=torch.int32, device=device)
    cnt = torch.zeros(1, dtype=torch.int32, device=device)
    rc = lib.rdrt_compact_static(h, C.c_void_p(mask.data_ptr()), STATIC_MAX + 1, C.c_void_p(out.data_ptr()),
                                 C.c_void_p(cnt.data_ptr()))
    assert rc == -7  # RDRT_ERR_CAPACITY
    assert b"at most 2097152" in lib.rdrt_last_error()
    s.close()
