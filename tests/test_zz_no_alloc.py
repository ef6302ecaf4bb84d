"""cs_hip_exclusive_scan_u32_static, the scan that allocates nothing, against cs_hip_exclusive_scan_u32 (the same scan with
a workspace) through the C boundary: bit for bit on calls in a row on one device, the second call smaller than the first
(the fixed workspace keeps the last call's contents; each scan must leave it zeroed), then n = 1 and n = 0, and an n past
the fixed workspace refused before anything is launched."""
import ctypes as C

import numpy as np
import pytest

import oracle as O
from cospanmdao_amd import CsError, check, hip
from helpers import DeviceArray, init_device

pytestmark = pytest.mark.gpu

MAX_N = 1 << 26  # CS_HIP_SCAN_STATIC_MAX_N


def pattern(n, seed):
    i = np.arange(n, dtype=np.uint64)
    return ((i * np.uint64(2654435761) + np.uint64(seed)) % np.uint64(1000003)).astype(np.uint32)


def test_static_scan_equals_workspace_scan():
    init_device()
    for n, seed in ((3_000_001, 7), (70_001, 11), (1, 3), (0, 5)):
        print("static scan n =", n, flush=True)
        v = pattern(n, seed)
        din = DeviceArray(v) if n else DeviceArray(nbytes=16)
        a, b = DeviceArray(nbytes=(n + 1) * 4), DeviceArray(nbytes=(n + 1) * 4)
        check(hip.cs_hip_exclusive_scan_u32(din.ptr, a.ptr, n, None, None))
        check(hip.cs_hip_exclusive_scan_u32_static(din.ptr, b.ptr, n, None))
        check(hip.cs_hip_stream_synchronize(None))
        got, want = b.get(np.uint32, n + 1), a.get(np.uint32, n + 1)
        assert np.array_equal(got, want), n
        assert np.array_equal(got, O.exclusive_scan_u32(v)), n
        bits = C.c_uint32(0xFFFFFFFF)
        check(hip.cs_hip_take_fault(C.byref(bits)))
        assert bits.value == 0


def test_static_scan_refuses_n_past_its_workspace():
    init_device()
    tiny = DeviceArray(nbytes=16)
    with pytest.raises(CsError) as e:
        check(hip.cs_hip_exclusive_scan_u32_static(tiny.ptr, tiny.ptr, MAX_N + 1, None))  # rejected before any launch
    assert e.value.status == 1 and "at most" in e.value.message
