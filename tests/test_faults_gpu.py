"""Bounded device waits on the GPU. The host sets the spin limit to 0, so every wait gives up at once without polling (the
condition is forced from the host: no kernel hangs or faults, each runs to its normal end); the give-up must be reported
(the device's fault word through cs_hip_take_fault, and Model::wait's error through the workload handles), the word must
be cleared by that report, and the next work with the default limit must be correct again (the scan workspace and the
ticket state are left zeroed even by work whose waits gave up)."""
import ctypes as C

import numpy as np
import pytest

import oracle as O
from cospanmdao_amd import CsError, check, hip, workloads
from helpers import DeviceArray, init_device

pytestmark = pytest.mark.gpu

FAULT_TICKET_WAIT, FAULT_SCAN_LOOKBACK = 1, 2
DEFAULT_LIMIT = 1 << 24


def take_fault():
    bits = C.c_uint32(0xFFFFFFFF)
    check(hip.cs_hip_take_fault(C.byref(bits)))
    return bits.value


def test_scan_lookback_give_up_is_reported():
    init_device()
    n = (1 << 20) + 77  # 513 tiles: every tile but the first looks back
    v = (np.arange(n, dtype=np.uint64) * 2654435761 % 11).astype(np.uint32)
    want = O.exclusive_scan_u32(v)
    din, dout = DeviceArray(v), DeviceArray(nbytes=(n + 1) * 4)
    ws_bytes = hip.cs_hip_scan_workspace_bytes(n)
    ws = DeviceArray(nbytes=ws_bytes)
    check(hip.cs_hip_scan_workspace_init(ws.ptr, ws_bytes, None))
    assert take_fault() == 0
    try:
        print("scan with spin limit 0, n =", n, flush=True)
        check(hip.cs_hip_set_spin_limit(0))
        check(hip.cs_hip_exclusive_scan_u32(din.ptr, dout.ptr, n, ws.ptr, None))
        check(hip.cs_hip_stream_synchronize(None))
        assert take_fault() & FAULT_SCAN_LOOKBACK
        assert take_fault() == 0, "the fault word is cleared by the read that reports it"
    finally:
        check(hip.cs_hip_set_spin_limit(DEFAULT_LIMIT))
    assert hip.cs_hip_spin_limit() == DEFAULT_LIMIT
    print("scan with the default limit", flush=True)
    check(hip.cs_hip_exclusive_scan_u32(din.ptr, dout.ptr, n, ws.ptr, None))
    check(hip.cs_hip_stream_synchronize(None))
    assert np.array_equal(dout.get(np.uint32, n + 1), want)
    assert take_fault() == 0


def test_fused_group_wait_give_up_is_reported():
    """The compressible NS residual's fused variant (one fused group with group waits: ticket mode) with the limit at 0:
    cs_nsc_wait fails with the bounded-wait message; with the limit restored the next runs equal the unfused variant bit
    for bit (the ticket state was zeroed by the run whose waits gave up)."""
    init_device()
    n = 256

    def make(variant):
        h = C.c_void_p()
        check(workloads.cs_nsc_create(C.byref(h), variant, n, n, 0, 1, 8))
        check(workloads.cs_nsc_init(h, 1.0))
        return h

    def residual(h):
        out = np.empty((4, n, n), np.float32)
        check(workloads.cs_nsc_read(h, b"R", out.ctypes.data, out.nbytes))
        return out

    ref = make(0)
    fused = make(1)
    try:
        check(workloads.cs_nsc_run(ref))
        check(workloads.cs_nsc_wait(ref))
        want = residual(ref)
        try:
            print("fused NS residual with spin limit 0", flush=True)
            check(hip.cs_hip_set_spin_limit(0))
            check(workloads.cs_nsc_run(fused))
            with pytest.raises(CsError) as e:
                check(workloads.cs_nsc_wait(fused))
            assert "gave up" in e.value.message and "fused group" in e.value.message, e.value.message
        finally:
            check(hip.cs_hip_set_spin_limit(DEFAULT_LIMIT))
        assert take_fault() == 0, "the wait that reported the fault cleared the word"
        print("fused NS residual with the default limit", flush=True)
        for _ in range(3):  # eager (a new limit is a new launch), then captured and replayed
            check(workloads.cs_nsc_run(fused))
        check(workloads.cs_nsc_wait(fused))
        assert np.array_equal(residual(fused), want)
    finally:
        check(workloads.cs_nsc_destroy(fused))
        check(workloads.cs_nsc_destroy(ref))
