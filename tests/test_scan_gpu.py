"""The device scan and the ragged column build against the oracle and the reference's known answers (bit-exact)."""
import ctypes as C
import json
import os

import numpy as np
import pytest

import oracle as O
from cospanmdao_amd import check, hip
from helpers import DeviceArray, init_device

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))


def scan_pattern(n):
    return np.array([(i * 2654435761 + 7) % (2 ** 64) % 11 for i in range(n)], np.uint32)


def device_scan(values, workspace=None):
    values = np.ascontiguousarray(values, np.uint32)
    n = values.size
    din = DeviceArray(values if n else np.zeros(1, np.uint32))
    dout = DeviceArray(nbytes=4 * (n + 1))
    check(hip.cs_hip_exclusive_scan_u32(din.ptr, dout.ptr, n, workspace.ptr if workspace else None, None))
    return dout.get(np.uint32, n + 1)


def setup_module():
    init_device()


def test_scan_known_answers():
    with open(os.path.join(HERE, "golden", "scan_kat.json")) as f:
        cases = json.load(f)["cases"]
    for case in cases:
        v = scan_pattern(case["n"]) if case["input"] == "pattern" else np.array(case["values"], np.uint32)
        print("scan KAT", case["name"], flush=True)
        out = device_scan(v)
        assert np.array_equal(out, O.exclusive_scan_u32(v)), case["name"]
        if "expected" in case:
            assert out.tolist() == case["expected"]


@pytest.mark.parametrize("n", [1, 63, 64, 65, 2047, 2048, 2049, 64 * 2048 + 1, 300_000, 3_000_001])
def test_scan_random_sizes(n):
    rng = np.random.default_rng(n)
    v = rng.integers(0, 2 ** 32, n, dtype=np.uint64).astype(np.uint32)  # wraps freely
    print("scan random", n, flush=True)
    assert np.array_equal(device_scan(v), O.exclusive_scan_u32(v))


def test_scan_caller_workspace_is_left_zeroed():
    n = 100_000
    ws = DeviceArray(nbytes=hip.cs_hip_scan_workspace_bytes(n))
    check(hip.cs_hip_scan_workspace_init(ws.ptr, ws.nbytes, None))
    for seed in range(3):  # the same workspace, three scans in a row
        v = np.random.default_rng(seed).integers(0, 1000, n).astype(np.uint32)
        assert np.array_equal(device_scan(v, ws), O.exclusive_scan_u32(v))
    assert not ws.get(np.uint32, ws.nbytes // 4).any()


def test_scan_in_a_captured_graph_replays():
    n = 50_000
    v = np.random.default_rng(1).integers(0, 100, n).astype(np.uint32)
    din, dout = DeviceArray(v), DeviceArray(nbytes=4 * (n + 1))
    ws = DeviceArray(nbytes=hip.cs_hip_scan_workspace_bytes(n))
    check(hip.cs_hip_scan_workspace_init(ws.ptr, ws.nbytes, None))
    check(hip.cs_hip_stream_synchronize(None))
    s = C.c_void_p()
    check(hip.cs_hip_stream_create(C.byref(s)))
    g = C.c_void_p()
    check(hip.cs_hip_graph_begin(s))
    check(hip.cs_hip_exclusive_scan_u32(din.ptr, dout.ptr, n, ws.ptr, s))
    check(hip.cs_hip_graph_end(s, C.byref(g)))
    want = O.exclusive_scan_u32(v)
    for rep in range(4):
        dout.fill_u32(0xDEADBEEF)
        check(hip.cs_hip_stream_synchronize(None))
        check(hip.cs_hip_graph_launch(g, s))
        check(hip.cs_hip_stream_synchronize(s))
        assert np.array_equal(dout.get(np.uint32, n + 1), want), rep
    check(hip.cs_hip_graph_destroy(g))
    check(hip.cs_hip_stream_destroy(s))


DTYPES = {np.dtype(np.int32): 0, np.dtype(np.uint32): 1, np.dtype(np.int64): 2, np.dtype(np.uint64): 3}


def device_column(src, offsets_mode=False, k=1):
    src = np.ascontiguousarray(src)
    count = src.shape[0]
    n = count - 1 if offsets_mode else count
    dsrc = DeviceArray(src if src.size else np.zeros(1, src.dtype))
    dext = DeviceArray(nbytes=4 * max(n, 1))
    doffs = DeviceArray(nbytes=4 * (n + 1))
    dres = DeviceArray(nbytes=8)
    check(hip.cs_hip_build_column(DTYPES[src.dtype], dsrc.ptr, count, k, 1 if offsets_mode else 0, dext.ptr, doffs.ptr, dres.ptr, None, None))
    res = dres.get(np.uint32, 2)
    return dext.get(np.uint32, n), doffs.get(np.uint32, n + 1), int(res[0]), int(res[1])


def same_column(src, offsets_mode=False, k=1):
    print("column", src.dtype, src.shape, "offsets" if offsets_mode else "counts", flush=True)
    got = device_column(src, offsets_mode, k)
    want = O.build_column(src, offsets_mode, k)
    assert np.array_equal(got[0], want[0]) and np.array_equal(got[1], want[1]) and got[2:] == want[2:], (got[2:], want[2:])
    return got


def test_column_counts_known_answer():
    # tests/test_shapes.cpp:656-682 of the reference
    for dt in DTYPES:
        ext, offs, total, flag = same_column(np.array([2, 0, 3, 1], dt))
        assert offs.tolist() == [0, 2, 2, 5, 6] and total == 6 and flag == 0


def test_column_offsets_known_answer():
    ext, offs, total, flag = same_column(np.array([0, 2, 2, 5, 6], np.int64), offsets_mode=True)
    assert ext.tolist() == [2, 0, 3, 1] and total == 6 and flag == 0


def test_column_flags():
    same_column(np.array([2, -1, 3], np.int32))
    same_column(np.array([1, 2 ** 32, 1], np.uint64))
    same_column(np.array([1, -5, 2 ** 40], np.int64))
    same_column(np.array([2 ** 31, 2 ** 31, 1], np.uint32))
    same_column(np.array([1, 2, 3], np.uint32), offsets_mode=True)
    same_column(np.array([0, 4, 3], np.int32), offsets_mode=True)
    same_column(np.array([0, -3, 3], np.int32), offsets_mode=True)
    same_column(np.zeros(0, np.int32))
    same_column(np.zeros(1, np.int64), offsets_mode=True)


def test_column_large_and_rank2():
    rng = np.random.default_rng(5)
    src = rng.integers(-3, 50, (200_001, 3)).astype(np.int64)
    same_column(src, k=3)
    offs = np.concatenate([[0], np.cumsum(rng.integers(0, 9, 100_000))]).astype(np.uint32)
    same_column(offs, offsets_mode=True)
    # wraps in the middle of a long run: exact detection
    big = np.full(70_000, 2 ** 20, np.uint32)
    same_column(big)


def test_column_rejects_an_empty_offsets_source():
    d = DeviceArray(nbytes=16)
    st = hip.cs_hip_build_column(0, d.ptr, 0, 1, 1, d.ptr, d.ptr, d.ptr, None, None)
    assert st == 1
