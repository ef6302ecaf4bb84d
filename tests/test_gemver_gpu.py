"""GEMVER on the device, every style, against the oracle: the f32 restatement of the reference's components bit for bit
where the evaluation order is the same, the FP64 oracle within a stated tolerance everywhere."""
import ctypes as C

import numpy as np
import pytest

import oracle as O
from cospanmdao_amd import check, workloads

pytestmark = pytest.mark.gpu

UNFUSED, FUSED, NAIVE, TWOPASS, TWOPASS_1 = range(5)
STYLES = {"unfused": UNFUSED, "fused": FUSED, "naive": NAIVE, "twopass": TWOPASS, "twopass_1": TWOPASS_1}
ALPHA, BETA = 1.5, 1.2
# f32 against FP64: rel-L2 (the reference's metric, chain_data.hpp:149-161; its gate is 1e-4, main.cpp:47)
TOL_L2 = 1e-5


class Gemver:
    def __init__(self, style, n, R=64, store=0, seed=0, rt2=None):
        self.h = C.c_void_p()
        self.n, self.R, self.store = n, R, store
        check(workloads.cs_gemver_create(C.byref(self.h), style, store, n, R, ALPHA, BETA, seed))
        if rt2:
            check(workloads.cs_gemver_set_tile(self.h, rt2))

    def run(self):
        check(workloads.cs_gemver_run(self.h))
        check(workloads.cs_gemver_wait(self.h))

    def read(self, name, dtype, count):
        out = np.empty(count, dtype)
        check(workloads.cs_gemver_read(self.h, name.encode(), out.ctypes.data, out.nbytes))
        return out

    def close(self):
        if self.h:
            check(workloads.cs_gemver_destroy(self.h))
            self.h = C.c_void_p()


def components_f32(inp, n, R):
    """The f32 restatement of the reference's four components: B, P, x, w."""
    L = O.lib()
    B = np.empty((n, n), np.float32)
    P = np.empty((n // R, n), np.float32)
    x = np.empty(n, np.float32)
    w = np.empty(n, np.float32)
    L.or_rank2_update_f32(n, R, O.ptr(inp["A"]), *[O.ptr(inp[k]) for k in ("u1", "v1", "u2", "v2")], O.ptr(B))
    L.or_gemvt_partial_f32(n, R, O.ptr(B), O.ptr(inp["y"]), O.ptr(P))
    L.or_col_reduce_f32(n, n // R, O.ptr(P), O.ptr(inp["z"]), BETA, O.ptr(x))
    L.or_gemv_f32(n, O.ptr(B), O.ptr(x), ALPHA, O.ptr(w))
    return B, P, x, w


@pytest.mark.parametrize("n,R", [(256, 64), (1024, 64), (1024, 128)])
@pytest.mark.parametrize("style", ["unfused", "fused", "naive"])
def test_reference_styles_bitwise(style, n, R):
    print("gemver", style, n, R, flush=True)
    g = Gemver(STYLES[style], n, R)
    try:
        g.run()
        inp = O.gemver_inputs(n)
        assert np.array_equal(g.read("A", np.float32, n * n).reshape(n, n), inp["A"])
        B, P, x, w = components_f32(inp, n, R)
        assert np.array_equal(g.read("B", np.float32, n * n).reshape(n, n), B)
        assert np.array_equal(g.read("P", np.float32, (n // R) * n).reshape(n // R, n), P)
        gx, gw = g.read("x", np.float32, n), g.read("w", np.float32, n)
        assert np.array_equal(gx, x)
        assert np.array_equal(gw, w)
        xr, wr = O.gemver_f64(inp, ALPHA, BETA)
        assert O.rel_l2(gx, xr) < TOL_L2 and O.rel_l2(gw, wr) < TOL_L2
        g.run()  # a second forward gives the same outputs
        assert np.array_equal(g.read("x", np.float32, n), gx) and np.array_equal(g.read("w", np.float32, n), gw)
    finally:
        g.close()


@pytest.mark.parametrize("rt2", [4, 8, 16, 32])
@pytest.mark.parametrize("n,R", [(256, 64), (1024, 64), (1040, 64), (2048, 256), (4096, 128)])
@pytest.mark.parametrize("style", ["twopass", "twopass_1"])
def test_twopass_f32(style, n, R, rt2):
    print("gemver", style, n, R, rt2, flush=True)
    g = Gemver(STYLES[style], n, R, rt2=rt2)
    try:
        g.run()
        inp = O.gemver_inputs(n)
        gx, gw = g.read("x", np.float32, n), g.read("w", np.float32, n)
        # pass 1 keeps the components' evaluation order: x equals the f32 restatement bit for bit (when R divides n)
        if n % R == 0:
            _, P, x, _ = components_f32(inp, n, R)
            assert np.array_equal(g.read("P", np.float32, (n // R) * n).reshape(n // R, n), P)
            assert np.array_equal(gx, x)
        xr, wr = O.gemver_f64(inp, ALPHA, BETA)
        assert O.rel_l2(gx, xr) < TOL_L2, O.rel_l2(gx, xr)
        assert O.rel_l2(gw, wr) < TOL_L2, O.rel_l2(gw, wr)
        assert O.rel_linf(gw, wr) < 1e-4
    finally:
        g.close()


def test_twopass_full_size():
    """N = 16384 (A is 1 GiB): the headline configuration, against the FP64 oracle."""
    n, R = 16384, 64
    print("gemver twopass full", n, flush=True)
    g = Gemver(TWOPASS, n, R)
    try:
        g.run()
        gx, gw = g.read("x", np.float32, n), g.read("w", np.float32, n)
    finally:
        g.close()
    inp = O.gemver_inputs(n)
    xr, wr = O.gemver_f64(inp, ALPHA, BETA)
    ex, ew = O.rel_l2(gx, xr), O.rel_l2(gw, wr)
    print(f"N={n} rel-L2 x {ex:.3e} w {ew:.3e}", flush=True)
    assert ex < TOL_L2 and ew < TOL_L2


@pytest.mark.parametrize("rt2", [8, 16, 32])
@pytest.mark.parametrize("store", [1, 2])
@pytest.mark.parametrize("n", [1024, 2064])
def test_twopass_reduced_precision(store, n, rt2):
    """A stored bf16 (store 1) or e4m3 with per-row f32 scale (store 2): the oracle is FP64 on the dequantised stored A;
    the quantisation itself is checked bit for bit against the CPU quantiser. Pass-2 tiles of 8 rows (spread over the
    work-group) and 16, 32 (a quarter per wave; 2064 rows leave a partial last tile at 32)."""
    print("gemver reduced", store, n, rt2, flush=True)
    g = Gemver(TWOPASS, n, 64, store=store, rt2=rt2)
    try:
        g.run()
        inp = O.gemver_inputs(n)
        if store == 1:
            q = np.empty((n, n), np.uint16)
            O.lib().or_quantize_bf16(O.ptr(inp["A"]), O.ptr(q), n, n, n, n)
            assert np.array_equal(g.read("A", np.uint16, n * n).reshape(n, n), q)
            scale = None
        else:
            q = np.empty((n, n), np.uint8)
            scale = np.empty(n, np.float32)
            O.lib().or_quantize_e4m3_rows(O.ptr(inp["A"]), O.ptr(q), O.ptr(scale), n, n, n, n)
            assert np.array_equal(g.read("A", np.uint8, n * n).reshape(n, n), q)
            assert np.array_equal(g.read("scale", np.float32, n), scale)
        gx, gw = g.read("x", np.float32, n), g.read("w", np.float32, n)
        xr, wr = O.gemver_f64(inp, ALPHA, BETA, store=store, A=q, row_scale=scale)
        assert O.rel_l2(gx, xr) < TOL_L2 and O.rel_l2(gw, wr) < TOL_L2, (O.rel_l2(gx, xr), O.rel_l2(gw, wr))
    finally:
        g.close()


def test_direct_replay_matches_graph_replay():
    """CS_REPLAY_DIRECT re-issues the captured launches on the model's stream instead of launching the graph: the same x and
    w, bit for bit, over eager, captured and replayed runs, with and without a wait between runs."""
    n = 1024
    a, b = Gemver(TWOPASS, n, 128, seed=3), Gemver(TWOPASS, n, 128, seed=3)
    try:
        check(workloads.cs_gemver_set_replay(b.h, 1))
        for i in range(6):
            if i == 4:  # three runs back to back, one wait
                for _ in range(3):
                    check(workloads.cs_gemver_run(a.h))
                    check(workloads.cs_gemver_run(b.h))
            a.run()
            b.run()
            for name in ("x", "w"):
                assert np.array_equal(a.read(name, np.float32, n), b.read(name, np.float32, n)), (i, name)
        for g in (a, b):
            st = (C.c_uint64 * 4)()
            check(workloads.cs_gemver_stats(g.h, st))
            assert st[2] == 1 and st[3] >= 6, list(st)  # one capture, the later runs replayed (by either mechanism)
        assert workloads.cs_gemver_set_replay(b.h, 7) != 0
    finally:
        a.close()
        b.close()


def test_bad_arguments_are_errors():
    h = C.c_void_p()
    assert workloads.cs_gemver_create(C.byref(h), UNFUSED, 0, 1000, 64, 1.0, 1.0, 0) == 1  # n % 256 != 0
    assert workloads.cs_gemver_create(C.byref(h), UNFUSED, 1, 1024, 64, 1.0, 1.0, 0) == 1  # bf16 needs two-pass
    assert workloads.cs_gemver_create(C.byref(h), 9, 0, 1024, 64, 1.0, 1.0, 0) == 1
    check(workloads.cs_gemver_create(C.byref(h), TWOPASS, 0, 1024, 64, 1.0, 1.0, 0))
    try:
        for rows in (0, 3, 12, 64):  # pass-2 tiles are 4, 8, 16 or 32 rows
            assert workloads.cs_gemver_set_tile(h, rows) == 1, rows
    finally:
        check(workloads.cs_gemver_destroy(h))
