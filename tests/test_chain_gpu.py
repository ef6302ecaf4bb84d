"""The GEMVER -> BiCGK -> GEMVER chain on the device in the reference's six styles (examples/gemver-chain/systems.hpp), the
reference driver's --check contract (main.cpp:192-234: every output within 1e-4 relative L2 of the FP64 chain, B2 too for
n <= 4096; unfused, parallel and fused bit-identical), and more: every style bit for bit against the oracle's f32
restatement of its evaluation order (B1 and B2 included)."""
import ctypes as C

import numpy as np
import pytest

import oracle as O
from cospanmdao_amd import check, workloads

pytestmark = pytest.mark.gpu

STYLES = {"unfused": 0, "parallel": 1, "fused": 2, "onepass": 3, "naive": 4, "crafted": 5}
ONEPASS = {"onepass", "naive", "crafted"}  # BiCGK from one pass over B1
TOL = 1e-4  # the reference's --check tolerance (main.cpp:47)


class Chain:
    def __init__(self, style, n, R, seed=0):
        self.h = C.c_void_p()
        check(workloads.cs_chain_create(C.byref(self.h), STYLES[style], n, R, seed))
        self.n = n

    def run(self):
        check(workloads.cs_chain_run(self.h))
        check(workloads.cs_chain_wait(self.h))

    def read(self, name):
        shape = (self.n, self.n) if name.startswith("B") else (self.n,)
        out = np.empty(shape, np.float32)
        check(workloads.cs_chain_read(self.h, name.encode(), out.ctypes.data, out.nbytes))
        return out

    def outputs(self):
        return {k: self.read(k) for k in ("B1", "x1", "w1", "q", "s", "B2", "x2", "w2")}

    def stats(self):
        s = (C.c_uint64 * 4)()
        check(workloads.cs_chain_stats(self.h, s))
        return list(s)

    def close(self):
        if self.h:
            check(workloads.cs_chain_destroy(self.h))
            self.h = C.c_void_p()


def run_style(style, n, R, runs=1):
    c = Chain(style, n, R)
    try:
        for _ in range(runs):
            c.run()
        return c.outputs(), c.stats()
    finally:
        c.close()


@pytest.fixture(scope="module")
def small():
    """n = 512, R = 64: every style's outputs, the f32 restatements and the FP64 chain."""
    n, R = 512, 64
    inp = O.gemver_inputs(n)
    ab = float(O.chain_scale(n))
    ref = O.chain_f64(inp, ab, ab, want_B2=True)
    f32 = {"plain": O.chain_f32(inp, False, R), "onepass": O.chain_f32(inp, True, R), "crafted": O.chain_f32(inp, True, 16)}
    got = {}
    for style in STYLES:
        print("chain", style, "n", n, flush=True)
        got[style], _ = run_style(style, n, R)
    return {"n": n, "R": R, "ref": ref, "f32": f32, "got": got}


@pytest.mark.parametrize("style", list(STYLES))
def test_style_bitwise_against_its_f32_order(small, style):
    want = small["f32"]["crafted" if style == "crafted" else "onepass" if style in ONEPASS else "plain"]
    got = small["got"][style]
    for k in ("B1", "x1", "w1", "q", "s", "B2", "x2", "w2"):
        assert np.array_equal(got[k], want[k]), (style, k)


@pytest.mark.parametrize("style", list(STYLES))
def test_style_against_fp64_chain(small, style):
    got, ref = small["got"][style], small["ref"]
    errs = {k: O.rel_l2(got[k], ref[k]) for k in O.CHAIN_OUT}
    errs["B2"] = O.rel_l2(got["B2"].ravel(), ref["B2"].ravel())
    print(style, {k: f"{v:.2e}" for k, v in errs.items()}, flush=True)
    assert all(e <= TOL for e in errs.values()), errs


def test_same_order_styles_are_bit_identical(small):
    """The reference's check: unfused, parallel and fused sum in the same order (main.cpp:129-131)."""
    g = small["got"]
    for style in ("parallel", "fused"):
        for k in g["unfused"]:
            assert np.array_equal(g[style][k], g["unfused"][k]), (style, k)
    for k in g["onepass"]:  # naive = the same BiCGK components fused with a barrier
        assert np.array_equal(g["naive"][k], g["onepass"][k]), k
    for k in ("B1", "x1", "w1"):  # every style's G1 is the same arithmetic
        for style in ("onepass", "naive"):
            assert np.array_equal(g[style][k], g["unfused"][k]), (style, k)


def test_crafted_equals_onepass_at_its_strip_height():
    n = 1024
    a, _ = run_style("crafted", n, 64)
    b, _ = run_style("onepass", n, 16)
    for k in a:
        assert np.array_equal(a[k], b[k]), k


@pytest.mark.parametrize("style", ["unfused", "onepass", "crafted"])
def test_chain_n4096_within_tolerance(style):
    """n = 4096 (B2 included, as the reference's check does up to 4096), after a capture and replays."""
    n = 4096
    inp = O.gemver_inputs(n)
    ab = float(O.chain_scale(n))
    ref = O.chain_f64(inp, ab, ab, want_B2=True)
    got, st = run_style(style, n, 64, runs=4)
    assert st[2] >= 1 and st[3] >= 2, st  # captured, then replayed
    errs = {k: O.rel_l2(got[k], ref[k]) for k in O.CHAIN_OUT}
    errs["B2"] = O.rel_l2(got["B2"].ravel(), ref["B2"].ravel())
    print(style, {k: f"{v:.2e}" for k, v in errs.items()}, flush=True)
    assert all(e <= TOL for e in errs.values()), errs


def test_replays_give_the_eager_results():
    n = 512
    eager, _ = run_style("naive", n, 64, runs=1)
    replayed, st = run_style("naive", n, 64, runs=5)
    assert st[3] >= 3, st
    for k in eager:
        assert np.array_equal(eager[k], replayed[k]), k


def test_bad_problems_are_errors():
    h = C.c_void_p()
    assert workloads.cs_chain_create(C.byref(h), 0, 300, 64, 0) == 1  # n not a multiple of 256
    assert workloads.cs_chain_create(C.byref(h), 0, 512, 48, 0) == 1  # n not a multiple of R
    assert workloads.cs_chain_create(C.byref(h), 6, 512, 64, 0) == 1  # no such style
