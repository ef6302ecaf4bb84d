"""The incompressible NS residual on the device, fused (one fused group) and unfused, against the FP64 oracle."""
import ctypes as C

import numpy as np
import pytest

import oracle as O
from cospanmdao_amd import check, workloads

pytestmark = pytest.mark.gpu

FUSED, UNFUSED = 0, 1
NU = 1e-3
# f32 arithmetic against FP64: rel-L2 per residual field (the reference's NS gate is rel-L2 <= 1e-4,
# examples/ns-residual/main.cpp:321)
TOL_L2 = 1e-5


class Ns:
    def __init__(self, style, nx, ny, store=0, seed=0, ty=16):
        self.h = C.c_void_p()
        self.nx, self.ny, self.store = nx, ny, store
        check(workloads.cs_ns_create(C.byref(self.h), style, store, nx, ny, 1.0 / nx, NU, seed, ty))

    def run(self):
        check(workloads.cs_ns_run(self.h))
        check(workloads.cs_ns_wait(self.h))

    def read(self, name):
        if name in ("sq", "norm"):
            out = np.empty(3 if name == "sq" else 4, np.float64)
        elif name in ("u", "v", "p"):
            out = np.empty(self.nx * self.ny, np.uint16 if self.store == 1 else np.float32)
        else:
            out = np.empty(self.nx * self.ny, np.float32)
        check(workloads.cs_ns_read(self.h, name.encode(), out.ctypes.data, out.nbytes))
        return out

    def write(self, name, arr):
        arr = np.ascontiguousarray(arr)
        check(workloads.cs_ns_write(self.h, name.encode(), arr.ctypes.data, arr.nbytes))

    def close(self):
        if self.h:
            check(workloads.cs_ns_destroy(self.h))
            self.h = C.c_void_p()


def oracle_fields(seed, nx, ny):
    u, v, p = (np.empty(nx * ny, np.float32) for _ in range(3))
    O.lib().or_ns_fields(seed, nx, ny, O.ptr(u), O.ptr(v), O.ptr(p))
    return u, v, p


def to_bf16(a):
    q = np.empty(a.size, np.uint16)
    O.lib().or_quantize_bf16(O.ptr(a), O.ptr(q), 1, a.size, a.size, a.size)
    return q


def compare(g, store, fields):
    R = [g.read(k) for k in ("Ru", "Rv", "Rc")]
    stats = np.empty(6)
    sq = np.empty(3)
    O.lib().or_ns_compare(g.nx, g.ny, 1.0 / g.nx, NU, store, *[O.ptr(f) for f in fields], *[O.ptr(r) for r in R], O.ptr(stats), O.ptr(sq))
    return R, stats, sq


@pytest.mark.parametrize("style", [FUSED, UNFUSED])
# nx = 8: two live lanes, both halos the periodic wrap; nx = 1032: the last wave of a row holds two live lanes; ty = 7 and
# ty = 64 with a partial last strip: the five-row steady state, its tail, and a tile shorter than the steady state;
# 1032 x 4072: the unfused variant's 1026 partials (one per 4096 cells, the last one partial) take the norm's four-wave
# path (up to 1024 it is one wave)
@pytest.mark.parametrize("nx,ny,ty", [(64, 48, 16), (1024, 40, 16), (2048, 64, 7), (1032, 33, 32), (8, 24, 16), (1032, 130, 64),
                                      (4096, 12, 4), (1024, 520, 16), (1032, 4072, 16)])
def test_ns_small_against_fp64(style, nx, ny, ty):
    print("ns", style, nx, ny, ty, flush=True)
    g = Ns(style, nx, ny, ty=ty, seed=3)
    try:
        g.run()
        fields = oracle_fields(3, nx, ny)
        for name, f in zip("uvp", fields):
            assert np.array_equal(g.read(name), f), name  # the device generator equals the host one bit for bit
        R, stats, sq = compare(g, 0, fields)
        assert np.all(stats[0::2] < TOL_L2), stats
        gsq = g.read("sq")
        own = np.array([np.sum(r.astype(np.float64) ** 2) for r in R])
        assert np.allclose(gsq, own, rtol=1e-12), (gsq, own)  # the device norm of its own residual (FP64)
        assert np.allclose(gsq, sq, rtol=1e-5)
        norm = g.read("norm")
        assert np.allclose(norm[:3], np.sqrt(gsq), rtol=1e-14) and np.isclose(norm[3], np.sqrt(gsq.sum()), rtol=1e-14)
    finally:
        g.close()


@pytest.mark.parametrize("nx,ny,ty", [(512, 96, 16), (1032, 200, 64)])
def test_ns_fused_equals_unfused_bitwise(nx, ny, ty):
    a, b = Ns(FUSED, nx, ny, seed=5, ty=ty), Ns(UNFUSED, nx, ny, seed=5, ty=ty)
    try:
        a.run()
        b.run()
        for k in ("Ru", "Rv", "Rc"):
            assert np.array_equal(a.read(k), b.read(k)), k
    finally:
        a.close()
        b.close()


@pytest.mark.parametrize("style", [FUSED, UNFUSED])
def test_ns_uniform_field_gives_exactly_zero(style):
    # a uniform periodic flow has R = 0 exactly (examples/ns-residual/main.cpp:183-187,320)
    nx, ny = 256, 64
    g = Ns(style, nx, ny)
    try:
        g.write("u", np.full(nx * ny, 0.75, np.float32))
        g.write("v", np.full(nx * ny, -1.25, np.float32))
        g.write("p", np.full(nx * ny, 3.0, np.float32))
        g.run()
        for k in ("Ru", "Rv", "Rc"):
            assert not np.any(g.read(k)), k
        assert np.all(g.read("sq") == 0.0)
    finally:
        g.close()


def test_ns_direct_replay_matches_graph_replay():
    """CS_REPLAY_DIRECT: the same residuals and norms as graph replay, bit for bit, across a field update (update_input keeps
    the capture) and runs back to back."""
    nx, ny = 1024, 136
    a, b = Ns(FUSED, nx, ny, seed=4, ty=64), Ns(FUSED, nx, ny, seed=4, ty=64)
    try:
        check(workloads.cs_ns_set_replay(b.h, 1))
        u = np.random.default_rng(7).uniform(-1, 1, nx * ny).astype(np.float32)
        for i in range(6):
            if i == 3:
                a.write("u", u)
                b.write("u", u)
            if i == 4:
                for _ in range(3):
                    check(workloads.cs_ns_run(a.h))
                    check(workloads.cs_ns_run(b.h))
            a.run()
            b.run()
            for k in ("Ru", "Rv", "Rc", "sq", "norm"):
                assert np.array_equal(a.read(k), b.read(k)), (i, k)
        assert workloads.cs_ns_set_replay(b.h, -1) != 0
    finally:
        a.close()
        b.close()


def test_ns_graph_replay_keeps_results():
    nx, ny = 1024, 128
    g = Ns(FUSED, nx, ny)
    try:
        g.run()
        first = [g.read(k) for k in ("Ru", "Rv", "Rc", "sq")]
        for _ in range(4):  # the second identical run is captured, later ones replay the graph
            g.run()
            for k, f in zip(("Ru", "Rv", "Rc", "sq"), first):
                assert np.array_equal(g.read(k), f), k
    finally:
        g.close()


def test_ns_full_size_f32():
    """4096 x 4096, h = 1/4096, nu = 1e-3: the measured configuration."""
    n = 4096
    print("ns full", n, flush=True)
    g = Ns(FUSED, n, n, ty=16)
    try:
        g.run()
        fields = oracle_fields(0, n, n)
        assert np.array_equal(g.read("u"), fields[0])
        R, stats, sq = compare(g, 0, fields)
        print("rel-L2 Ru Rv Rc", stats[0::2], "rel-inf", stats[1::2], flush=True)
        assert np.all(stats[0::2] < TOL_L2), stats
        assert np.allclose(g.read("sq"), sq, rtol=1e-5)
    finally:
        g.close()


@pytest.mark.parametrize("nx,ny,ty", [(2048, 2048, 16), (1032, 70, 64)])
def test_ns_bf16_fields(nx, ny, ty):
    """u, v, p stored bf16 (rounded to nearest even from the f32 fields), f32 arithmetic, FP64 norms; the oracle is FP64 on
    the stored bf16 values."""
    g = Ns(FUSED, nx, ny, store=1, seed=2, ty=ty)
    try:
        g.run()
        fields = oracle_fields(2, nx, ny)
        stored = [to_bf16(f) for f in fields]
        for name, s in zip("uvp", stored):
            assert np.array_equal(g.read(name), s), name
        R, stats, sq = compare(g, 1, stored)
        assert np.all(stats[0::2] < TOL_L2), stats
    finally:
        g.close()


def test_ns_bad_arguments():
    h = C.c_void_p()
    assert workloads.cs_ns_create(C.byref(h), 0, 0, 100, 64, 0.01, NU, 0, 16) == 1  # nx % 8
    assert workloads.cs_ns_create(C.byref(h), 5, 0, 128, 64, 0.01, NU, 0, 16) == 1
    assert workloads.cs_ns_create(C.byref(h), 0, 2, 128, 64, 0.01, NU, 0, 16) == 1
