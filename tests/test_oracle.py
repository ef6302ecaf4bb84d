"""The oracle pinned: against the reference's own code (oracle/_ref), its known-answer tests (tests/golden), closed forms,
exact arithmetic and invariants. CPU only."""
import ctypes as C
import json
import math
import os
from fractions import Fraction

import numpy as np
import pytest

import oracle as O

HERE = os.path.dirname(os.path.abspath(__file__))
GOLDEN = os.path.join(HERE, "golden")


def scan_pattern(n):
    # (i * 2654435761u + 7u) % 11u with i a std::size_t: 64-bit arithmetic (tests/test_device_scan.cpp:57-61)
    return np.array([(i * 2654435761 + 7) % (2 ** 64) % 11 for i in range(n)], np.uint32)


def fnv1a(a):
    h = 1469598103934665603
    for b in np.ascontiguousarray(a, np.uint32).view(np.uint8).tobytes():
        h ^= b
        h = (h * 1099511628211) % (1 << 64)
    return h


def kat_cases():
    with open(os.path.join(GOLDEN, "scan_kat.json")) as f:
        return json.load(f)["cases"]


# ---- hash generator and FP64 GEMVER against the reference's own chain_data.hpp ------------------------------------------

@pytest.mark.skipif(O.ref() is None, reason="oracle/_ref not built (the reference checkout is absent)")
def test_hash_matches_reference_generator():
    L, R = O.lib(), O.ref()
    idx = list(range(2000)) + [2 ** 32 - 1, 2 ** 32, 2 ** 32 + 5, 2 ** 40 + 7, 2 ** 63 + 11]
    for seed in range(9):
        for i in idx:
            assert L.or_hash_unit(seed, i) == R.ref_hash_unit(seed, i)
    n = 64
    a = np.empty((n, n), np.float32)
    R.ref_make_matrix_A0(n, O.ptr(a))
    assert np.array_equal(a, O.hash_matrix(1, n, n))


@pytest.mark.skipif(O.ref() is None, reason="oracle/_ref not built (the reference checkout is absent)")
def test_gemver_f64_matches_reference_chain():
    n = 256
    inp = O.gemver_inputs(n)
    outs = [np.empty(n) for _ in range(6)]
    O.ref().ref_chain(n, O.ptr(inp["A"]), *[O.ptr(inp[k]) for k in ("u1", "v1", "u2", "v2", "y", "z")],
                      *[O.ptr(o) for o in outs])
    ab = 1.0 / math.sqrt(n)
    x, w = O.gemver_f64(inp, ab, ab)
    assert np.array_equal(x, outs[0])  # x1 of the reference's chain, bit for bit
    assert np.array_equal(w, outs[1])  # w1


def test_gemver_f64_closed_forms():
    n = 96
    inp = O.gemver_inputs(n, base_seed=3)
    A = inp["A"].astype(np.float64)
    # u = v = 0: the textbook GEMV pair
    zero = dict(inp, u1=np.zeros(n, np.float32), u2=np.zeros(n, np.float32))
    x, w = O.gemver_f64(zero, 1.5, 1.2)
    xr = inp["z"] + 1.2 * (A.T @ inp["y"].astype(np.float64))
    assert np.allclose(x, xr, rtol=1e-13, atol=1e-13)
    assert np.allclose(w, 1.5 * (A @ xr), rtol=1e-12, atol=1e-12)
    # the full rank-2 update against numpy in float64
    x, w = O.gemver_f64(inp, 1.5, 1.2)
    d = {k: inp[k].astype(np.float64) for k in ("u1", "v1", "u2", "v2")}
    B = A + (np.outer(d["u1"], d["v1"]) + np.outer(d["u2"], d["v2"]))
    xr = inp["z"] + 1.2 * (B.T @ inp["y"].astype(np.float64))
    assert np.allclose(x, xr, rtol=1e-12, atol=1e-12)
    assert np.allclose(w, 1.5 * (B @ xr), rtol=1e-11, atol=1e-11)
    # y = 0: x = z exactly
    x, _ = O.gemver_f64(dict(inp, y=np.zeros(n, np.float32)), 1.5, 1.2)
    assert np.array_equal(x, inp["z"].astype(np.float64))


def round_f32(fr):
    """The f32 nearest to an exact rational, ties to even."""
    f = np.float32(float(fr))
    best = None
    for c in (np.nextafter(f, np.float32(-np.inf)), f, np.nextafter(f, np.float32(np.inf))):
        d = abs(Fraction(float(c)) - fr)
        key = (d, int(np.array(c, np.float32).view(np.uint32)) & 1)
        if best is None or key < best[0]:
            best = (key, c)
    return best[1]


def test_rank2_update_is_two_exact_fmas():
    n, R = 16, 4
    inp = O.gemver_inputs(n, base_seed=5)
    B = np.empty((n, n), np.float32)
    O.lib().or_rank2_update_f32(n, R, O.ptr(inp["A"]), *[O.ptr(inp[k]) for k in ("u1", "v1", "u2", "v2")], O.ptr(B))
    for i in range(n):
        for j in range(n):
            t = round_f32(Fraction(float(inp["u1"][i])) * Fraction(float(inp["v1"][j])) + Fraction(float(inp["A"][i, j])))
            b = round_f32(Fraction(float(inp["u2"][i])) * Fraction(float(inp["v2"][j])) + Fraction(float(t)))
            assert B[i, j] == b


def test_f32_components_against_fp64():
    n, R = 512, 64
    inp = O.gemver_inputs(n)
    L = O.lib()
    B = np.empty((n, n), np.float32)
    P = np.empty((n // R, n), np.float32)
    x = np.empty(n, np.float32)
    w = np.empty(n, np.float32)
    L.or_rank2_update_f32(n, R, O.ptr(inp["A"]), *[O.ptr(inp[k]) for k in ("u1", "v1", "u2", "v2")], O.ptr(B))
    L.or_gemvt_partial_f32(n, R, O.ptr(B), O.ptr(inp["y"]), O.ptr(P))
    L.or_col_reduce_f32(n, n // R, O.ptr(P), O.ptr(inp["z"]), 1.2, O.ptr(x))
    L.or_gemv_f32(n, O.ptr(B), O.ptr(x), 1.5, O.ptr(w))
    xr, wr = O.gemver_f64(inp, 1.5, 1.2)
    assert O.rel_l2(x, xr) < 1e-6 and O.rel_l2(w, wr) < 1e-6
    # the fused two-pass CPU form: x bitwise equal to the components' x (same order), w within tolerance
    P2 = np.empty_like(P)
    x2 = np.empty(n, np.float32)
    w2 = np.empty(n, np.float32)
    L.or_gemver_twopass_f32(n, R, O.ptr(inp["A"]), *[O.ptr(inp[k]) for k in ("u1", "v1", "u2", "v2", "y", "z")], 1.5, 1.2,
                            O.ptr(P2), O.ptr(x2), O.ptr(w2))
    assert np.array_equal(x2, x) and np.array_equal(P2, P)
    assert O.rel_l2(w2, wr) < 1e-6


def test_gemv_tree_order_is_pairwise():
    # one row, B = 1, x = the lane index pattern: the pairwise tree over 256 lanes of exactly representable values
    n = 512
    B = np.ones((n, n), np.float32)
    x = (np.arange(n) % 7).astype(np.float32)
    y = np.empty(n, np.float32)
    O.lib().or_gemv_f32(n, O.ptr(B), O.ptr(x), 1.0, O.ptr(y))
    assert np.all(y == np.float32(x.astype(np.float64).sum()))
    # a sum whose value depends on the order: 2^24 + 1 + 1 ... pins the tree (pairs (0,1), then ((0,1),(2,3)), ...)
    x = np.zeros(n, np.float32)
    x[0], x[2], x[3] = 2.0 ** 24, 1.0, 1.0  # lanes 0, 2, 3 (one element each)
    O.lib().or_gemv_f32(n, O.ptr(B), O.ptr(x), 1.0, O.ptr(y))
    # (2^24 + 0) + (1 + 1) = 2^24 + 2 exactly under the tree; a left-to-right sum gives 2^24 (each +1 ties to even)
    assert y[0] == np.float32(2.0 ** 24 + 2)


# ---- scans and ragged columns ----------------------------------------------------------------------------------------------

@pytest.mark.parametrize("case", kat_cases(), ids=lambda c: c["name"])
def test_scan_known_answers(case):
    v = scan_pattern(case["n"]) if case["input"] == "pattern" else np.array(case["values"], np.uint32)
    out = O.exclusive_scan_u32(v)
    if "expected" in case:
        assert out.tolist() == case["expected"]
    else:
        assert fnv1a(out) == int(case["expected_fnv1a"]) and int(out[-1]) == case["expected_total"]
        for i, e in case["probes"]:
            assert int(out[i]) == e


def test_scan_wraps_mod_2_32():
    v = np.array([2 ** 31, 2 ** 31, 5], np.uint32)
    assert O.exclusive_scan_u32(v).tolist() == [0, 2 ** 31, 0, 5]


def test_build_column_counts_mode():
    # tests/test_shapes.cpp:656-682 of the reference: counts {2, 0, 3, 1} -> offsets {0, 2, 2, 5, 6}, total 6
    for dt in (np.int32, np.uint32, np.int64, np.uint64):
        ext, offs, total, flag = O.build_column(np.array([2, 0, 3, 1], dt))
        assert ext.tolist() == [2, 0, 3, 1] and offs.tolist() == [0, 2, 2, 5, 6] and total == 6 and flag == 0
    # a rank-2 source [N, k]: column 0
    src = np.array([[2, 9], [0, 9], [3, 9], [1, 9]], np.int32)
    ext, offs, total, flag = O.build_column(src, k=2)
    assert offs.tolist() == [0, 2, 2, 5, 6] and flag == 0
    # negative: bit 1, the extent is 0
    ext, offs, total, flag = O.build_column(np.array([2, -1, 3], np.int32))
    assert flag == 1 and ext.tolist() == [2, 0, 3] and total == 5
    # past 32 bits: bit 2
    ext, offs, total, flag = O.build_column(np.array([1, 2 ** 32, 1], np.uint64))
    assert flag == 2 and ext.tolist() == [1, 0, 1]
    # a wrap of the running sum: bit 2
    ext, offs, total, flag = O.build_column(np.array([2 ** 31, 2 ** 31, 1], np.uint32))
    assert flag == 2
    # empty
    ext, offs, total, flag = O.build_column(np.zeros(0, np.int32))
    assert offs.tolist() == [0] and total == 0 and flag == 0


def test_build_column_offsets_mode():
    ext, offs, total, flag = O.build_column(np.array([0, 2, 2, 5, 6], np.int64), offsets_mode=True)
    assert ext.tolist() == [2, 0, 3, 1] and offs.tolist() == [0, 2, 2, 5, 6] and total == 6 and flag == 0
    # offs[0] != 0: bit 1
    assert O.build_column(np.array([1, 2, 3], np.uint32), offsets_mode=True)[3] == 1
    # a decrease: bit 1, extent 0
    ext, offs, total, flag = O.build_column(np.array([0, 4, 3], np.int32), offsets_mode=True)
    assert flag == 1 and ext.tolist() == [4, 0]
    # a negative entry: bit 1
    assert O.build_column(np.array([0, -3, 3], np.int32), offsets_mode=True)[3] & 1


# ---- storage formats -------------------------------------------------------------------------------------------------------

def test_bf16_round_trip_and_rounding():
    L = O.lib()
    for b in range(0, 1 << 16, 7):
        f = L.or_bf16_to_f32(b)
        if math.isnan(f):
            assert math.isnan(L.or_bf16_to_f32(L.or_f32_to_bf16(f)))
        else:
            assert L.or_f32_to_bf16(f) == b
    # ties to even: 1 + 2^-8 is halfway between 1 and 1 + 2^-7
    assert L.or_f32_to_bf16(1.0 + 2.0 ** -8) == L.or_f32_to_bf16(1.0)
    assert L.or_f32_to_bf16(1.0 + 3 * 2.0 ** -8) == L.or_f32_to_bf16(1.0 + 2.0 ** -6)


def e4m3_spec(b):
    """OCP E4M3 decode from the format's definition."""
    s = -1.0 if b & 0x80 else 1.0
    e, m = (b >> 3) & 0xF, b & 7
    if e == 0xF and m == 7:
        return float("nan")
    if e == 0:
        return s * m * 2.0 ** -9
    return s * (1 + m / 8) * 2.0 ** (e - 7)


def test_e4m3_every_code():
    L = O.lib()
    for b in range(256):
        want = e4m3_spec(b)
        got = L.or_e4m3_to_f32(b)
        if math.isnan(want):
            assert math.isnan(got)
            continue
        assert got == want
        if b != 0x80:  # -0 encodes as -0 (0x80), fine either way
            assert L.or_f32_to_e4m3(want) == b
    assert L.or_e4m3_to_f32(0x7E) == 448.0
    assert L.or_f32_to_e4m3(1e9) == 0x7E and L.or_f32_to_e4m3(-1e9) == 0xFE
    # ties to even: 1 + 1/16 is halfway between 1 (m = 0) and 1 + 1/8 (m = 1) -> 1
    assert L.or_f32_to_e4m3(1.0625) == L.or_f32_to_e4m3(1.0)
    assert L.or_f32_to_e4m3(1.1875) == L.or_f32_to_e4m3(1.25)
    # the smallest subnormal and half of it (ties to even: 0)
    assert L.or_f32_to_e4m3(2.0 ** -9) == 1 and L.or_f32_to_e4m3(2.0 ** -10) == 0


def test_e4m3_nearest_is_nearest():
    L = O.lib()
    codes = sorted({e4m3_spec(b) for b in range(256) if not math.isnan(e4m3_spec(b))})
    rng = np.random.default_rng(0)
    for v in rng.uniform(-460, 460, 2000).astype(np.float32):
        q = L.or_e4m3_to_f32(L.or_f32_to_e4m3(float(v)))
        d = min(abs(c - float(v)) for c in codes)
        assert abs(q - float(v)) == d or abs(float(v)) > 448


def test_quantize_e4m3_rows_scale():
    a = O.hash_matrix(9, 8, 100)
    a[3] = 0.0
    q = np.empty((8, 100), np.uint8)
    s = np.empty(8, np.float32)
    O.lib().or_quantize_e4m3_rows(O.ptr(a), O.ptr(q), O.ptr(s), 8, 100, 100, 100)
    assert s[3] == 1.0 and np.all(q[3] == 0)
    for i in (0, 5):
        assert s[i] == np.float32(np.abs(a[i]).max()) / np.float32(448.0)
        assert np.abs(q[i].astype(np.int32) & 0x7F).max() == 0x7E  # the row maximum maps to 448


# ---- NS residual ----------------------------------------------------------------------------------------------------------

def test_ns_uniform_field_gives_zero():
    nx = ny = 32
    u = np.full(nx * ny, 0.75, np.float32)
    v = np.full(nx * ny, -1.25, np.float32)
    p = np.full(nx * ny, 3.0, np.float32)
    R = [np.empty(nx * ny) for _ in range(3)]
    sq = np.empty(3)
    O.lib().or_ns_residual_f64(nx, ny, 1.0 / nx, 1e-3, 0, O.ptr(u), O.ptr(v), O.ptr(p), *[O.ptr(r) for r in R], O.ptr(sq))
    assert all(np.all(r == 0.0) for r in R) and np.all(sq == 0.0)


def test_ns_single_mode_matches_analytic_derivatives():
    n = 256
    h = 1.0 / n
    nu = 1e-3
    jj, ii = np.meshgrid(np.arange(n), np.arange(n))
    X, Y = jj * h, ii * h
    k = 2 * np.pi
    u = np.cos(k * X) * np.sin(k * Y)
    v = -np.sin(k * X) * np.cos(k * Y)
    p = 0.25 * (np.cos(2 * k * X) + np.cos(2 * k * Y))
    uf, vf, pf = (a.astype(np.float32).ravel() for a in (u, v, p))
    R = [np.empty(n * n) for _ in range(3)]
    O.lib().or_ns_residual_f64(n, n, h, nu, 0, O.ptr(uf), O.ptr(vf), O.ptr(pf), *[O.ptr(r) for r in R], None)
    U, V, Pp = (a.astype(np.float32).astype(np.float64) for a in (u, v, p))
    ux = -k * np.sin(k * X) * np.sin(k * Y)
    uy = k * np.cos(k * X) * np.cos(k * Y)
    vx = -k * np.cos(k * X) * np.cos(k * Y)
    vy = k * np.sin(k * X) * np.sin(k * Y)
    px = -0.5 * k * np.sin(2 * k * X)
    py = -0.5 * k * np.sin(2 * k * Y)
    lu, lv = -2 * k * k * u, -2 * k * k * v
    Ru = U * ux + V * uy + px - nu * lu
    Rv = U * vx + V * vy + py - nu * lv
    Rc = ux + vy
    for got, want in zip(R, (Ru, Rv, Rc)):
        err = np.abs(got.reshape(n, n) - want).max()
        assert err < 2e-3 * max(1.0, np.abs(want).max())  # O(h^2) truncation (k h ~ 0.025) plus f32 rounding of the inputs


def test_ns_compare_agrees_with_residual():
    nx, ny = 64, 48
    u, v, p = (np.empty(nx * ny, np.float32) for _ in range(3))
    O.lib().or_ns_fields(7, nx, ny, O.ptr(u), O.ptr(v), O.ptr(p))
    R = [np.empty(nx * ny) for _ in range(3)]
    sq = np.empty(3)
    O.lib().or_ns_residual_f64(nx, ny, 1.0 / nx, 1e-3, 0, O.ptr(u), O.ptr(v), O.ptr(p), *[O.ptr(r) for r in R], O.ptr(sq))
    Rf = [r.astype(np.float32) for r in R]
    stats = np.empty(6)
    sq2 = np.empty(3)
    O.lib().or_ns_compare(nx, ny, 1.0 / nx, 1e-3, 0, O.ptr(u), O.ptr(v), O.ptr(p), *[O.ptr(r) for r in Rf], O.ptr(stats), O.ptr(sq2))
    assert np.all(stats < 1e-7) and np.allclose(sq, sq2, rtol=1e-12)
    assert np.allclose(sq, [np.sum(r * r) for r in R], rtol=1e-10)
    # the fields are not trivial: every field has spread
    assert min(u.std(), v.std(), p.std()) > 1e-3


def test_ragged_sizes_bounds():
    s = np.empty(512, np.uint32)
    O.lib().or_ragged_sizes(0, 512, 64, 4096, O.ptr(s))
    assert s.min() >= 64 and s.max() <= 4096
    logs = np.log(s.astype(np.float64) / 64) / np.log(64)
    assert 0.4 < logs.mean() < 0.6 and logs.min() < 0.05 and logs.max() > 0.95


def test_strip_partials_match_components_and_ragged_strips():
    L = O.lib()
    n, R = 256, 64
    inp = O.gemver_inputs(n, base_seed=2)
    B = np.empty((n, n), np.float32)
    P = np.empty((n // R, n), np.float32)
    L.or_rank2_update_f32(n, R, O.ptr(inp["A"]), *[O.ptr(inp[k]) for k in ("u1", "v1", "u2", "v2")], O.ptr(B))
    L.or_gemvt_partial_f32(n, R, O.ptr(B), O.ptr(inp["y"]), O.ptr(P))
    P2 = np.empty_like(P)
    L.or_strip_partials_f32(n, n, R, O.ptr(inp["A"]), *[O.ptr(inp[k]) for k in ("u1", "v1", "u2", "v2", "y")], O.ptr(P2))
    assert np.array_equal(P, P2)
    # a ragged n: the last strip is shorter; the strip sums equal the FP64 column sums of the updated matrix
    n = 150
    inp = O.gemver_inputs(n, base_seed=4)
    P3 = np.empty(((n + R - 1) // R, n), np.float32)
    L.or_strip_partials_f32(n, n, R, O.ptr(inp["A"]), *[O.ptr(inp[k]) for k in ("u1", "v1", "u2", "v2", "y")], O.ptr(P3))
    d = {k: inp[k].astype(np.float64) for k in ("u1", "v1", "u2", "v2", "y")}
    Bd = inp["A"].astype(np.float64) + np.outer(d["u1"], d["v1"]) + np.outer(d["u2"], d["v2"])
    want = (Bd * d["y"][:, None]).sum(axis=0)
    assert np.allclose(P3.astype(np.float64).sum(axis=0), want, rtol=1e-5, atol=1e-4)


def test_hash_with_base_is_the_flat_stream():
    L = O.lib()
    full = O.hash_matrix(11, 5, 40)
    part = np.empty((3, 40), np.float32)
    L.or_fill_hash_f32_base(11, O.ptr(part), 3, 40, 40, 2 * 40)
    assert np.array_equal(part, full[2:])


# ---- the GEMVER -> BiCGK -> GEMVER chain --------------------------------------------------------------------------------
@pytest.mark.skipif(O.ref() is None, reason="oracle/_ref not built (the reference checkout is absent)")
@pytest.mark.parametrize("n", [256, 512])
def test_chain_f64_matches_reference_chain(n):
    """or_chain_f64 restates chain_data.hpp:88-147; the reference's own code (oracle/_ref) gives the same bits."""
    inp = O.gemver_inputs(n)
    ab = 1.0 / math.sqrt(n)
    want = {k: np.empty(n) for k in O.CHAIN_OUT}
    O.ref().ref_chain(n, O.ptr(inp["A"]), *[O.ptr(inp[k]) for k in ("u1", "v1", "u2", "v2", "y", "z")],
                      *[O.ptr(want[k]) for k in O.CHAIN_OUT])
    got = O.chain_f64(inp, ab, ab, want_B2=True)
    for k in O.CHAIN_OUT:
        assert np.array_equal(got[k], want[k]), k
    # B2 = A0 + 2 (u1 v1^T + u2 v2^T), the closed form of G2's update of B1
    u1, v1, u2, v2 = (inp[k].astype(np.float64) for k in ("u1", "v1", "u2", "v2"))
    B2 = inp["A"].astype(np.float64) + 2.0 * (np.outer(u1, v1) + np.outer(u2, v2))
    assert np.array_equal(got["B2"], B2)


def test_chain_f64_closed_form_bicgk():
    """q = B1 x1 and s = B1^T w1 against numpy's dense products in FP64 (tolerance: another summation order)."""
    n = 256
    inp = O.gemver_inputs(n, base_seed=3)
    got = O.chain_f64(inp, 0.5, 0.25)
    u1, v1, u2, v2 = (inp[k].astype(np.float64) for k in ("u1", "v1", "u2", "v2"))
    B1 = inp["A"].astype(np.float64) + np.outer(u1, v1) + np.outer(u2, v2)
    x1 = inp["z"] + 0.25 * (B1.T @ inp["y"].astype(np.float64))
    w1 = 0.5 * (B1 @ x1)
    assert np.allclose(got["x1"], x1, rtol=1e-12, atol=1e-12) and np.allclose(got["w1"], w1, rtol=1e-12, atol=1e-12)
    assert np.allclose(got["q"], B1 @ got["x1"], rtol=1e-12, atol=1e-10)
    assert np.allclose(got["s"], B1.T @ got["w1"], rtol=1e-12, atol=1e-10)
    B2 = B1 + np.outer(u1, v1) + np.outer(u2, v2)
    x2 = got["s"] + 0.25 * (B2.T @ got["q"])
    assert np.allclose(got["x2"], x2, rtol=1e-12, atol=1e-9)
    assert np.allclose(got["w2"], 0.5 * (B2 @ got["x2"]), rtol=1e-12, atol=1e-9)


def test_bicgk_onepass_on_integers_is_exact():
    """On small integers every sum is exact in f32, so the one-pass partials must equal the exact block sums: pins which
    rows and columns feed each entry of Ps and Pq, whatever the order."""
    n, R = 512, 64
    rng = np.random.default_rng(7)
    B = rng.integers(-8, 9, (n, n)).astype(np.float32)
    p = rng.integers(-4, 5, n).astype(np.float32)
    r = rng.integers(-4, 5, n).astype(np.float32)
    Ps = np.empty((n // R, n), np.float32)
    Pq = np.empty((n, n // 256), np.float32)
    O.lib().or_bicgk_onepass_f32(n, R, O.ptr(B), O.ptr(p), O.ptr(r), O.ptr(Ps), O.ptr(Pq))
    Bi, pi, ri = B.astype(np.int64), p.astype(np.int64), r.astype(np.int64)
    assert np.array_equal(Ps.astype(np.int64), (Bi * ri[:, None]).reshape(n // R, R, n).sum(axis=1))
    assert np.array_equal(Pq.astype(np.int64), (Bi * pi[None, :]).reshape(n, n // 256, 256).sum(axis=2))
    q = np.empty(n, np.float32)
    O.lib().or_row_reduce_f32(n, n // 256, O.ptr(Pq), O.ptr(q))
    assert np.array_equal(q.astype(np.int64), Bi @ pi)


@pytest.mark.parametrize("onepass,R", [(False, 64), (True, 64), (True, 16)])
def test_chain_f32_against_fp64(onepass, R):
    """The f32 restatements of every style's evaluation order stay within the reference's tolerance of the FP64 chain."""
    n = 512
    inp = O.gemver_inputs(n)
    ab = O.chain_scale(n)
    want = O.chain_f64(inp, float(ab), float(ab), want_B2=True)
    got = O.chain_f32(inp, onepass, R)
    for k in O.CHAIN_OUT:
        assert O.rel_l2(got[k], want[k]) < 1e-5, k
    assert O.rel_l2(got["B2"].ravel(), want["B2"].ravel()) < 1e-6


# ---- the error metric against the reference's own relative_l2 (chain_data.hpp:147-161) ----------------------------------

def _rel_cases():
    rng = np.random.default_rng(11)
    cases = []
    for n in (1, 2, 7, 1000, 4097):
        want = rng.standard_normal(n)
        got = (want + rng.standard_normal(n) * 10.0 ** rng.integers(-9, 0)).astype(np.float32)
        cases.append((f"random n={n}", got, want))
    cases.append(("both zero", np.zeros(5, np.float32), np.zeros(5)))
    cases.append(("want zero, got not", np.array([0, 1e-30, 0], np.float32), np.zeros(3)))
    cases.append(("NaN in got", np.array([1, np.nan, 3], np.float32), np.array([1.0, 2.0, 3.0])))
    cases.append(("inf in got", np.array([1, np.inf, 3], np.float32), np.array([1.0, 2.0, 3.0])))
    cases.append(("exact", np.array([1.5, -2.25, 1e-20], np.float32), np.array([1.5, -2.25, float(np.float32(1e-20))])))
    cases.append(("got zero", np.zeros(4, np.float32), np.array([3.0, -4.0, 0.0, 12.0])))
    cases.append(("empty", np.zeros(0, np.float32), np.zeros(0)))
    cases.append(("huge and tiny", np.array([1e30, 1e-30], np.float32), np.array([1e30, 2e-30])))
    return cases


@pytest.mark.skipif(O.ref() is None, reason="oracle/_ref not built (the reference checkout is absent)")
@pytest.mark.parametrize("name,got,want", _rel_cases(), ids=[c[0] for c in _rel_cases()])
def test_relative_l2_matches_reference(name, got, want):
    """or_relative_l2_f32 / _f64 give the reference's relative_l2 bit for bit (the same value, or both infinite) on
    random vectors, two zero vectors (0), a zero want (infinity), NaN or infinity in got, empty vectors."""
    L, R = O.lib(), O.ref()
    got = np.ascontiguousarray(got, np.float32)
    want = np.ascontiguousarray(want, np.float64)
    g64 = got.astype(np.float64)
    ours = L.or_relative_l2_f32(O.ptr(got), O.ptr(want), got.size)
    ref = R.ref_relative_l2(O.ptr(got), O.ptr(want), got.size)
    ours64 = L.or_relative_l2_f64(O.ptr(g64), O.ptr(want), got.size)
    ref64 = R.ref_relative_l2_f64(O.ptr(g64), O.ptr(want), got.size)
    for a, b in ((ours, ref), (ours64, ref64)):
        assert (math.isinf(a) and math.isinf(b)) or np.float64(a).tobytes() == np.float64(b).tobytes(), (name, a, b)


def test_relative_l2_closed_forms():
    """Values the definition fixes: got = (1 + e) want gives e; got = 0 gives 1; one wrong entry among equal ones gives
    its share; a NaN anywhere gives infinity (the reference's rule); the f32 and f64 forms agree on f32-exact data."""
    L = O.lib()
    want = np.array([3.0, 4.0])  # ||want|| = 5
    for got, expect in (([3.0, 4.0], 0.0), ([0.0, 0.0], 1.0), ([3.0, 4.5], 0.1), ([6.0, 8.0], 1.0), ([-3.0, -4.0], 2.0)):
        g = np.array(got, np.float32)
        assert L.or_relative_l2_f32(O.ptr(g), O.ptr(want), 2) == pytest.approx(expect, abs=1e-15)
        g64 = np.array(got)
        assert L.or_relative_l2_f64(O.ptr(g64), O.ptr(want), 2) == pytest.approx(expect, abs=1e-15)
    assert O.rel_linf(np.array([3.0, 4.5], np.float32), want) == pytest.approx(0.125)
    assert O.rel_linf(np.array([3.0, 4.0], np.float32), want) == 0.0
    assert math.isinf(O.rel_linf(np.array([np.nan, 4.0], np.float32), want))
    assert math.isinf(O.rel_linf(np.array([1.0], np.float32), np.zeros(1)))


# ---- the NS fields: what the Fourier transform fixes ----------------------------------------------------------------------

def test_ns_fields_spectrum():
    """u, v, p = sum of 8 modes a_m / |k_m|^2 cos(2 pi (kx x + ky y) + phi_m), integer k in [-4, 4]^2 minus 0, a_m ~ N(0, 1),
    phi_m ~ U[0, 2 pi). Over 50 seeds x 3 fields the FFT of each field must have: all its energy on integer wavenumbers
    inside the box (a non-integer wavenumber or a non-periodic field leaks across every bin), no mean, at most 8
    conjugate pairs and about 7.3 on average (8 draws from 40 pairs; one mode in place of eight gives 1), sum over pairs of
    |c_k|^2 |k|^4 / 8 with mean 1 (the 1/|k|^2 law with unit-variance amplitudes; 1/|k| gives about 13, 1/|k|^3 about
    0.2), and phases spread uniformly."""
    nx, ny = 64, 48
    pairs_per_field, law, phases = [], [], []
    for seed in range(50):
        fields = [np.empty(nx * ny, np.float32) for _ in range(3)]
        O.lib().or_ns_fields(seed, nx, ny, *[O.ptr(f) for f in fields])
        for f in fields:
            F = np.fft.fft2(f.reshape(ny, nx).astype(np.float64)) / (nx * ny)
            total = np.sum(np.abs(F) ** 2)
            ky = np.fft.fftfreq(ny, 1.0 / ny).round().astype(int)[:, None] * np.ones((1, nx), int)
            kx = np.fft.fftfreq(nx, 1.0 / nx).round().astype(int)[None, :] * np.ones((ny, 1), int)
            inside = (np.abs(kx) <= 4) & (np.abs(ky) <= 4)
            assert np.sum(np.abs(F[~inside]) ** 2) <= 1e-12 * total, "energy outside the integer wavenumber box"
            assert abs(F[0, 0]) ** 2 <= 1e-12 * total, "the field has a mean"
            live = np.abs(F) ** 2 > 1e-10 * total
            # one representative per conjugate pair: ky > 0, or ky == 0 and kx > 0
            rep = live & ((ky > 0) | ((ky == 0) & (kx > 0)))
            pairs = int(np.sum(rep))
            assert 1 <= pairs <= 8, pairs
            assert int(np.sum(live)) == 2 * pairs, "bins must come in conjugate pairs (a real field)"
            pairs_per_field.append(pairs)
            c = 2.0 * F[rep]  # a cos(theta + phi) puts a e^{i phi} / 2 on bin +k
            k2 = (kx[rep] ** 2 + ky[rep] ** 2).astype(np.float64)
            law.append(float(np.sum(np.abs(c) ** 2 * k2 ** 2)) / 8.0)
            phases.extend(list(c / np.abs(c)))
    mean_pairs = float(np.mean(pairs_per_field))
    assert 6.8 <= mean_pairs <= 7.8, mean_pairs
    assert 0.8 <= float(np.mean(law)) <= 1.2, float(np.mean(law))
    assert abs(np.mean(phases)) < 0.1, abs(np.mean(phases))
