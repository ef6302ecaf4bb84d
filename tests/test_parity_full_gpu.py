"""Parity at the full sizes of the measured configurations, with the per-output errors against the FP64 oracle recorded
(tests/helpers.py record_errors; the committed summary is profiles/parity_errors_full_size.md):

  * GEMVER N = 16384 f32, fused two-pass (pass-2 tiles of 16 and 4 rows) and the reference's four unfused components
    (B, P, x, w bit for bit against the f32 restatement of the components, x and w within 1e-5 rel-L2 of FP64);
  * GEMVER N = 32768 with A stored f32, bf16 and e4m3 + per-row f32 scale: the stored A bit for bit against the CPU
    quantisers, x and w within 1e-5 rel-L2 of FP64 on the dequantised stored A (the bar), at every strip height R and
    pass-2 tile the benchmark picks from; the quantisation error (FP64 on the original f32 A) recorded beside it;
  * the NS residual at 4096^2 (f32, fused and unfused), 8192^2 (f32) and 8192^2 with bf16 fields: Ru, Rv, Rc within 1e-5
    rel-L2 of FP64 (on the stored fields), the FP64 norms within 1e-5 relative, the quantisation error of bf16 fields
    recorded.

Errors are rel-L2 (the reference's metric, examples/gemver-chain/common/chain_data.hpp:147-161) and
||got - want||_inf / ||want||_inf."""
import ctypes as C

import numpy as np
import pytest

import oracle as O
from cospanmdao_amd import check, workloads
from helpers import record_errors

pytestmark = pytest.mark.gpu

UNFUSED, TWOPASS = 0, 3
ALPHA, BETA = 1.5, 1.2
TOL_L2 = 1e-5
NU = 1e-3


def errors(got, want):
    return {"rel_l2": O.rel_l2(got, want), "rel_inf": O.rel_linf(got, want)}


class Gemver:
    def __init__(self, style, n, R, store=0, rt2=None):
        self.h = C.c_void_p()
        check(workloads.cs_gemver_create(C.byref(self.h), style, store, n, R, ALPHA, BETA, 0))
        if rt2:
            check(workloads.cs_gemver_set_tile(self.h, rt2))

    def run(self, times=1):
        for _ in range(times):
            check(workloads.cs_gemver_run(self.h))
        check(workloads.cs_gemver_wait(self.h))

    def read(self, name, dtype, count):
        out = np.empty(count, dtype)
        check(workloads.cs_gemver_read(self.h, name.encode(), out.ctypes.data, out.nbytes))
        return out

    def close(self):
        check(workloads.cs_gemver_destroy(self.h))


def test_gemver_f32_n16384_fused_and_unfused():
    n = 16384
    inp = O.gemver_inputs(n)
    xr, wr = O.gemver_f64(inp, ALPHA, BETA)
    rec = {}
    print("gemver twopass n", n, flush=True)
    g = Gemver(TWOPASS, n, 128, rt2=16)
    try:
        g.run(times=3)  # eager, captured, replayed
        gx, gw = g.read("x", np.float32, n), g.read("w", np.float32, n)
    finally:
        g.close()
    rec["twopass"] = {"x": errors(gx, xr), "w": errors(gw, wr), "R": 128, "pass2_rows": 16}
    g = Gemver(TWOPASS, n, 128, rt2=4)
    try:
        g.run(times=3)
        x4, w4 = g.read("x", np.float32, n), g.read("w", np.float32, n)
    finally:
        g.close()
    assert np.array_equal(x4, gx), "x does not depend on the pass-2 tile"
    rec["twopass_pass2_rows_4"] = {"w": errors(w4, wr)}
    print("gemver unfused n", n, flush=True)
    g = Gemver(UNFUSED, n, 64)
    try:
        g.run(times=3)
        B = g.read("B", np.float32, n * n).reshape(n, n)
        P = g.read("P", np.float32, (n // 64) * n).reshape(n // 64, n)
        ux, uw = g.read("x", np.float32, n), g.read("w", np.float32, n)
    finally:
        g.close()
    L = O.lib()
    Bf = np.empty((n, n), np.float32)
    L.or_rank2_update_f32(n, 64, O.ptr(inp["A"]), *[O.ptr(inp[k]) for k in ("u1", "v1", "u2", "v2")], O.ptr(Bf))
    assert np.array_equal(B, Bf)
    del Bf
    Pf = np.empty((n // 64, n), np.float32)
    L.or_gemvt_partial_f32(n, 64, O.ptr(B), O.ptr(inp["y"]), O.ptr(Pf))
    assert np.array_equal(P, Pf)
    xf, wf = np.empty(n, np.float32), np.empty(n, np.float32)
    L.or_col_reduce_f32(n, n // 64, O.ptr(P), O.ptr(inp["z"]), BETA, O.ptr(xf))
    L.or_gemv_f32(n, O.ptr(B), O.ptr(ux), ALPHA, O.ptr(wf))
    assert np.array_equal(ux, xf) and np.array_equal(uw, wf)
    rec["unfused"] = {"x": errors(ux, xr), "w": errors(uw, wr), "R": 64, "bitwise_vs_f32_components": True}
    record_errors("gemver_n16384_f32", rec)
    print(rec, flush=True)
    for style in ("twopass", "unfused"):
        assert rec[style]["x"]["rel_l2"] < TOL_L2 and rec[style]["w"]["rel_l2"] < TOL_L2, rec


@pytest.mark.parametrize("store,name", [(0, "f32"), (1, "bf16"), (2, "e4m3")])
def test_gemver_n32768_reduced_precision_storage(store, name):
    n = 32768
    inp = O.gemver_inputs(n)
    L = O.lib()
    if store == 0:  # the f32 baseline: the stored A is the generated one
        q, scale = inp["A"], None
    elif store == 1:
        q = np.empty((n, n), np.uint16)
        L.or_quantize_bf16(O.ptr(inp["A"]), O.ptr(q), n, n, n, n)
        scale = None
    else:
        q = np.empty((n, n), np.uint8)
        scale = np.empty(n, np.float32)
        L.or_quantize_e4m3_rows(O.ptr(inp["A"]), O.ptr(q), O.ptr(scale), n, n, n, n)
    rec = {"n": n, "storage": name}
    outs = {}
    for R, rt2 in [(R, t) for R in (128, 256, 512) for t in (4, 8, 16, 32)]:  # every pair the benchmark chooses among
        print("gemver", name, "n", n, "R", R, "rt2", rt2, flush=True)
        g = Gemver(TWOPASS, n, R, store=store, rt2=rt2)
        try:
            g.run(times=3)
            if (R, rt2) == (128, 4):
                assert np.array_equal(g.read("A", q.dtype, n * n).reshape(n, n), q), "stored A differs from the CPU quantiser"
                if scale is not None:
                    assert np.array_equal(g.read("scale", np.float32, n), scale)
            outs[(R, rt2)] = (g.read("x", np.float32, n), g.read("w", np.float32, n))
        finally:
            g.close()
    for R, rt2 in outs:
        assert np.array_equal(outs[(R, 4)][0], outs[(R, rt2)][0]), ("x does not depend on the pass-2 tile", R, rt2)
    xs, ws = O.gemver_f64(inp, ALPHA, BETA, store=store, A=q, row_scale=scale)  # FP64 on the dequantised stored A
    del q
    xo, wo = O.gemver_f64(inp, ALPHA, BETA) if store else (xs, ws)  # FP64 on the original f32 A
    for (R, rt2), (gx, gw) in outs.items():
        rec[f"R{R}_pass2_rows_{rt2}"] = {"x": errors(gx, xs), "w": errors(gw, ws)}
    gx, gw = outs[(128, 4)]
    rec["quantisation_error_vs_original_A"] = {"x": errors(gx, xo), "w": errors(gw, wo)}
    record_errors(f"gemver_n32768_{name}", rec)
    print(rec, flush=True)
    for key in outs:
        e = rec[f"R{key[0]}_pass2_rows_{key[1]}"]
        assert e["x"]["rel_l2"] < TOL_L2 and e["w"]["rel_l2"] < TOL_L2, rec
    # the storage's own rounding: bf16 keeps 8 significant bits, e4m3 4 (sanity bounds, not the bar)
    assert rec["quantisation_error_vs_original_A"]["x"]["rel_l2"] < (5e-3 if store <= 1 else 5e-2), rec


class Ns:
    def __init__(self, style, n, store=0, ty=32):
        self.h = C.c_void_p()
        self.n, self.store = n, store
        check(workloads.cs_ns_create(C.byref(self.h), style, store, n, n, 1.0 / n, NU, 0, ty))

    def run(self, times=1):
        for _ in range(times):
            check(workloads.cs_ns_run(self.h))
        check(workloads.cs_ns_wait(self.h))

    def read(self, name):
        cells = self.n * self.n
        if name in ("sq", "norm"):
            out = np.empty(3 if name == "sq" else 4, np.float64)
        elif name in ("u", "v", "p"):
            out = np.empty(cells, np.uint16 if self.store == 1 else np.float32)
        else:
            out = np.empty(cells, np.float32)
        check(workloads.cs_ns_read(self.h, name.encode(), out.ctypes.data, out.nbytes))
        return out

    def close(self):
        check(workloads.cs_ns_destroy(self.h))


def ns_fields(n):
    u, v, p = (np.empty(n * n, np.float32) for _ in range(3))
    O.lib().or_ns_fields(0, n, n, O.ptr(u), O.ptr(v), O.ptr(p))
    return [u, v, p]


def ns_errors(n, store, fields, R, norm):
    """rel-L2 and rel-inf of Ru, Rv, Rc and the relative error of the four FP64 norms against FP64 on `fields`."""
    stats, sq = np.empty(6), np.empty(3)
    O.lib().or_ns_compare(n, n, 1.0 / n, NU, store, *[O.ptr(f) for f in fields], *[O.ptr(r) for r in R], O.ptr(stats), O.ptr(sq))
    want = np.array([np.sqrt(sq[0]), np.sqrt(sq[1]), np.sqrt(sq[2]), np.sqrt(sq.sum())])
    out = {k: {"rel_l2": float(stats[2 * i]), "rel_inf": float(stats[2 * i + 1])} for i, k in enumerate(("Ru", "Rv", "Rc"))}
    out["norms_rel"] = [float(abs(norm[i] - want[i]) / want[i]) for i in range(4)]
    return out


@pytest.mark.parametrize("n,store,styles", [(4096, 0, ("fused", "unfused")), (8192, 0, ("fused",)), (8192, 1, ("fused",))])
def test_ns_residual_full_sizes(n, store, styles):
    fields = ns_fields(n)
    stored = fields
    if store == 1:
        stored = []
        for f in fields:
            s = np.empty(f.size, np.uint16)
            O.lib().or_quantize_bf16(O.ptr(f), O.ptr(s), 1, f.size, f.size, f.size)
            stored.append(s)
    tag = f"ns{n}_{'bf16' if store else 'f32'}"
    rec = {}
    for style in styles:
        print("ns", style, n, "store", store, flush=True)
        g = Ns(0 if style == "fused" else 1, n, store, ty=64 if style == "fused" else 32)
        try:
            g.run(times=3)
            for name, s in zip("uvp", stored):
                assert np.array_equal(g.read(name), s), name
            R = [g.read(k) for k in ("Ru", "Rv", "Rc")]
            norm = g.read("norm")
        finally:
            g.close()
        rec[style] = ns_errors(n, store, stored, R, norm)
        if store == 1:
            rec[style]["quantisation_error_vs_f32_fields"] = ns_errors(n, 0, fields, R, norm)
    record_errors(tag, rec)
    print(tag, rec, flush=True)
    for style in styles:
        e = rec[style]
        assert all(e[k]["rel_l2"] < TOL_L2 for k in ("Ru", "Rv", "Rc")), rec
        assert max(e["norms_rel"]) < 1e-5, rec
