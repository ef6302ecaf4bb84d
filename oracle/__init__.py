"""ctypes bindings of the CPU oracle (oracle/oracle.h). TEST INFRASTRUCTURE ONLY: imported by tests/, by
__graft_entry__.smoke() as the checker, and by bench.py's cpu_baseline leg; never by the product."""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
_P = C.c_void_p
_S = C.c_uint64


def _load(path):
    if not os.path.exists(path):
        raise ImportError(f"{path} is not built: make -C oracle")
    return C.CDLL(path)


_lib = None
_ref = None


def lib():
    global _lib
    if _lib is None:
        L = _load(os.path.join(HERE, "liboracle.so"))
        sig = {
            "or_mix32": (C.c_uint32, [C.c_uint32]),
            "or_hash_unit": (C.c_float, [C.c_uint32, _S]),
            "or_fill_hash_f32": (None, [C.c_uint32, _P, _S, _S, _S]),
            "or_fill_hash_f32_base": (None, [C.c_uint32, _P, _S, _S, _S, _S]),
            "or_strip_partials_f32": (None, [_S, _S, _S, _P, _P, _P, _P, _P, _P, _P]),
            "or_array_seed": (C.c_uint32, [C.c_uint32, C.c_uint32]),
            "or_relative_l2_f32": (C.c_double, [_P, _P, _S]),
            "or_relative_l2_f64": (C.c_double, [_P, _P, _S]),
            "or_relative_linf_f32": (C.c_double, [_P, _P, _S]),
            "or_f32_to_bf16": (C.c_uint16, [C.c_float]),
            "or_bf16_to_f32": (C.c_float, [C.c_uint16]),
            "or_f32_to_e4m3": (C.c_uint8, [C.c_float]),
            "or_e4m3_to_f32": (C.c_float, [C.c_uint8]),
            "or_quantize_bf16": (None, [_P, _P, _S, _S, _S, _S]),
            "or_quantize_e4m3_rows": (None, [_P, _P, _P, _S, _S, _S, _S]),
            "or_gemver_f64": (None, [_S, C.c_int, _P, _S, _P, _P, _P, _P, _P, _P, _P, C.c_double, C.c_double, _P, _P]),
            "or_rank2_update_f32": (None, [_S, _S, _P, _P, _P, _P, _P, _P]),
            "or_gemvt_partial_f32": (None, [_S, _S, _P, _P, _P]),
            "or_col_reduce_f32": (None, [_S, _S, _P, _P, C.c_float, _P]),
            "or_gemv_f32": (None, [_S, _P, _P, C.c_float, _P]),
            "or_gemver_twopass_f32": (None, [_S, _S, _P, _P, _P, _P, _P, _P, _P, C.c_float, C.c_float, _P, _P, _P]),
            "or_chain_f64": (None, [_S] + [_P] * 7 + [C.c_double, C.c_double] + [_P] * 7),
            "or_bicgk_onepass_f32": (None, [_S, _S, _P, _P, _P, _P, _P]),
            "or_row_reduce_f32": (None, [_S, _S, _P, _P]),
            "or_exclusive_scan_u32": (None, [_P, _P, _S]),
            "or_exclusive_scan_u64": (None, [_P, _P, _S]),
            "or_build_column": (C.c_uint32, [C.c_int, _P, _S, _S, C.c_int, _P, _P, _P]),
            "or_ragged_sizes": (None, [C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32, _P]),
            "or_ns_fields": (None, [C.c_uint32, C.c_int, C.c_int, _P, _P, _P]),
            "or_ns_residual_f64": (None, [C.c_int, C.c_int, C.c_double, C.c_double, C.c_int, _P, _P, _P, _P, _P, _P, _P]),
            "or_ns_compare": (None, [C.c_int, C.c_int, C.c_double, C.c_double, C.c_int, _P, _P, _P, _P, _P, _P, _P, _P]),
            "or_nsc_residual_f64": (None, [C.c_int, C.c_int, C.c_int, C.c_int, _P, _P, _P]),
            "or_nsc_prim_f64": (None, [_P, _P]),
            "or_nsc_slope_f64": (C.c_double, [C.c_double, C.c_double, C.c_int]),
            "or_nsc_hllc_f64": (None, [_P, _P, C.c_int, _P]),
            "or_nsc_visc_f64": (None, [_P, _P, _P, _P, C.c_int, _P]),
            "or_nsc_cell_diag_f64": (None, [_P, _P, _P]),
            "or_nsc_init_f32": (None, [C.c_int, C.c_int, C.c_float, _P]),
            "or_threads": (C.c_int, []),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def ref():
    """The reference's own chain_data.hpp behind oracle/ref_chain.cpp, or None where it was not built."""
    global _ref
    if _ref is None:
        path = os.path.join(HERE, "_ref", "libchain_ref.so")
        if not os.path.exists(path):
            return None
        R = C.CDLL(path)
        R.ref_hash_unit.restype = C.c_float
        R.ref_hash_unit.argtypes = [C.c_uint32, _S]
        R.ref_make_vector.argtypes = [C.c_uint32, _S, _P]
        R.ref_make_matrix_A0.argtypes = [_S, _P]
        R.ref_chain.argtypes = [_S] + [_P] * 13
        R.ref_relative_l2.restype = C.c_double
        R.ref_relative_l2.argtypes = [_P, _P, _S]
        R.ref_relative_l2_f64.restype = C.c_double
        R.ref_relative_l2_f64.argtypes = [_P, _P, _S]
        _ref = R
    return _ref


def ptr(a):
    return C.c_void_p(a.ctypes.data)


# ---- convenience wrappers -----------------------------------------------------------------------------------------
SEEDS = {"A": 1, "u1": 2, "v1": 3, "u2": 4, "v2": 5, "y": 6, "z": 7}


def hash_matrix(seed, rows, cols):
    a = np.empty((rows, cols), np.float32)
    lib().or_fill_hash_f32(seed, ptr(a), rows, cols, cols)
    return a


def gemver_inputs(n, base_seed=0):
    """A and the six vectors, as the models generate them on the device."""
    out = {"A": hash_matrix(base_seed * 16 + SEEDS["A"], n, n)}
    for k in ("u1", "v1", "u2", "v2", "y", "z"):
        out[k] = hash_matrix(base_seed * 16 + SEEDS[k], 1, n)[0]
    return out


def gemver_f64(inp, alpha, beta, store=0, A=None, row_scale=None):
    n = inp["u1"].shape[0]
    A = inp["A"] if A is None else A
    x = np.empty(n)
    w = np.empty(n)
    lib().or_gemver_f64(n, store, ptr(A), n, ptr(row_scale) if row_scale is not None else None,
                        *[ptr(inp[k]) for k in ("u1", "v1", "u2", "v2", "y", "z")], alpha, beta, ptr(x), ptr(w))
    return x, w


CHAIN_OUT = ("x1", "w1", "q", "s", "x2", "w2")


def chain_scale(n):
    """alpha = beta = 1/sqrt(n) in f32, as the chain's Systems initialise them."""
    return np.float32(1.0) / np.sqrt(np.float32(n))


def chain_f64(inp, alpha, beta, want_B2=False):
    """The chain in FP64 (or_chain_f64): {x1 w1 q s x2 w2 [B2]}."""
    n = inp["u1"].shape[0]
    out = {k: np.empty(n) for k in CHAIN_OUT}
    B2 = np.empty((n, n)) if want_B2 else None
    lib().or_chain_f64(n, ptr(inp["A"]), *[ptr(inp[k]) for k in ("u1", "v1", "u2", "v2", "y", "z")], float(alpha), float(beta),
                       *[ptr(out[k]) for k in CHAIN_OUT], ptr(B2) if want_B2 else None)
    if want_B2:
        out["B2"] = B2
    return out


def chain_f32(inp, onepass, R):
    """The chain in f32 with the components' evaluation order: the unfused / parallel / fused styles (onepass False) or
    the onepass / naive / crafted styles (onepass True; crafted at R = 16). {B1 x1 w1 q s B2 x2 w2}."""
    L = lib()
    n = inp["u1"].shape[0]
    ab = float(chain_scale(n))
    vec = [ptr(inp[k]) for k in ("u1", "v1", "u2", "v2")]

    def gemver(A, y, z):
        B = np.empty((n, n), np.float32)
        L.or_rank2_update_f32(n, R, ptr(A), *vec, ptr(B))
        P = np.empty((n // R, n), np.float32)
        L.or_gemvt_partial_f32(n, R, ptr(B), ptr(y), ptr(P))
        x = np.empty(n, np.float32)
        L.or_col_reduce_f32(n, n // R, ptr(P), ptr(z), ab, ptr(x))
        w = np.empty(n, np.float32)
        L.or_gemv_f32(n, ptr(B), ptr(x), ab, ptr(w))
        return B, x, w

    B1, x1, w1 = gemver(inp["A"], inp["y"], inp["z"])
    q = np.empty(n, np.float32)
    sv = np.empty(n, np.float32)
    Ps = np.empty((n // R, n), np.float32)
    if onepass:
        Pq = np.empty((n, n // 256), np.float32)
        L.or_bicgk_onepass_f32(n, R, ptr(B1), ptr(x1), ptr(w1), ptr(Ps), ptr(Pq))
        L.or_row_reduce_f32(n, n // 256, ptr(Pq), ptr(q))
    else:
        L.or_gemv_f32(n, ptr(B1), ptr(x1), 1.0, ptr(q))
        L.or_gemvt_partial_f32(n, R, ptr(B1), ptr(w1), ptr(Ps))
    L.or_col_reduce_f32(n, n // R, ptr(Ps), None, 1.0, ptr(sv))
    B2, x2, w2 = gemver(B1, q, sv)
    return {"B1": B1, "x1": x1, "w1": w1, "q": q, "s": sv, "B2": B2, "x2": x2, "w2": w2}


def rel_l2(got, want):
    got = np.ascontiguousarray(got, np.float32)
    want = np.ascontiguousarray(want, np.float64)
    return lib().or_relative_l2_f32(ptr(got), ptr(want), got.size)


def rel_linf(got, want):
    got = np.ascontiguousarray(got, np.float32)
    want = np.ascontiguousarray(want, np.float64)
    return lib().or_relative_linf_f32(ptr(got), ptr(want), got.size)


def exclusive_scan_u32(v):
    v = np.ascontiguousarray(v, np.uint32)
    out = np.empty(v.size + 1, np.uint32)
    lib().or_exclusive_scan_u32(ptr(v), ptr(out), v.size)
    return out


def build_column(src, offsets_mode=False, k=1):
    """(extents, offs, total, flag) of shape_hook.hpp:215-348 restated."""
    src = np.ascontiguousarray(src)
    dt = {np.dtype(np.int32): 0, np.dtype(np.uint32): 1, np.dtype(np.int64): 2, np.dtype(np.uint64): 3}[src.dtype]
    count = src.shape[0]
    n = count - 1 if offsets_mode else count
    ext = np.zeros(max(n, 0), np.uint32)
    offs = np.zeros(max(n + 1, 1), np.uint32)
    total = C.c_uint32(0)
    flag = lib().or_build_column(dt, ptr(src), count, k, 1 if offsets_mode else 0, ptr(ext), ptr(offs), C.byref(total))
    return ext, offs, total.value, flag


# ---- compressible NS residual (examples/ns-residual/physics.hpp, reference.hpp) ----------------------------------------
NSC_GAMMA, NSC_MU, NSC_PR, NSC_CFL = 1.4, 0.01, 0.72, 0.5


def _d(a, n):
    a = np.ascontiguousarray(a, np.float64)
    assert a.size == n
    return a


def nsc_init(nx, ny, amp=1.0):
    U = np.empty((4, ny, nx), np.float32)
    lib().or_nsc_init_f32(nx, ny, amp, ptr(U))
    return U


def nsc_residual(U, walls=True, vanleer=False):
    """(R [4, ny, nx] f64, diag {dt, ws, sq[4]}) of the FP64 restatement."""
    U = np.ascontiguousarray(U, np.float32)
    _, ny, nx = U.shape
    R = np.empty((4, ny, nx))
    d = np.empty(6)
    lib().or_nsc_residual_f64(nx, ny, 1 if walls else 0, 1 if vanleer else 0, ptr(U), ptr(R), ptr(d))
    return R, {"dt": d[0], "ws": d[1], "sq": d[2:6].copy()}


def nsc_prim(U):
    W = np.empty(4)
    lib().or_nsc_prim_f64(ptr(_d(U, 4)), ptr(W))
    return W


def nsc_slope(a, b, vanleer=False):
    return lib().or_nsc_slope_f64(float(a), float(b), 1 if vanleer else 0)


def nsc_hllc(wl, wr, dir):
    F = np.empty(4)
    lib().or_nsc_hllc_f64(ptr(_d(wl, 4)), ptr(_d(wr, 4)), dir, ptr(F))
    return F


def nsc_visc(wl, wr, gl, gr, dir):
    F = np.empty(4)
    lib().or_nsc_visc_f64(ptr(_d(wl, 4)), ptr(_d(wr, 4)), ptr(_d(gl, 6)), ptr(_d(gr, 6)), dir, ptr(F))
    return F


def nsc_cell_diag(W):
    dt, ws = C.c_double(), C.c_double()
    lib().or_nsc_cell_diag_f64(ptr(_d(W, 4)), C.byref(dt), C.byref(ws))
    return dt.value, ws.value
