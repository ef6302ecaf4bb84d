"""The compressible NS residual on the device (workloads/nsc_*.hpp through cs_nsc_*) against the FP64 oracle, with the
reference driver's --check contract (examples/ns-residual/main.cpp:304-329): R within 1e-4 relative L2; the deterministic
diagnostics within 1e-4 relative; the atomic min dt and max wave speed EQUAL to the deterministic ones; a uniform periodic
flow gives R exactly 0. Further: the three variants give bit-identical residuals and partials (one tile arithmetic), the
error stage's values are the atomic minus the deterministic diagnostics, the atomic accumulators are reset every run, and
repeated (graph-replayed) runs give the eager results."""
import ctypes as C

import numpy as np
import pytest

import oracle as O
from cospanmdao_amd import check, workloads

pytestmark = pytest.mark.gpu

VARIANTS = {"unfused": 0, "fused": 1, "launches": 2}
TOL = 1e-4  # main.cpp:321


class Nsc:
    def __init__(self, variant, nx, ny, vanleer=False, walls=True, ty=8):
        self.h = C.c_void_p()
        check(workloads.cs_nsc_create(C.byref(self.h), VARIANTS[variant], nx, ny, 1 if vanleer else 0, 1 if walls else 0, ty))
        self.nx, self.ny = nx, ny
        self.tiles = (nx // 64) * (ny // ty)

    def init(self, amp=1.0):
        check(workloads.cs_nsc_init(self.h, amp))

    def write(self, U):
        U = np.ascontiguousarray(U, np.float32)
        check(workloads.cs_nsc_write(self.h, U.ctypes.data, U.nbytes))

    def run(self, n=1):
        for _ in range(n):
            check(workloads.cs_nsc_run(self.h))
        check(workloads.cs_nsc_wait(self.h))

    def read(self, name):
        shapes = {"U": (4, self.ny, self.nx), "R": (4, self.ny, self.nx), "Part": (6, self.tiles), "Sq": (4,), "DetSq": (4,), "Err": (6,)}
        out = np.empty(shapes.get(name, (1,)), np.float32)
        check(workloads.cs_nsc_read(self.h, name.encode(), out.ctypes.data, out.nbytes))
        return out

    def outputs(self):
        return {k: self.read(k) for k in ("U", "R", "Part", "Dt", "Ws", "Sq", "DetDt", "DetWs", "DetSq", "Err")}

    def stats(self):
        s = (C.c_uint64 * 4)()
        check(workloads.cs_nsc_stats(self.h, s))
        return list(s)

    def close(self):
        if self.h:
            check(workloads.cs_nsc_destroy(self.h))
            self.h = C.c_void_p()


def evaluate(variant, nx, ny, vanleer=False, walls=True, ty=8, amp=1.0, runs=1, U=None):
    m = Nsc(variant, nx, ny, vanleer, walls, ty)
    try:
        if U is None:
            m.init(amp)
        else:
            m.write(U)
        m.run(runs)
        return m.outputs(), m.stats()
    finally:
        m.close()


def rel(a, b):
    return abs(float(a) - float(b)) / max(abs(float(b)), 1e-300)


def check_against_oracle(got, vanleer, walls):
    R, d = O.nsc_residual(got["U"], walls=walls, vanleer=vanleer)
    l2 = O.rel_l2(got["R"].ravel(), R.ravel())
    det = max([rel(got["DetDt"][0], d["dt"]), rel(got["DetWs"][0], d["ws"])] + [rel(got["DetSq"][q], d["sq"][q]) for q in range(4)])
    print(f"R rel-L2 {l2:.2e}  det rel {det:.2e}  atomic-norm gap {max(rel(got['Sq'][q], got['DetSq'][q]) for q in range(4)):.2e}", flush=True)
    assert l2 <= TOL, l2
    assert det <= TOL, det
    assert got["Dt"][0] == got["DetDt"][0] and got["Ws"][0] == got["DetWs"][0]  # min and max are exact
    err = np.concatenate([got["Dt"] - got["DetDt"], got["Ws"] - got["DetWs"], got["Sq"] - got["DetSq"]])
    np.testing.assert_array_equal(got["Err"], err)


@pytest.mark.parametrize("variant", list(VARIANTS))
@pytest.mark.parametrize("vanleer", [False, True])
@pytest.mark.parametrize("walls", [True, False])
def test_against_the_fp64_residual(variant, vanleer, walls):
    print("nsc", variant, "vanleer", vanleer, "walls", walls, "256 x 128", flush=True)
    got, _ = evaluate(variant, 256, 128, vanleer, walls)
    check_against_oracle(got, vanleer, walls)


@pytest.mark.parametrize("ty", [8, 16])
def test_variants_are_bit_identical(ty):
    res = {}
    for v in VARIANTS:
        print("nsc", v, "ty", ty, flush=True)
        res[v], _ = evaluate(v, 192, 96 if ty == 8 else 112, ty=ty)
    for v in ("fused", "launches"):
        for k in ("R", "Part", "DetDt", "DetWs", "DetSq", "Dt", "Ws"):
            np.testing.assert_array_equal(res[v][k], res["unfused"][k], err_msg=f"{v} {k}")


@pytest.mark.parametrize("variant", list(VARIANTS))
@pytest.mark.parametrize("vanleer", [False, True])
def test_uniform_periodic_flow_gives_exactly_zero(variant, vanleer):
    got, _ = evaluate(variant, 128, 64, vanleer, walls=False, amp=0.0)
    assert np.all(got["R"] == 0.0)
    assert np.all(got["DetSq"] == 0.0) and np.all(got["Sq"] == 0.0)


@pytest.mark.parametrize("variant", list(VARIANTS))
def test_repeated_runs_reset_the_accumulators(variant):
    """Five runs (the model captures a graph and replays it): R and the diagnostics equal the first run's; the atomic
    accumulators start again from their reset values every run."""
    once, _ = evaluate(variant, 128, 64)
    many, st = evaluate(variant, 128, 64, runs=5)
    print("stats", st, flush=True)
    for k in ("R", "Part", "DetDt", "DetWs", "DetSq", "Dt", "Ws"):
        np.testing.assert_array_equal(many[k], once[k], err_msg=k)
    np.testing.assert_allclose(many["Sq"], once["DetSq"], rtol=1e-5)
    assert st[3] >= 1, st  # replayed


def test_fused_waits_publish_each_runs_partials():
    """The fused variant with group waits reads every tile's partials in later segments of the same launch, after the
    tiles published them (an L2 write-back shared by the work-groups of an XCD). Four different states in turn through
    one model (eager, captured, then replayed), 2048 tiles over every XCD: the partials and diagnostics equal the
    launch-boundary variant's on the same state bit for bit, so none was read stale from the run before."""
    rng = np.random.default_rng(5)
    nx, ny = 1024, 1024
    base = O.nsc_init(nx, ny)
    fused, ref = Nsc("fused", nx, ny), Nsc("launches", nx, ny)
    try:
        for it in range(4):
            U = (base * (1 + 0.05 * (it + 1) * rng.standard_normal(base.shape))).astype(np.float32)
            for m in (fused, ref):
                m.write(U)
                m.run(2 if it == 3 else 1)
            print("state", it, "stats", fused.stats(), flush=True)
            got, want = fused.outputs(), ref.outputs()
            for k in ("R", "Part", "DetDt", "DetWs", "DetSq", "Dt", "Ws"):  # the atomic sums' order is free: Sq, Err
                np.testing.assert_array_equal(got[k], want[k], err_msg=f"state {it} {k}")
    finally:
        fused.close()
        ref.close()


def test_host_state_with_random_perturbation():
    rng = np.random.default_rng(11)
    U = O.nsc_init(256, 64)
    U = (U * (1 + 0.02 * rng.standard_normal(U.shape))).astype(np.float32)
    got, _ = evaluate("launches", 256, 64, vanleer=True, walls=True, U=U)
    np.testing.assert_array_equal(got["U"], U)
    check_against_oracle(got, True, True)


@pytest.mark.parametrize("ty", [8, 16])
def test_1024_square(ty):
    got, _ = evaluate("fused", 1024, 1024, ty=ty)
    check_against_oracle(got, False, True)


def test_bad_problems_are_errors():
    h = C.c_void_p()
    assert workloads.cs_nsc_create(C.byref(h), 0, 100, 64, 0, 1, 8) == 1  # nx not a multiple of 64
    assert workloads.cs_nsc_create(C.byref(h), 0, 128, 60, 0, 1, 8) == 1  # ny not a multiple of the tile rows
    assert workloads.cs_nsc_create(C.byref(h), 0, 128, 64, 0, 1, 12) == 1  # tile rows 8 or 16
    assert workloads.cs_nsc_create(C.byref(h), 3, 128, 64, 0, 1, 8) == 1  # no such variant
    assert workloads.cs_nsc_create(C.byref(h), 0, 128, 64, 2, 1, 8) == 1  # no such limiter
