"""The batched ragged GEMVER on the device: the offsets scan bit for bit, every instance's x and w against the oracle, the
phased form (one launch per phase) and the persistent form (one work-queue launch) under graph replay, the autotuner."""
import ctypes as C
import os

import numpy as np
import pytest

import oracle as O
from cospanmdao_amd import CsError, check, workloads

pytestmark = pytest.mark.gpu

# The persistent form: one work-queue launch of G work-groups whose units come from the unit map; a work-group publishes
# the units it ran once per phase and waits once per phase (its protocol is simulated on the CPU by
# tests/host/batch_host.cpp over adversarial interleavings and residency caps).
PERSISTENT = pytest.mark.gpu  # no extra condition: the persistent tests run
# whether the library lets cs_batch_set_mode select the persistent form (batch_lib.hip kPersistentFormEnabled)
PERSISTENT_ENABLED = True
PHASED, PERSISTENT_MODE = 0, 1

ALPHA, BETA = 1.5, 1.2
TOL_L2 = 1e-5


class Batch:
    def __init__(self, count, lo, hi, seed=0, mode=PHASED):
        self.h = C.c_void_p()
        check(workloads.cs_batch_create(C.byref(self.h), count, lo, hi, ALPHA, BETA, seed))
        check(workloads.cs_batch_set_mode(self.h, mode))
        self.count, self.seed = count, seed
        self.sizes = np.empty(count, np.uint32)
        check(workloads.cs_batch_sizes(self.h, self.sizes.ctypes.data_as(C.POINTER(C.c_uint32)), count))

    def params(self, R, RT, G):
        check(workloads.cs_batch_set_params(self.h, R, RT, G))

    def run(self):
        check(workloads.cs_batch_run(self.h))
        check(workloads.cs_batch_wait(self.h))

    def read(self, name, dtype, count):
        out = np.empty(count, dtype)
        check(workloads.cs_batch_read(self.h, name.encode(), out.ctypes.data, out.nbytes))
        return out

    def stats(self):
        s = (C.c_uint64 * 4)()
        check(workloads.cs_batch_stats(self.h, s))
        return list(s)

    def close(self):
        if self.h:
            check(workloads.cs_batch_destroy(self.h))
            self.h = C.c_void_p()


def counts_of(sizes, R, RT):
    K = sizes.size
    n = sizes.astype(np.uint64)
    ld = (n + 31) // 32 * 32  # rows padded to 128-byte lines
    strips = (n + R - 1) // R
    tiles = (n + 4095) // 4096  # pass-1 tiles are R rows x 4096 columns
    blocks = (n + 1023) // 1024  # column-reduction blocks
    units = [strips * tiles, blocks, (n + RT - 1) // RT]
    # the work order: instances by size, largest first, ties in index order (a stable sort of -n)
    order = np.argsort(-sizes.astype(np.int64), kind="stable")
    c = np.concatenate([n, n * ld // 4, strips * n] + units + [u[order] for u in units])
    assert c.size == 9 * K
    return c.astype(np.uint32)


def instance_inputs(sizes, k, seed):
    """Instance k's A and vectors as the model generates them: one flat hash stream per array."""
    L = O.lib()
    n = int(sizes[k])
    a_base = int(np.sum(sizes[:k].astype(np.uint64) ** 2))
    v_base = int(np.sum(sizes[:k].astype(np.uint64)))
    A = np.empty((n, n), np.float32)
    L.or_fill_hash_f32_base(seed * 16 + 1, O.ptr(A), n, n, n, a_base)
    inp = {"A": A}
    for i, name in enumerate(("u1", "v1", "u2", "v2", "y", "z")):
        v = np.empty(n, np.float32)
        L.or_fill_hash_f32_base(seed * 16 + 2 + i, O.ptr(v), 1, n, n, v_base)
        inp[name] = v
    return inp


def check_instances(b, which, R):
    v_off = np.concatenate([[0], np.cumsum(b.sizes.astype(np.uint64))]).astype(np.int64)
    total = int(v_off[-1])
    x = b.read("x", np.float32, total)
    w = b.read("w", np.float32, total)
    worst = 0.0
    for k in which:
        n = int(b.sizes[k])
        inp = instance_inputs(b.sizes, k, b.seed)
        gx, gw = x[v_off[k]:v_off[k] + n], w[v_off[k]:v_off[k] + n]
        xr, wr = O.gemver_f64(inp, ALPHA, BETA)
        ex, ew = O.rel_l2(gx, xr), O.rel_l2(gw, wr)
        worst = max(worst, ex, ew)
        assert ex < TOL_L2 and ew < TOL_L2, (k, n, ex, ew)
        # x keeps the components' order: bit for bit against the strip partials + column reduction of the oracle
        strips = (n + R - 1) // R
        P = np.empty((strips, n), np.float32)
        O.lib().or_strip_partials_f32(n, n, R, O.ptr(inp["A"]), *[O.ptr(inp[q]) for q in ("u1", "v1", "u2", "v2", "y")], O.ptr(P))
        xf = np.empty(n, np.float32)
        O.lib().or_col_reduce_f32(n, strips, O.ptr(P), O.ptr(inp["z"]), BETA, O.ptr(xf))
        assert np.array_equal(gx, xf), (k, n)
    return worst


MODES = [pytest.param(PHASED, id="phased"), pytest.param(PERSISTENT_MODE, id="persistent", marks=PERSISTENT)]


@pytest.mark.parametrize("mode", MODES)
@pytest.mark.parametrize("R,RT,G", [(32, 4, 3), (32, 8, 64), (64, 8, 17), (128, 4, 512), (64, 16, 5), (256, 32, 9)])
def test_small_batch(R, RT, G, mode):
    print("batch small", R, RT, G, "mode", mode, flush=True)
    b = Batch(24, 16, 700, seed=1, mode=mode)
    try:
        sizes = np.empty(24, np.uint32)
        O.lib().or_ragged_sizes(1, 24, 16, 700, O.ptr(sizes))
        assert np.array_equal(b.sizes, sizes)
        b.params(R, RT, G)
        b.run()
        counts = b.read("counts", np.uint32, 9 * 24)
        assert np.array_equal(counts, counts_of(b.sizes, R, RT))
        offs = b.read("offs", np.uint32, 9 * 24 + 1)
        assert np.array_equal(offs, O.exclusive_scan_u32(counts))  # the device scan, bit for bit
        res = b.read("result", np.uint32, 2)
        assert res[0] == offs[-1] and res[1] == 0
        check_instances(b, range(24), R)
    finally:
        b.close()


@pytest.mark.parametrize("mode", MODES)
def test_wide_instances(mode):
    """Instances wider than one pass-1 tile (4096 columns): two and three column blocks, partial last blocks."""
    print("batch wide, mode", mode, flush=True)
    b = Batch(5, 3000, 9000, seed=5, mode=mode)
    try:
        assert int(b.sizes.max()) > 4096, b.sizes
        b.params(64, 4, 128)
        b.run()
        counts = b.read("counts", np.uint32, 9 * 5)
        assert np.array_equal(counts, counts_of(b.sizes, 64, 4))
        worst = check_instances(b, range(5), 64)
        print(f"sizes {b.sizes.tolist()}, worst rel-L2 {worst:.2e}", flush=True)
    finally:
        b.close()


@pytest.mark.parametrize("mode", MODES)
@pytest.mark.parametrize("R,RT,G", [(32, 4, 7), (64, 32, 48)])
def test_tiny_and_odd_instances(R, RT, G, mode):
    """48 instances of 1 to 40 rows (seed 7: sixteen of them 1, 2 or 3 rows): rows shorter than one 4-column chunk, the
    masked last chunk of pass 1, pass 2's scalar column tail, strips and pass-2 groups with fewer rows than they hold."""
    print("batch tiny", R, RT, G, "mode", mode, flush=True)
    b = Batch(48, 1, 40, seed=7, mode=mode)
    try:
        assert int(b.sizes.min()) == 1 and int((b.sizes < 4).sum()) == 16, b.sizes
        b.params(R, RT, G)
        b.run()
        b.run()
        counts = b.read("counts", np.uint32, 9 * 48)
        assert np.array_equal(counts, counts_of(b.sizes, R, RT))
        check_instances(b, range(48), R)
    finally:
        b.close()


@pytest.mark.parametrize("mode", MODES)
def test_replays_and_parameter_changes(mode):
    print("batch replays, mode", mode, flush=True)
    b = Batch(40, 64, 1200, seed=2, mode=mode)
    try:
        b.params(64, 8, 96)
        for _ in range(4):
            b.run()
        st = b.stats()
        assert st[2] >= 1 and st[3] >= 2, st  # captured once, then replayed
        check_instances(b, [0, 7, 39], 64)
        b.params(32, 4, 40)  # new parameters: new counts, new grid; eager, then captured again
        for _ in range(3):
            b.run()
        check_instances(b, [3, 20, 39], 32)
    finally:
        b.close()


@pytest.mark.parametrize("mode", MODES)
def test_full_batch_512(mode):
    """512 instances, N log-uniform in [64, 4096], seed 0: the measured configuration (about 4 GB of A)."""
    print("batch full 512, mode", mode, flush=True)
    b = Batch(512, 64, 4096, seed=0, mode=mode)
    try:
        b.params(64, 8, 1024)
        b.run()
        b.run()
        b.run()  # a replay
        counts = b.read("counts", np.uint32, 9 * 512)
        offs = b.read("offs", np.uint32, 9 * 512 + 1)
        assert np.array_equal(offs, O.exclusive_scan_u32(counts))
        assert np.array_equal(offs[:513].astype(np.int64), np.concatenate([[0], np.cumsum(b.sizes.astype(np.int64))]))
        order = np.argsort(b.sizes)
        which = sorted(set(order[:3].tolist() + order[-3:].tolist() + list(range(0, 512, 64))))
        worst = check_instances(b, which, 64)
        print(f"worst rel-L2 over {len(which)} instances: {worst:.2e}", flush=True)
    finally:
        b.close()


@pytest.mark.parametrize("mode", MODES)
def test_autotune_picks_a_candidate_and_caches(tmp_path, mode):
    print("batch autotune, mode", mode, flush=True)
    b = Batch(32, 64, 900, seed=3, mode=mode)
    try:
        best = (C.c_uint32 * 3)()
        path = str(tmp_path / "autotune.txt")
        check(workloads.cs_batch_autotune(b.h, 2, path.encode(), best))
        assert best[0] in (64, 128, 256) and best[1] in (8, 16, 32) and best[2] > 0
        assert os.path.exists(path)
        again = (C.c_uint32 * 3)()
        check(workloads.cs_batch_autotune(b.h, 2, path.encode(), again))  # from the cache
        assert list(again) == list(best)
        b.run()
        check_instances(b, [0, 31], int(best[0]))
    finally:
        b.close()


def test_bad_parameters_are_errors():
    b = Batch(4, 64, 128)
    try:
        assert workloads.cs_batch_set_params(b.h, 16, 8, 64) == 1
        assert workloads.cs_batch_set_params(b.h, 64, 5, 64) == 1
        assert workloads.cs_batch_set_params(b.h, 64, 8, 0) == 1
        assert workloads.cs_batch_set_mode(b.h, 2) == 1
        assert workloads.cs_batch_set_params(b.h, 64, 12, 64) == 1
        if PERSISTENT_ENABLED:
            check(workloads.cs_batch_set_mode(b.h, PERSISTENT_MODE))
            check(workloads.cs_batch_set_params(b.h, 64, 16, 64))  # wave-level pass 2 in both forms
            assert workloads.cs_batch_set_params(b.h, 64, 12, 64) == 1
        else:
            with pytest.raises(CsError) as e:
                check(workloads.cs_batch_set_mode(b.h, PERSISTENT_MODE))
            assert e.value.status == 1 and "switched off" in e.value.message
    finally:
        b.close()
