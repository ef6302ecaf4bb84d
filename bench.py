"""Benchmark of the fused-run path on one MI355X: prints ONE JSON line.

Headline workload (the configuration the metric is quoted on): GEMVER N = 16384, f32, A uniform[-1, 1) (1 GiB),
alpha = 1.5, beta = 1.2, seed 0, run as the fused two-pass model (Â never stored; A read twice). A step is one forward of
the model (x = beta Â^T y + z, then w = alpha Â x) with every input resident in HBM; after the second step every step
replays one captured graph. `value` = algorithmic HBM bytes per step (2 N^2 * 4 + 13 N * 4 = 2 148 335 616 B) / measured
time per step, in GB/s.

Beside it (`extra`): the same GEMVER as the reference's four unfused components (B materialised); the NS residual at
4096^2 fused (one fused group) and unfused; the batched ragged GEMVER (512 instances, N log-uniform in [64, 4096]: instances/s,
p50/p99 per replay, per-byte rate against the fused N=16384 rate); GEMVER N = 32768 with A stored f32, bf16 and e4m3 +
row scale (speedups) and the NS residual at 8192^2 with bf16 fields. Errors against the FP64 oracle are measured by the
GPU tests (tests/test_parity_full_gpu.py writes them; `extra.errors_vs_fp64` carries the committed record of that run,
profiles/parity_errors_full_size.md), not here. The headline is also run cache-cold (`extra.cache_cold`: two and four
separate copies of A in turn). The roofline object is for the dominant kernel of the headline workload; the CPU baseline
follows the reference's protocol (median of five processes of 20 forwards) on the host cores (oracle/cpu_bench.c). More
than one GPU is not part of this build: --gpus N > 1 exits with a message.
"""
import argparse
import ctypes as C
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "fused GEMVER/NS-residual HBM GB/s (% of 8 TB/s) vs plain; batched instances/s"
PEAK_GBS = 8000.0
UNFUSED, FUSED, NAIVE, TWOPASS, TWOPASS_1 = range(5)
REPLAY_GRAPH, REPLAY_DIRECT = 0, 1  # CS_REPLAY_GRAPH, CS_REPLAY_DIRECT


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def gemver_bytes_fused(n, elem=4):
    return 2 * n * n * elem + 13 * n * 4


def gemver_bytes_plain(n, R):
    return 4 * n * n * 4 + 2 * (n // R) * n * 4 + 13 * n * 4


class Gemver:
    def __init__(self, style, n, R, store=0, rt2=8, seed=0):
        from cospanmdao_amd import check, workloads

        self.w, self.check = workloads, check
        self.h = C.c_void_p()
        check(workloads.cs_gemver_create(C.byref(self.h), style, store, n, R, 1.5, 1.2, seed))
        if style >= TWOPASS:
            check(workloads.cs_gemver_set_tile(self.h, rt2))

    def run(self):
        self.check(self.w.cs_gemver_run(self.h))

    def wait(self):
        self.check(self.w.cs_gemver_wait(self.h))

    def launches(self, iters):
        ms = (C.c_float * 16)()
        cnt = C.c_int()
        self.check(self.w.cs_gemver_time_launches(self.h, iters, ms, 16, C.byref(cnt)))
        return [ms[k] for k in range(cnt.value)]

    def stage_ms(self, leaf, iters):
        ms = C.c_float()
        self.check(self.w.cs_gemver_time_stage(self.h, leaf, iters, C.byref(ms)))
        return ms.value

    def set_replay(self, mode):
        self.check(self.w.cs_gemver_set_replay(self.h, mode))

    def close(self):
        self.check(self.w.cs_gemver_destroy(self.h))


def profiled_traffic(n, R, rt2):
    """HBM bytes per launch of the two passes from the committed rocprofv3 counter summary of this exact configuration
    (FETCH_SIZE and WRITE_SIZE, each scaled by what the counter reads for a 1 GiB copy), or None."""
    path = os.path.join(ROOT, "profiles", "gemver_twopass_n16384_hbm_counters.md")
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith("traffic_bytes "):
            kv = dict(t.split("=") for t in line.split()[1:])
            if kv.get("config") == f"n{n}_R{R}_RT{rt2}":
                return {"pass1": int(kv["pass1"]), "pass2": int(kv["pass2"])}
    return None


class Ns:
    def __init__(self, style, n, store=0, ty=32):
        from cospanmdao_amd import check, workloads

        self.w, self.check = workloads, check
        self.h = C.c_void_p()
        check(workloads.cs_ns_create(C.byref(self.h), style, store, n, n, 1.0 / n, 1e-3, 0, ty))

    def run(self):
        self.check(self.w.cs_ns_run(self.h))

    def wait(self):
        self.check(self.w.cs_ns_wait(self.h))

    def launches(self, iters):
        ms = (C.c_float * 16)()
        cnt = C.c_int()
        self.check(self.w.cs_ns_time_launches(self.h, iters, ms, 16, C.byref(cnt)))
        return [ms[k] for k in range(cnt.value)]

    def set_replay(self, mode):
        self.check(self.w.cs_ns_set_replay(self.h, mode))

    def close(self):
        self.check(self.w.cs_ns_destroy(self.h))


def timed_steps(model, steps, warmup):
    """Wall-clock ms per step: `warmup` untimed steps, then `steps` steps bracketed by device synchronisation."""
    from cospanmdao_amd import check, hip

    for _ in range(warmup):
        model.run()
    model.wait()
    check(hip.cs_hip_device_synchronize())
    t0 = time.perf_counter()
    for _ in range(steps):
        model.run()
    model.wait()
    check(hip.cs_hip_device_synchronize())
    return (time.perf_counter() - t0) * 1e3 / steps


def gbs(nbytes, ms):
    return nbytes / (ms * 1e-3) / 1e9


def bench_ns(extra, steps):
    """The NS residual at 4096^2 and 8192^2 (f32 fields) and 8192^2 (bf16 fields). The fused form's tile rows are picked by
    timing 32 and 64; its steps are timed with graph replay and with direct replay (the headline of each line)."""
    for n, store, tag in ((4096, 0, "ns4096_f32"), (8192, 0, "ns8192_f32"), (8192, 1, "ns8192_bf16")):
        cells = n * n
        fbytes = cells * (3 * (2 if store else 4) + 3 * 4)
        tried = {}
        for ty in (32, 64):
            m = Ns(0, n, store, ty)
            tried[ty] = timed_steps(m, 20, 3)
            m.close()
        ty = min(tried, key=tried.get)
        rec = {"tile_rows": ty, "tile_rows_tried_ms": {str(k): round(v, 4) for k, v in tried.items()}}
        m = Ns(0, n, store, ty)
        gms = timed_steps(m, steps, 3)
        m.set_replay(REPLAY_DIRECT)
        ms = timed_steps(m, steps, 3)
        ln = m.launches(3)
        m.close()
        rec["fused"] = {"ms": round(ms, 4), "GBps": round(gbs(fbytes, ms), 1), "pct_of_peak": round(100 * gbs(fbytes, ms) / PEAK_GBS, 2),
                        "launch_ms": [round(t, 4) for t in ln], "replay": "direct"}
        rec["fused_graph_replay"] = {"ms": round(gms, 4), "GBps": round(gbs(fbytes, gms), 1)}
        if store == 0:
            m = Ns(1, n, store)
            ums = timed_steps(m, steps, 3)
            ln = m.launches(3)
            m.close()
            rec["unfused"] = {"ms": round(ums, 4), "GBps": round(gbs(fbytes, ums), 1), "pct_of_peak": round(100 * gbs(fbytes, ums) / PEAK_GBS, 2),
                              "launch_ms": [round(t, 4) for t in ln]}
            rec["fused_speedup_vs_unfused"] = round(ums / ms, 3)
        rec["algorithmic_bytes"] = fbytes
        extra[tag] = rec
        log(f"{tag}: {rec}")
    extra["ns8192_bf16"]["speedup_vs_f32"] = round(extra["ns8192_f32"]["fused"]["ms"] / extra["ns8192_bf16"]["fused"]["ms"], 3)


# The reference's published n = 16384 chain on an RTX 4090 (examples/gemver-chain/README.md, results tables): ms per chain.
CHAIN_REFERENCE_4090_MS = {"unfused": 11.41, "parallel": 11.43, "fused": 9.150, "onepass": 8.655, "naive": 8.531, "crafted": 8.359}
CHAIN_STYLES = ("unfused", "parallel", "fused", "onepass", "naive", "crafted")


def bench_chain(extra, iters, n=16384, R=64):
    """The GEMVER -> BiCGK -> GEMVER chain in the reference's six styles: ms per chain over `iters` back-to-back chains
    (graph replays), the reference's nominal GB/s (8 n^2 x 4 B per chain) and the launches of one eager chain."""
    from cospanmdao_amd import check, workloads as W

    rec = {"n": n, "R": R, "nominal_bytes": 8 * n * n * 4,
           "reference_rtx4090_ms": CHAIN_REFERENCE_4090_MS, "styles": {}}
    for k, style in enumerate(CHAIN_STYLES):
        h = C.c_void_p()
        check(W.cs_chain_create(C.byref(h), k, n, R, 0))
        for _ in range(3):
            check(W.cs_chain_run(h))
        check(W.cs_chain_wait(h))
        t = C.c_float()
        check(W.cs_chain_time(h, iters, C.byref(t)))
        lm = (C.c_float * 16)()
        cnt = C.c_int()
        check(W.cs_chain_time_launches(h, 2, lm, 16, C.byref(cnt)))
        check(W.cs_chain_destroy(h))
        ms = t.value
        rec["styles"][style] = {"ms": round(ms, 4), "GBps_nominal": round(gbs(8 * n * n * 4, ms), 1),
                                "launch_ms": [round(lm[i], 4) for i in range(min(cnt.value, 16))],
                                "speedup_vs_reference_rtx4090": round(CHAIN_REFERENCE_4090_MS[style] / ms, 2)}
    base = rec["styles"]["unfused"]["ms"]
    for st in rec["styles"].values():
        st["speedup_vs_unfused"] = round(base / st["ms"], 3)
    extra["gemver_chain"] = rec
    log(f"gemver chain n={n}: " + ", ".join(f"{k} {v['ms']:.3f} ms ({v['GBps_nominal']:.0f} GB/s nominal)" for k, v in rec["styles"].items()))


# examples/ns-residual/README.md:87-103 (RTX 4090, walls, minmod): ms per evaluation at n = nx = ny
NSC_REFERENCE_4090_MS = {1024: {"unfused": 3.260, "fused": 1.532, "handwritten": 0.268},
                         2048: {"unfused": 3.263, "fused": 1.842, "handwritten": 0.945},
                         4096: {"unfused": 9.685, "fused": 5.810, "handwritten": 3.550}}
NSC_VARIANTS = ("unfused", "fused", "launches")


def bench_nsc(extra, iters, sizes=(1024, 2048, 4096)):
    """The compressible NS residual (walls, minmod, as the reference's table): ms per evaluation of each variant and tile
    height at n = 1024, 2048, 4096, the reference's nominal GB/s (32 B per cell) and the per-launch times of one run."""
    from cospanmdao_amd import check, workloads as W

    rec = {"config": "walls, minmod, nx = ny = n", "nominal_bytes_per_cell": 32, "reference_rtx4090_ms": NSC_REFERENCE_4090_MS,
           "runs": {}}
    for n in sizes:
        res = {}
        for v, variant in enumerate(NSC_VARIANTS):
            for ty in (8, 16):
                h = C.c_void_p()
                check(W.cs_nsc_create(C.byref(h), v, n, n, 0, 1, ty))
                check(W.cs_nsc_init(h, 1.0))
                for _ in range(3):
                    check(W.cs_nsc_run(h))
                check(W.cs_nsc_wait(h))
                t = C.c_float()
                check(W.cs_nsc_time(h, iters, C.byref(t)))
                lm = (C.c_float * 8)()
                cnt = C.c_int()
                check(W.cs_nsc_time_launches(h, 2, lm, 8, C.byref(cnt)))
                check(W.cs_nsc_destroy(h))
                ms = t.value
                res[f"{variant}-ty{ty}"] = {"ms": round(ms, 4), "GBps_nominal": round(gbs(32 * n * n, ms), 1),
                                            "launch_ms": [round(lm[i], 4) for i in range(min(cnt.value, 8))]}
        best = min(res, key=lambda k: res[k]["ms"])
        ref = NSC_REFERENCE_4090_MS.get(n, {})
        rec["runs"][n] = {"variants": res, "best": best,
                          "speedup_vs_reference_rtx4090": {k: round(ref[k] / res[best]["ms"], 2) for k in ref}}
        log(f"nsc n={n}: " + ", ".join(f"{k} {v['ms']:.3f} ms" for k, v in res.items()) + f"; best {best}")
    extra["ns_compressible"] = rec


def bench_precision(extra, steps):
    """GEMVER N = 32768 with A stored f32, bf16 and e4m3 + row scale, each at its fastest strip height R (128, 256, 512:
    pass 1 runs best with about two work-groups per CU, and R sets their number) and pass-2 tile height (4, 8, 16 or 32
    rows); every combination is recorded, and the speedups compare the storage types at their own best."""
    n = 32768
    rec = {}
    for store, name, elem in ((0, "f32", 4), (1, "bf16", 2), (2, "e4m3", 1)):
        nbytes = gemver_bytes_fused(n, elem) + (n * 4 if store == 2 else 0)
        tried = {}
        for R in (128, 256, 512):
            for rt2 in (4, 8, 16, 32):
                g = Gemver(TWOPASS, n, R, store=store, rt2=rt2)
                g.set_replay(REPLAY_DIRECT)
                ms = timed_steps(g, steps, 3)
                tried[f"R{R}_rt{rt2}"] = {"ms": round(ms, 4), "GBps": round(gbs(nbytes, ms), 1), "R": R, "pass2_rows": rt2,
                                          "launch_ms": [round(t, 4) for t in g.launches(3)]}
                g.close()
        best = min(tried, key=lambda k: tried[k]["ms"])
        rec[name] = dict(tried[best], algorithmic_bytes=nbytes, tried=tried)
    rec["bf16_speedup_vs_f32"] = round(rec["f32"]["ms"] / rec["bf16"]["ms"], 3)
    rec["e4m3_speedup_vs_f32"] = round(rec["f32"]["ms"] / rec["e4m3"]["ms"], 3)
    extra["gemver32768_storage"] = rec
    log(f"gemver N=32768 storage: {rec}")


def bench_cache_cold(extra, steps, n=16384, R=128, rt2=16):
    """The headline workload with A not left in any cache by the step before: two and four models (separate 1 GiB copies
    of A, different seeds) run in turn, so every step reads an A that the last step did not touch (the last-level cache
    holds 256 MB). Per-step rate of the rotation, and the dominant pass's per-launch time inside the rotation (device
    events between the launches of each run)."""
    rec = {}
    for copies in (2, 4):
        gs = [Gemver(TWOPASS, n, R, rt2=rt2, seed=s) for s in range(copies)]
        for g in gs:
            g.set_replay(REPLAY_DIRECT)

        class Rotation:
            def run(self):
                for g in gs:
                    g.run()

            def wait(self):
                for g in gs:
                    g.wait()

        ms = timed_steps(Rotation(), max(50, steps // copies), 3) / copies
        p2 = []
        for _ in range(3):
            for g in gs:
                p2.append(g.launches(1)[-1])
        for g in gs:
            g.close()
        rec[f"{copies}_copies"] = {"ms_per_step": round(ms, 4), "GBps": round(gbs(gemver_bytes_fused(n), ms), 1),
                                   "pct_of_peak": round(100 * gbs(gemver_bytes_fused(n), ms) / PEAK_GBS, 2),
                                   "pass2_launch_ms": round(sum(p2) / len(p2), 4)}
    extra["cache_cold"] = rec
    log(f"cache cold: {rec}")
    return rec


def batch_form(mode, reps):
    """One form of the batched ragged GEMVER (512 instances, N log-uniform in [64, 4096], seed 0): autotuned, then `reps`
    graph replays timed one by one with device events."""
    from cospanmdao_amd import check, workloads as W
    import numpy as np

    h = C.c_void_p()
    check(W.cs_batch_create(C.byref(h), 512, 64, 4096, 1.5, 1.2, 0))
    check(W.cs_batch_set_mode(h, mode))
    sizes = np.empty(512, np.uint32)
    check(W.cs_batch_sizes(h, sizes.ctypes.data_as(C.POINTER(C.c_uint32)), 512))
    best = (C.c_uint32 * 3)()
    cache = os.path.join(os.environ.get("XDG_CACHE_HOME", os.path.join(os.path.expanduser("~"), ".cache")), "cospanmdao_amd")
    os.makedirs(cache, exist_ok=True)
    t0 = time.perf_counter()
    check(W.cs_batch_autotune(h, 3, os.path.join(cache, "autotune.txt").encode(), best))
    tune_s = time.perf_counter() - t0
    for _ in range(3):
        check(W.cs_batch_run(h))
    check(W.cs_batch_wait(h))
    each = (C.c_float * reps)()
    check(W.cs_batch_time_each(h, reps, each))
    ms = np.array(list(each), np.float64)
    st = (C.c_uint64 * 4)()
    check(W.cs_batch_stats(h, st))
    phase_ms = []
    # unit map, then pass 1, column reductions, pass 2 (phased) or the one work-queue launch (persistent), back to back
    for leaf in ((2, 3, 4, 5) if mode == 0 else (2, 3)):
        t = C.c_float()
        check(W.cs_batch_time_stage(h, leaf, 10, C.byref(t)))
        phase_ms.append(round(t.value, 4))
    check(W.cs_batch_destroy(h))
    n = sizes.astype(np.float64)
    nbytes = float(np.sum(2 * n * n * 4 + 13 * n * 4))
    mean = float(ms.mean())
    rec = {"instances": 512, "instances_per_s": round(512 / (mean * 1e-3), 1), "ms_mean": round(mean, 4),
           "ms_p50": round(float(np.percentile(ms, 50)), 4), "ms_p99": round(float(np.percentile(ms, 99)), 4),
           "algorithmic_bytes": int(nbytes), "GBps": round(gbs(nbytes, mean), 1), "params_R_RT_G": list(best),
           "autotune_s": round(tune_s, 2), "replays": reps, "executor_stats": list(st)}
    if phase_ms:
        rec["phase_ms_back_to_back"] = phase_ms
    return rec


def bench_batch(extra, fused_gbps, reps):
    """The batched ragged GEMVER in both forms: the persistent one (plan, device scan, unit map, then ONE work-queue launch
    that runs every instance's pass 1, column reductions and pass 2) and the phased one (plan, scan, unit map, then one
    launch per phase); each replayed as one captured graph. The line reports the faster; both are kept."""
    forms = {"persistent": batch_form(1, reps), "phased": batch_form(0, reps)}
    forms["persistent"]["form"] = ("persistent: plan, device scan, unit map, one work-queue launch for every unit of the three "
                                   "phases (four launches, one replayed graph)")
    forms["phased"]["form"] = "phased: plan, device scan, unit map, pass 1, column reductions, pass 2 (six launches, one replayed graph)"
    for f in forms.values():
        f["per_byte_rate_vs_fused_16384"] = round(f["GBps"] / fused_gbps, 3) if fused_gbps else None
    best = min(forms, key=lambda k: forms[k]["ms_mean"])
    rec = dict(forms[best])
    rec["forms"] = forms
    extra["batched_ragged"] = rec
    log(f"batched: best {best}; " + "; ".join(f"{k} {v['ms_mean']:.3f} ms (p99 {v['ms_p99']:.3f})" for k, v in forms.items()))


def cpu_baseline(n, R):
    """The CPU baseline with the reference's timing protocol (examples/gemver-chain/main.cpp:154-169,
    examples/run_benchmarks.nu:9-22): every figure is the median of five processes, each timing 20 forwards after 3 warm-ups
    (one discarded process first). A forward is the oracle's f32 restatement on every OpenMP thread of the host
    (oracle/cpu_bench.c): the fused two-pass form at the headline's N (the same workload as `value`) and, as the stand-in of
    the reference's CPU path, its four components at N = 1024 (the reference's CPU-runnable configuration)."""
    import statistics
    import subprocess

    exe = os.path.join(ROOT, "oracle", "cpu_bench")
    if not os.path.exists(exe):
        raise RuntimeError("oracle/cpu_bench is not built (make -C oracle)")

    def protocol(mode, nn, rr):
        runs = []
        for k in range(6):
            r = subprocess.run([exe, mode, str(nn), str(rr)], capture_output=True, text=True, check=True, timeout=300)
            if k:  # the first process is discarded
                runs.append(json.loads(r.stdout.strip().splitlines()[-1]))
        ms = [x["ms"] for x in runs]
        med = statistics.median(ms)
        return {"ms_median": round(med, 4), "ms_min": round(min(ms), 4), "ms_max": round(max(ms), 4), "threads": runs[0]["threads"],
                "GBps": round(gemver_bytes_fused(nn) / (med * 1e-3) / 1e9, 2)}

    two = protocol("twopass", n, R)
    comp = protocol("components", 1024, 64)
    return {
        "value": two["GBps"],
        "unit": "GB/s",
        "cores": int(two["threads"]),
        "kind": "port",
        "sample": f"f32 two-pass GEMVER N={n} R={R} (oracle/cpu_bench.c: or_gemver_twopass_f32, OpenMP): median of 5 processes x "
        f"20 forwards after 3 warm-ups, {two['ms_median']:.2f} ms per forward (min {two['ms_min']:.2f}, max {two['ms_max']:.2f}); "
        "GB/s counts the fused algorithmic bytes 2 N^2 x 4 + 13 N x 4",
        "reference_path_n1024": dict(comp, what="the reference's four GEMVER components restated in f32 (Rank2Update, GemvTPartial, "
                                     "ColReduce, Gemv) at N=1024 R=64, the same protocol; GB/s counts the fused algorithmic bytes"),
    }


def measured_errors():
    """Per-output errors against the FP64 oracle at the full measured sizes, as measured by tests/test_parity_full_gpu.py on
    the GPU and committed in profiles/parity_errors_full_size.md (this program does not run the oracle itself)."""
    path = os.path.join(ROOT, "profiles", "parity_errors_full_size.md")
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith("errors_json "):
            return {"source": "profiles/parity_errors_full_size.md (tests/test_parity_full_gpu.py)", "cases": json.loads(line[len("errors_json "):])}
    return None


def copy_calibration(gib=1):
    """A plain device-to-device copy of `gib` GiB, three times (torch's copy kernel): the reference point that scales
    rocprofv3's FETCH_SIZE and WRITE_SIZE counters to bytes, since each copy reads and writes exactly gib GiB."""
    import torch

    x = torch.empty(gib * 2**28, dtype=torch.float32, device="cuda:0")
    y = torch.empty_like(x)
    for _ in range(3):
        y.copy_(x)
    torch.cuda.synchronize()
    print(json.dumps({"copy_calibration_bytes": x.numel() * 4, "copies": 3}), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--n", type=int, default=16384)
    ap.add_argument("--R", type=int, default=128)
    ap.add_argument("--rt2", type=int, default=16, help="rows per pass-2 work-group (4, 8, 16, 32)")
    ap.add_argument("--style", default="twopass", choices=["twopass", "twopass_1"])
    ap.add_argument("--quick", action="store_true", help="the headline workload only (no secondary workloads, no CPU baseline)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--copy-calibration", action="store_true", help="only the 1 GiB copy that scales the HBM counters, then exit")
    ap.add_argument("--no-batch", action="store_true", help="skip the batched ragged GEMVER")
    ap.add_argument("--ns-only", action="store_true",
                    help="only the NS residual lines (4096^2 and 8192^2 f32, 8192^2 bf16); prints their record, not the bench line")
    ap.add_argument("--nsc-only", action="store_true",
                    help="only the compressible NS residual at n = 1024, 2048, 4096 in three variants; prints its record, not the bench line")
    ap.add_argument("--precision-only", action="store_true",
                    help="only GEMVER N=32768 with A stored f32, bf16 and e4m3; prints its record, not the bench line")
    ap.add_argument("--nsc-sizes", default="1024,2048,4096", help="grid sizes n (nx = ny) of --nsc-only, comma separated")
    ap.add_argument("--chain-only", action="store_true",
                    help="only the GEMVER -> BiCGK -> GEMVER chain at n = 16384 in six styles; prints its record, not the bench line")
    ap.add_argument("--plain-only", action="store_true",
                    help="only the GEMVER N=16384 as the reference's four unfused components; prints its record, not the bench line")
    ap.add_argument("--batch-only", action="store_true",
                    help="only the batched ragged GEMVER (512 instances, N in [64, 4096]); prints its record, not the bench line")
    a = ap.parse_args()
    if a.gpus != 1:
        sys.exit(f"bench.py: this build measures one GPU; --gpus {a.gpus} is not supported (how the path would shard is "
                 "described in DESIGN.md)")
    if a.copy_calibration:
        copy_calibration()
        return
    from cospanmdao_amd import check, hip

    check(hip.cs_hip_init(0))
    if a.ns_only:
        extra = {}
        bench_ns(extra, max(20, a.steps // 4))
        print(json.dumps({k: extra[k] for k in ("ns4096_f32", "ns8192_f32", "ns8192_bf16")}), flush=True)
        return
    if a.precision_only:
        extra = {}
        bench_precision(extra, max(20, a.steps // 8))
        print(json.dumps(extra["gemver32768_storage"]), flush=True)
        return
    if a.nsc_only:
        extra = {}
        bench_nsc(extra, max(20, a.steps // 4), tuple(int(v) for v in a.nsc_sizes.split(",")))
        print(json.dumps(extra["ns_compressible"]), flush=True)
        return
    if a.chain_only:
        extra = {}
        bench_chain(extra, max(10, a.steps // 10))
        print(json.dumps(extra["gemver_chain"]), flush=True)
        return
    if a.plain_only:
        p = Gemver(UNFUSED, a.n, 64)
        pms = timed_steps(p, a.steps, 3)
        ln = p.launches(3)
        p.close()
        print(json.dumps({"plain_ms": round(pms, 4), "GBps_own_bytes": round(gbs(gemver_bytes_plain(a.n, 64), pms), 1),
                          "launch_ms": [round(t, 4) for t in ln]}), flush=True)
        return
    if a.batch_only:
        extra = {}
        bench_batch(extra, None, max(30, a.steps // 2))
        print(json.dumps(extra["batched_ragged"]), flush=True)
        return
    n, R = a.n, a.R
    style = TWOPASS if a.style == "twopass" else TWOPASS_1

    g = Gemver(style, n, R, rt2=a.rt2)
    # graph replay first (reported beside the headline), then the headline: the captured launches re-issued on one stream
    ms_graph = timed_steps(g, a.steps, a.warmup)
    g.set_replay(REPLAY_DIRECT)
    ms = timed_steps(g, a.steps, a.warmup)
    launch_ms = g.launches(5)
    # the dominant kernel: pass 1 (reads A and u1 v1 u2 v2 y) or pass 2 (reads A and u1 v1 u2 v2 x, writes w), each timed
    # back to back between two device events on the model's stream
    stage_ms = [g.stage_ms(0, 20), g.stage_ms(2, 20)] if style == TWOPASS else None
    g.close()
    fused_bytes = gemver_bytes_fused(n)
    value = gbs(fused_bytes, ms)
    log(f"fused two-pass GEMVER N={n}: {ms * 1e3:.1f} us/step (graph replay {ms_graph * 1e3:.1f}), {value:.1f} GB/s; launches (us): {[round(t * 1e3, 1) for t in launch_ms]}"
        f"; passes back to back (us): {[round(t * 1e3, 1) for t in stage_ms] if stage_ms else None}")

    traffic = profiled_traffic(n, R, a.rt2)
    if stage_ms:
        k = 0 if stage_ms[0] >= stage_ms[1] else 1
        kbytes = n * n * 4 + (5 * n * 4 if k == 0 else 7 * n * 4)
        kname = "gemver pass 1 (AhatTPartial<f32>)" if k == 0 else f"gemver pass 2 (AhatGemv<f32,{a.rt2}>)"
        kms = stage_ms[k]
        ktraffic = traffic[("pass1", "pass2")[k]] if traffic else None
    else:
        kbytes, kname, kms, ktraffic = fused_bytes, "gemver fused two-pass (one fused group)", ms, None
    achieved = gbs(kbytes, kms)
    roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": PEAK_GBS, "unit": "GB/s", "frac": round(achieved / PEAK_GBS, 4),
                "traffic": ktraffic, "kernel": kname, "algorithmic_bytes": kbytes, "avg_launch_ms": round(kms, 4)}

    extra = {"fused_ms": round(ms, 4), "fused_pct_of_peak": round(100 * value / PEAK_GBS, 2), "launch_ms": [round(t, 4) for t in launch_ms],
             "graph_replay": {"ms": round(ms_graph, 4), "GBps": round(gbs(fused_bytes, ms_graph), 1)},
             "pass_ms_back_to_back": [round(t, 4) for t in stage_ms] if stage_ms else None, "traffic_bytes": traffic}
    extra["errors_vs_fp64"] = measured_errors()
    if style == TWOPASS:
        # the same workload with every step reading an A that the step before did not touch
        cold = bench_cache_cold(extra, a.steps, n, R, a.rt2)
        if stage_ms and k == 1:
            ca = gbs(kbytes, cold["4_copies"]["pass2_launch_ms"])
            roofline["cache_cold"] = {"achieved": round(ca, 1), "frac": round(ca / PEAK_GBS, 4), "avg_launch_ms": cold["4_copies"]["pass2_launch_ms"],
                                      "warm_avg_launch_ms_same_method": round(launch_ms[-1], 4),
                                      "method": "pass 2 per launch (device events between the launches of an eager run) while four "
                                                "separate 1 GiB copies of A run in turn; warm: the same measurement on one A"}
    if not a.quick:
        p = Gemver(UNFUSED, n, 64)
        pms = timed_steps(p, max(3, a.steps // 4), 3)
        p.close()
        extra["plain"] = {"ms": round(pms, 4), "GBps_own_bytes": round(gbs(gemver_bytes_plain(n, 64), pms), 1),
                          "GBps_fused_bytes": round(gbs(fused_bytes, pms), 1), "fused_speedup_vs_plain": round(pms / ms, 3)}
        log(f"plain (4 components, B materialised): {pms * 1e3:.1f} us/step; fused speedup {pms / ms:.2f}x")
        bench_ns(extra, max(20, a.steps // 4))
        if not a.no_batch:
            bench_batch(extra, value, max(30, a.steps // 2))
        bench_precision(extra, max(20, a.steps // 8))
        bench_chain(extra, max(10, a.steps // 10))
        bench_nsc(extra, max(10, a.steps // 10))

    cpu = None
    if not (a.no_cpu or a.quick):
        cpu = cpu_baseline(n, R)
        log(f"cpu baseline: {cpu}")

    out = {
        "metric": METRIC,
        "value": round(value, 1),
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(ms, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (counter-hash uniform[-1,1), seed 0, generated on the device)",
        "config": {"workload": f"GEMVER N={n} f32 fused two-pass (A {n * n * 4 / 2**30:.2f} GiB, read twice), alpha=1.5 beta=1.2",
                   "n": n, "R": R, "pass2_rows": a.rt2, "style": a.style,
                   "replay": "direct (the captured run's three launches re-issued on one stream; graph replay in extra)"},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "extra": extra,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
