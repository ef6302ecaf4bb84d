"""Benchmark of the fused-run path on one MI355X: prints ONE JSON line.

Workload (the configuration the metric is quoted on): GEMVER N = 16384, f32, A uniform[-1, 1] (1 GiB), alpha = 1.5,
beta = 1.2, seed 0, run as ONE fused two-pass model (Â never stored; A read twice). A step is one forward of the model
(x = beta Â^T y + z, then w = alpha Â x) with every input resident in HBM. `value` is algorithmic HBM bytes per step
(2 N^2 * 4 + 13 N * 4 = 2 148 335 616 B) over the measured time per step, in GB/s.

Beside it: the same GEMVER as the reference's four unfused components (B materialised), the per-launch times and the
roofline of the dominant kernel, and the CPU baseline (the oracle's f32 two-pass port with OpenMP on the host cores, on
a bounded sample). More than one GPU is not part of this build: --gpus N > 1 exits with a message.
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


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def gemver_bytes_fused(n):
    return 2 * n * n * 4 + 13 * n * 4


def gemver_bytes_plain(n, R):
    return 4 * n * n * 4 + 2 * (n // R) * n * 4 + 13 * n * 4


class Gemver:
    def __init__(self, style, n, R, store=0, rt2=8):
        from cospanmdao_amd import check, workloads

        self.w, self.check = workloads, check
        self.h = C.c_void_p()
        check(workloads.cs_gemver_create(C.byref(self.h), style, store, n, R, 1.5, 1.2, 0))
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

    def close(self):
        self.check(self.w.cs_gemver_destroy(self.h))


def timed_steps(model, steps, warmup):
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


def cpu_baseline(n, R, budget_s=12.0):
    """The oracle's f32 two-pass GEMVER (OpenMP) on the host cores: a bounded sample of the same workload."""
    import numpy as np

    import oracle as O

    L = O.lib()
    inp = O.gemver_inputs(n)
    P = np.empty((n // R, n), np.float32)
    x = np.empty(n, np.float32)
    w = np.empty(n, np.float32)
    args = [O.ptr(inp[k]) for k in ("A", "u1", "v1", "u2", "v2", "y", "z")]
    L.or_gemver_twopass_f32(n, R, *args, 1.5, 1.2, O.ptr(P), O.ptr(x), O.ptr(w))  # warm-up
    reps, t0 = 0, time.perf_counter()
    while reps < 20 and (reps == 0 or time.perf_counter() - t0 < budget_s):
        L.or_gemver_twopass_f32(n, R, *args, 1.5, 1.2, O.ptr(P), O.ptr(x), O.ptr(w))
        reps += 1
    sec = (time.perf_counter() - t0) / reps
    return {
        "value": round(gemver_bytes_fused(n) / sec / 1e9, 2),
        "unit": "GB/s",
        "cores": int(L.or_threads()),
        "kind": "port",
        "sample": f"{reps} forwards of the f32 two-pass GEMVER N={n} (oracle/oracle.c or_gemver_twopass_f32, OpenMP), "
        f"{sec * 1e3:.1f} ms each",
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n", type=int, default=16384)
    ap.add_argument("--R", type=int, default=128)
    ap.add_argument("--rt2", type=int, default=8)
    ap.add_argument("--style", default="twopass", choices=["twopass", "twopass_1"])
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    a = ap.parse_args()
    if a.gpus != 1:
        sys.exit(f"bench.py: this build measures one GPU; --gpus {a.gpus} is not supported (how the path would shard is "
                 "described in DESIGN.md)")
    from cospanmdao_amd import check, hip

    check(hip.cs_hip_init(0))
    n, R = a.n, a.R
    style = TWOPASS if a.style == "twopass" else TWOPASS_1

    g = Gemver(style, n, R, rt2=a.rt2)
    ms = timed_steps(g, a.steps, a.warmup)
    launch_ms = g.launches(max(5, a.steps // 2))
    g.close()
    fused_bytes = gemver_bytes_fused(n)
    value = fused_bytes / (ms * 1e-3) / 1e9
    log(f"fused two-pass GEMVER N={n}: {ms * 1e3:.1f} us/step, {value:.1f} GB/s; launches (us): "
        f"{[round(t * 1e3, 1) for t in launch_ms]}")

    # the dominant kernel: pass 1 (A + u1 v1 u2 v2 y) or pass 2 (A + u1 v1 u2 v2 x + w)
    if style == TWOPASS and len(launch_ms) == 3:
        k = 0 if launch_ms[0] >= launch_ms[2] else 2
        kbytes = n * n * 4 + (5 * n * 4 if k == 0 else 7 * n * 4)
        kname = "gemver pass 1 (AhatTPartial<f32>)" if k == 0 else f"gemver pass 2 (AhatGemv<f32,{a.rt2}>)"
        kms = launch_ms[k]
    else:
        kbytes, kname, kms = fused_bytes, "gemver fused two-pass (one launch)", launch_ms[0]
    achieved = kbytes / (kms * 1e-3) / 1e9
    roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / PEAK_GBS, 4), "traffic": None, "kernel": kname,
                "algorithmic_bytes": kbytes, "avg_launch_ms": round(kms, 4)}

    extra = {"fused_ms": round(ms, 4), "fused_pct_of_peak": round(100 * value / PEAK_GBS, 2),
             "launch_ms": [round(t, 4) for t in launch_ms]}
    if not a.no_plain:
        p = Gemver(UNFUSED, n, 64)
        pms = timed_steps(p, max(3, a.steps // 2), 2)
        p.close()
        extra.update({"plain_ms": round(pms, 4), "plain_GBps_own_bytes": round(gemver_bytes_plain(n, 64) / (pms * 1e-3) / 1e9, 1),
                      "plain_GBps_fused_bytes": round(fused_bytes / (pms * 1e-3) / 1e9, 1),
                      "fused_speedup_vs_plain": round(pms / ms, 3)})
        log(f"plain (4 components, B materialised): {pms * 1e3:.1f} us/step; fused speedup {pms / ms:.2f}x")

    cpu = None
    if not a.no_cpu:
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
                   "n": n, "R": R, "pass2_rows": a.rt2, "style": a.style},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "extra": extra,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
