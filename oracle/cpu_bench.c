/*
 * oracle/cpu_bench.c — one timed process of the GEMVER CPU baseline (TEST AND MEASUREMENT INFRASTRUCTURE: run only by
 * bench.py's cpu_baseline leg, never by the product).
 *
 * The reference's timing protocol (examples/gemver-chain/main.cpp:154-169 and examples/run_benchmarks.nu:9-22): inside the
 * process, 3 warm-up forwards each followed by a wait, then 20 forwards timed together; the caller runs five such
 * processes and takes the median. Its CPU path (AdaptiveCpp's OpenMP backend) cannot be built here, so a forward is the
 * oracle's f32 restatement of the same computation, on all OpenMP threads:
 *
 *   cpu_bench components N R   the reference's four components in order (Rank2Update -> B, GemvTPartial -> P,
 *                              ColReduce -> x, Gemv -> w), examples/gemver-chain/components.hpp:85-250
 *   cpu_bench twopass N R      the fused two-pass form (A read twice, B never stored), or_gemver_twopass_f32
 *
 * Inputs are the counter-hash arrays of chain_data.hpp:34-63 (seed 0 family), alpha 1.5, beta 1.2. Prints one line:
 *   {"mode": ..., "n": N, "R": R, "ms": <mean ms per forward over the 20>, "threads": T, "checksum": <sum of w>}
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include "oracle.h"

static double now_ms(void) {
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return (double)ts.tv_sec * 1e3 + (double)ts.tv_nsec * 1e-6;
}

static void* must(size_t bytes) {
    void* p = malloc(bytes ? bytes : 1);
    if (!p) {
        fprintf(stderr, "cpu_bench: out of memory (%zu bytes)\n", bytes);
        exit(2);
    }
    return p;
}

int main(int argc, char** argv) {
    if (argc != 4 || (strcmp(argv[1], "components") != 0 && strcmp(argv[1], "twopass") != 0)) {
        fprintf(stderr, "usage: cpu_bench components|twopass N R\n");
        return 2;
    }
    const int      components = strcmp(argv[1], "components") == 0;
    const uint64_t n = strtoull(argv[2], NULL, 10), R = strtoull(argv[3], NULL, 10);
    if (n == 0 || R == 0 || n % R != 0) {
        fprintf(stderr, "cpu_bench: N must be a positive multiple of R\n");
        return 2;
    }
    const float alpha = 1.5f, beta = 1.2f;
    float*      A     = must(sizeof(float) * n * n);
    float*      B     = components ? must(sizeof(float) * n * n) : NULL;
    float*      P     = must(sizeof(float) * (n / R) * n);
    float*      vec[6];
    for (int k = 0; k < 6; ++k) vec[k] = must(sizeof(float) * n);
    float* x = must(sizeof(float) * n);
    float* w = must(sizeof(float) * n);
    or_fill_hash_f32(or_array_seed(0, 1), A, n, n, n);
    for (int k = 0; k < 6; ++k) or_fill_hash_f32(or_array_seed(0, 2 + (uint32_t)k), vec[k], 1, n, n);
    float *u1 = vec[0], *v1 = vec[1], *u2 = vec[2], *v2 = vec[3], *y = vec[4], *z = vec[5];

    double t0 = 0.0;
    for (int it = 0; it < 23; ++it) {
        if (it == 3) t0 = now_ms();  /* 3 warm-ups, then 20 timed forwards */
        if (components) {
            or_rank2_update_f32(n, R, A, u1, v1, u2, v2, B);
            or_gemvt_partial_f32(n, R, B, y, P);
            or_col_reduce_f32(n, n / R, P, z, beta, x);
            or_gemv_f32(n, B, x, alpha, w);
        } else {
            or_gemver_twopass_f32(n, R, A, u1, v1, u2, v2, y, z, alpha, beta, P, x, w);
        }
    }
    const double ms = (now_ms() - t0) / 20.0;
    double       checksum = 0.0;
    for (uint64_t i = 0; i < n; ++i) checksum += (double)w[i];
    printf("{\"mode\": \"%s\", \"n\": %llu, \"R\": %llu, \"ms\": %.6f, \"threads\": %d, \"checksum\": %.9g}\n", argv[1],
           (unsigned long long)n, (unsigned long long)R, ms, or_threads(), checksum);
    free(A);
    free(B);
    free(P);
    for (int k = 0; k < 6; ++k) free(vec[k]);
    free(x);
    free(w);
    return 0;
}
