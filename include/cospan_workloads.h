/*
 * include/cospan_workloads.h — the example models of the fused-run path behind a C boundary.
 *
 * The reference drives its two examples from C++ programs (examples/gemver-chain/main.cpp, examples/ns-residual/main.cpp:
 * --variant, --n, --iters, --check). This library holds the same models, built through the CoSpan API of this
 * repository (Component / System / Model, fuse_components), and exposes each as an opaque handle so that the tests,
 * the benchmark and a host program in any language drive them through plain C: create, run, wait, read an output, time.
 * Inputs are generated on the device by the models' own initializers, bit-identical to the host generators of the
 * reference (examples/gemver-chain/common/chain_data.hpp:34-72) and of the oracle.
 */
#ifndef COSPAN_WORKLOADS_H
#define COSPAN_WORKLOADS_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- GEMVER --------------------------------------------------------------------------------------------------------
 * replaces: the Gemver System of examples/gemver-chain/systems.hpp:71-139 and its driver (main.cpp:154-234).
 * Styles (the first three are the reference's, systems.hpp:33-48):
 *   CS_GEMVER_UNFUSED   four components, four launches (B materialised)
 *   CS_GEMVER_FUSED     fuse_components(segment<range_mode::whole>(r2u, gtp)), then cred, gemv: three launches
 *   CS_GEMVER_NAIVE     one launch: segment<whole>(r2u, gtp), global_barrier, cred, global_barrier, gemv
 *   CS_GEMVER_TWOPASS   this build's form: Â never stored, A read twice (pass 1, cred, pass 2): three launches
 *   CS_GEMVER_TWOPASS_1 the same three stages fused with global barriers: one launch
 * Storage of A (store): 0 f32, 1 bf16, 2 e4m3 with a per-row f32 scale (TWOPASS styles only).
 * Inputs: A[i][j] = hash_unit(seed_A, i * n + j), u1 .. z likewise with their own seeds (base_seed * 16 + 1 .. 7, the
 * reference's seed_A .. seed_z for base_seed 0). alpha scales w, beta scales x.
 */
#define CS_GEMVER_UNFUSED 0
#define CS_GEMVER_FUSED 1
#define CS_GEMVER_NAIVE 2
#define CS_GEMVER_TWOPASS 3
#define CS_GEMVER_TWOPASS_1 4

typedef struct cs_gemver_s* cs_gemver;

int cs_gemver_create(cs_gemver* g, int style, int store, size_t n, size_t R, float alpha, float beta, uint32_t base_seed);
/* the work-group shapes of the TWOPASS styles: rows per pass-2 work-group (4, 8 or 16) */
int cs_gemver_set_tile(cs_gemver g, int pass2_rows);
int cs_gemver_run(cs_gemver g);
int cs_gemver_wait(cs_gemver g);
/* copies an output or input to host memory: "x", "w", "B", "P", "A" (as stored), "scale" (e4m3 row scales), "u1" .. "z" */
int cs_gemver_read(cs_gemver g, const char* name, void* host, size_t bytes);
/* `iters` back-to-back runs bracketed by device events on the model's stream; average milliseconds per run */
int cs_gemver_time(cs_gemver g, int iters, float* ms_per_run);
/* `iters` forwards with a device event around every launch: the average milliseconds of each launch of one forward, in
 * launch order (up to `cap` entries); *count = the launches of one forward */
int cs_gemver_time_launches(cs_gemver g, int iters, float* ms, int cap, int* count);
/* the model's executor counters: runs, kernel launches, graph captures, graph replays */
int cs_gemver_stats(cs_gemver g, uint64_t* stats4);
int cs_gemver_destroy(cs_gemver g);

#ifdef __cplusplus
}
#endif
#endif
