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

/* The identity of the sources this library was built from (16 hex digits, cospanmdao_amd/source_hash.py; no reference
 * counterpart). Tests compare it with the tree. */
const char* cs_workloads_build_id(void);

/* ---- GEMVER --------------------------------------------------------------------------------------------------------
 * replaces: the Gemver System of examples/gemver-chain/systems.hpp:71-139 and its driver (main.cpp:154-234).
 * Styles (the first three are the reference's, systems.hpp:33-48):
 *   CS_GEMVER_UNFUSED   four components, four launches (B materialised)
 *   CS_GEMVER_FUSED     fuse_components(segment<range_mode::whole>(r2u, gtp)), then cred, gemv: three launches
 *   CS_GEMVER_NAIVE     one fused group: segment<whole>(r2u, gtp), global_barrier, cred, global_barrier, gemv
 *   CS_GEMVER_TWOPASS   this build's form: Â never stored, A read twice (pass 1, cred, pass 2): three launches
 *   CS_GEMVER_TWOPASS_1 the same three stages as one fused group with global barriers
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
/* the work-group shapes of the TWOPASS styles: rows per pass-2 work-group (4 or 8: the rows spread over the work-group's
 * four waves; 16 or 32: each wave owns a quarter of the rows and reduces them alone) */
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
/* the steady-state milliseconds of leaf `leaf` (0-based, launch order) run `iters` times back to back between two device
 * events (the leaves before it run once first) */
int cs_gemver_time_stage(cs_gemver g, int leaf, int iters, float* ms);
/* the model's executor counters: runs, kernel launches, graph captures, graph replays */
int cs_gemver_stats(cs_gemver g, uint64_t* stats4);
/* how steady-state runs are replayed once captured (no counterpart in the reference, whose runs are always submitted
 * eagerly): CS_REPLAY_GRAPH (the default) one hipGraphLaunch per run; CS_REPLAY_DIRECT the captured launches re-issued in
 * order on the model's stream with no events, when the capture stayed on that one stream (a run spanning streams still
 * replays its graph). Same results either way. */
#define CS_REPLAY_GRAPH 0
#define CS_REPLAY_DIRECT 1
int cs_gemver_set_replay(cs_gemver g, int mode);
int cs_gemver_destroy(cs_gemver g);

/* ---- incompressible Navier-Stokes residual ---------------------------------------------------------------------------
 * replaces: the fused / unfused variants of examples/ns-residual (systems.hpp:88-140, components.hpp:636-852, driver
 * main.cpp:159-199) with this build's incompressible physics on a periodic nx x ny grid:
 *   Ru = u u_x + v u_y + p_x - nu lap u,  Rv = u v_x + v v_y + p_y - nu lap v,  Rc = u_x + v_y  (central differences),
 * then the FP64 sums of squares ("sq", 3 doubles) and L2 norms ("norm": |Ru|, |Rv|, |Rc|, all three; 4 doubles).
 *   CS_NS_FUSED    one fused group: the tile stencil, a global barrier, the fixed-order reduction of the tile partials
 *   CS_NS_UNFUSED  four launches: derivatives to global memory, combine, squares, norm
 * store: 0 f32 fields, 1 bf16 fields (arithmetic f32, norms FP64). tile_rows: rows per work-group of the fused stencil.
 * Fields: u, v, p from 8 random Fourier modes each (seeded), generated on the device. Names for cs_ns_read: "u" "v" "p"
 * (as stored), "Ru" "Rv" "Rc" (f32), "sq" "norm" (f64).
 */
#define CS_NS_FUSED 0
#define CS_NS_UNFUSED 1

typedef struct cs_ns_s* cs_ns;

int cs_ns_create(cs_ns* m, int style, int store, size_t nx, size_t ny, double h, double nu, uint32_t seed, int tile_rows);
int cs_ns_run(cs_ns m);
int cs_ns_wait(cs_ns m);
int cs_ns_read(cs_ns m, const char* name, void* host, size_t bytes);
/* replaces a field ("u", "v", "p") with host values in its storage format before the next run (update_input) */
int cs_ns_write(cs_ns m, const char* name, const void* host, size_t bytes);
int cs_ns_time(cs_ns m, int iters, float* ms_per_run);
int cs_ns_time_launches(cs_ns m, int iters, float* ms, int cap, int* count);
/* CS_REPLAY_GRAPH or CS_REPLAY_DIRECT, as cs_gemver_set_replay */
int cs_ns_set_replay(cs_ns m, int mode);
int cs_ns_destroy(cs_ns m);

/* ---- compressible Navier-Stokes residual ----------------------------------------------------------------------------
 * replaces: the unfused and fused variants of examples/ns-residual (systems.hpp:26-140) and the driver's actions
 * (main.cpp:159-199, 304-329): R = -div F_conv + div F_visc of U = (rho, rho u, rho v, E) [4, ny, nx] f32 (gamma 1.4, mu 0.01,
 * Pr 0.72, dx = dy = 1; MUSCL with the minmod or van Leer limiter, HLLC, viscous flux; x periodic, y periodic or no-slip
 * adiabatic walls) and the diagnostics min local timestep, max wave speed, sum R^2 per variable, atomic and deterministic.
 *   CS_NSC_UNFUSED   six launches: init, tile, three reductions, error
 *   CS_NSC_FUSED     one launch with group waits (the reference's layout: init | tile | barrier | reductions | error)
 *   CS_NSC_LAUNCHES  the same fused group with barriers after init, the tile and the reductions: four launches
 * nx a multiple of 64, ny of tile_rows (8 or 16: the rows of a work-group's tile). cs_nsc_init sets U from the device
 * generator (uniform flow plus a perturbation of amplitude amp; 0 gives uniform flow), cs_nsc_write from the host.
 * Names for cs_nsc_read (f32): "U", "R", "Part" [6, tiles], "Dt" "Ws" (1) "Sq" (4) atomic, "DetDt" "DetWs" "DetSq"
 * deterministic, "Err" (6: atomic - deterministic for dt, ws, sq0..3).
 */
#define CS_NSC_UNFUSED 0
#define CS_NSC_FUSED 1
#define CS_NSC_LAUNCHES 2
#define CS_NSC_MINMOD 0
#define CS_NSC_VANLEER 1

typedef struct cs_nsc_s* cs_nsc;

int cs_nsc_create(cs_nsc* m, int variant, size_t nx, size_t ny, int limiter, int walls, int tile_rows);
int cs_nsc_init(cs_nsc m, float amp);
int cs_nsc_write(cs_nsc m, const void* host_U, size_t bytes);
int cs_nsc_run(cs_nsc m);
int cs_nsc_wait(cs_nsc m);
int cs_nsc_read(cs_nsc m, const char* name, void* host, size_t bytes);
int cs_nsc_time(cs_nsc m, int iters, float* ms_per_run);
int cs_nsc_time_launches(cs_nsc m, int iters, float* ms, int cap, int* count);
/* the model's executor counters: runs, kernel launches, graph captures, graph replays */
int cs_nsc_stats(cs_nsc m, uint64_t* stats4);
int cs_nsc_destroy(cs_nsc m);

/* ---- the GEMVER -> BiCGK -> GEMVER chain ------------------------------------------------------------------------------
 * replaces: the Chain System of examples/gemver-chain/systems.hpp:226-346 in its six styles and the driver's actions
 * (main.cpp:154-188). G1: B1 = A + u1 v1^T + u2 v2^T, x1 = beta B1^T ya + za, w1 = alpha B1 x1; BiCGK: q = B1 x1,
 * s = B1^T w1; G2 on A = B1 with the same u1 .. v2, y = q, z = s: B2, x2, w2. alpha = beta = 1/sqrt(n) (the reference's
 * chain, chain_data.hpp:13). Inputs as for GEMVER (base_seed 0: the reference's generator); G2's u1 .. v2 equal G1's.
 * Styles (systems.hpp:33-53):
 *   CS_CHAIN_UNFUSED   every component its own launch
 *   CS_CHAIN_PARALLEL  as unfused, BiCGK's two products an ordinary System of independent children
 *   CS_CHAIN_FUSED     GEMVER's rank-2 update + partial column sums one launch; BiCGK's two products one launch
 *   CS_CHAIN_ONEPASS   fused GEMVERs, BiCGK from one pass over B1
 *   CS_CHAIN_NAIVE     three fused groups with global barriers (G1, BiCGK, G2)
 *   CS_CHAIN_CRAFTED   three fused groups over local-memory tiles of 16 rows (R is ignored): BiCGK and G2's rank-2 update
 *                      share one read of B1
 * n: a multiple of 256 and of R. Names for cs_chain_read (f32): "x1" "w1" "q" "s" "x2" "w2" (n), "B1" "B2" (n x n),
 * and the inputs "A" "u1a" .. "za" "u1b" .. "v2b".
 */
#define CS_CHAIN_UNFUSED 0
#define CS_CHAIN_PARALLEL 1
#define CS_CHAIN_FUSED 2
#define CS_CHAIN_ONEPASS 3
#define CS_CHAIN_NAIVE 4
#define CS_CHAIN_CRAFTED 5

typedef struct cs_chain_s* cs_chain;

int cs_chain_create(cs_chain* c, int style, size_t n, size_t R, uint32_t base_seed);
int cs_chain_run(cs_chain c);
int cs_chain_wait(cs_chain c);
int cs_chain_read(cs_chain c, const char* name, void* host, size_t bytes);
/* `iters` chains back to back between two device events, then one wait: *ms_per_run = the average */
int cs_chain_time(cs_chain c, int iters, float* ms_per_run);
/* per-launch device times of one eager run, averaged over `iters` runs (launch order) */
int cs_chain_time_launches(cs_chain c, int iters, float* ms, int cap, int* count);
int cs_chain_stats(cs_chain c, uint64_t* stats4);
int cs_chain_destroy(cs_chain c);

/* ---- batched ragged GEMVER -------------------------------------------------------------------------------------------
 * No reference counterpart as one launch (the reference refuses ragged ports inside fused groups,
 * cospanmdao/include/naga/cospan/fused_component.hpp:207-213); the offsets come from the column build it uses for
 * ragged shapes (shape_hook.hpp:215-348). `count` instances with N_k = round(lo (hi/lo)^U_k) (U_k from the hash (seed, k));
 * each x_k = beta Â_k^T y_k + z_k, w_k = alpha Â_k x_k as GEMVER above. One forward = a plan kernel, ONE device scan of
 * the per-instance counts, a unit map (every phase's units in the work order: largest instances first), then (phased
 * form, the default) one launch per phase over every instance: pass 1, the column reductions, pass 2 -- or (persistent
 * form) ONE work-queue launch that runs every unit of the three phases; runs from the second on replay one captured graph.
 * Names for cs_batch_read (u32 or f32 words): "sizes", "counts" (9 x count), "offs" (9 x count + 1: the scan),
 * "result" ([total, flag] of the scan), "x", "w", "u1" .. "z" (instances back to back).
 * params: R (strip rows, >= 32), RT (pass-2 rows per work-group: 4 or 8, the work-group shares a row tile; 16 or 32, a
 * wave per RT/4 rows), G (persistent work-groups; the phased form ignores it).
 */
typedef struct cs_batch_s* cs_batch;

int cs_batch_create(cs_batch* b, uint32_t count, uint32_t lo, uint32_t hi, float alpha, float beta, uint32_t base_seed);
int cs_batch_sizes(cs_batch b, uint32_t* sizes, size_t count);
int cs_batch_set_params(cs_batch b, uint32_t R, uint32_t RT, uint32_t groups);
int cs_batch_get_params(cs_batch b, uint32_t* params3);
/* 0: the phased form (default), 1: the persistent work-queue form; a change rebuilds the model at its next use */
int cs_batch_set_mode(cs_batch b, int persistent);
/* times every candidate (R in {64, 128, 256}, RT in {8, 16, 32}; persistent form also G in {2, 3, 4} x CUs) on this
 * model and keeps the fastest; cache_path (or NULL): a text file of "key R RT G" lines keyed by device, sizes and form,
 * read first and appended to */
int cs_batch_autotune(cs_batch b, int reps, const char* cache_path, uint32_t* best3);
int cs_batch_run(cs_batch b);
int cs_batch_wait(cs_batch b);
int cs_batch_read(cs_batch b, const char* name, void* host, size_t bytes);
int cs_batch_time(cs_batch b, int iters, float* ms_per_run);
/* `iters` back-to-back runs with a device event after each: ms[i] = the i-th run's time (for percentiles) */
int cs_batch_time_each(cs_batch b, int iters, float* ms);
/* one leaf of the model (0 plan, 1 scan, 2 unit map, then 3 pass 1, 4 column reductions, 5 pass 2 in the phased form,
 * or 3 the persistent kernel) run `iters` times back to back after the leaves before it: *ms = its average device time */
int cs_batch_time_stage(cs_batch b, int leaf, int iters, float* ms);
int cs_batch_stats(cs_batch b, uint64_t* stats4);
int cs_batch_destroy(cs_batch b);

#ifdef __cplusplus
}
#endif
#endif
