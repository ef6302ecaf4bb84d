/*
 * oracle/oracle.h -- the CPU oracle of the fused-run path (TEST INFRASTRUCTURE ONLY).
 *
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load this library, and only as the checker
 * (or, for cpu_baseline, as the timed CPU port). The product path (libcospan_hip and the workload libraries) never
 * links, loads or calls anything here.
 *
 * Every function states the reference file:line it restates (paths relative to the reference checkout). Where a
 * function has no check against the reference's own vectors, closed forms or brute force, its comment says
 * "parity unpinned".
 */
#ifndef CS_ORACLE_H
#define CS_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- deterministic inputs: examples/gemver-chain/common/chain_data.hpp:34-72 ---------------------------------- */
uint32_t or_mix32(uint32_t x);
float    or_hash_unit(uint32_t seed, uint64_t index);
/* rows x cols row-major with leading dimension ld (elements); element (i, j) = hash_unit(seed, i * cols + j) */
void or_fill_hash_f32(uint32_t seed, float* out, uint64_t rows, uint64_t cols, uint64_t ld);
/* the same with the flat index starting at `base` */
void or_fill_hash_f32_base(uint32_t seed, float* out, uint64_t rows, uint64_t cols, uint64_t ld, uint64_t base);
/* pass 1 of the two-pass GEMVER for any n: strips of R rows (the last may be shorter), P is ceil(n/R) x n */
void or_strip_partials_f32(uint64_t n, uint64_t lda, uint64_t R, const float* A, const float* u1, const float* v1, const float* u2,
                           const float* v2, const float* y, float* P);
/* seed of array `id` (1 = A .. 7 = z, chain_data.hpp:32) for a base seed */
uint32_t or_array_seed(uint32_t base_seed, uint32_t id);

/* ---- comparison: chain_data.hpp:147-161 ------------------------------------------------------------------------ */
double or_relative_l2_f32(const float* got, const double* want, uint64_t n);
double or_relative_l2_f64(const double* got, const double* want, uint64_t n);
double or_relative_linf_f32(const float* got, const double* want, uint64_t n);

/* ---- storage formats (no reference counterpart: CoSpan's DType has no bf16/fp8, dtype.hpp:17-20) --------------- */
uint16_t or_f32_to_bf16(float f);   /* round to nearest even, NaN stays NaN */
float    or_bf16_to_f32(uint16_t b);
uint8_t  or_f32_to_e4m3(float f);   /* OCP E4M3 (finite-only), round to nearest even, saturating at +-448 */
float    or_e4m3_to_f32(uint8_t b);
void     or_quantize_bf16(const float* in, uint16_t* out, uint64_t rows, uint64_t cols, uint64_t ld_in, uint64_t ld_out);
/* per-row scale s_i = max_j |a_ij| / 448 (1 for an all-zero row); q_ij = e4m3(a_ij / s_i) */
void     or_quantize_e4m3_rows(const float* in, uint8_t* out, float* scale, uint64_t rows, uint64_t cols, uint64_t ld_in,
                               uint64_t ld_out);

/* ---- GEMVER, FP64: chain_data.hpp:93-117 (B recomputed on the fly), alpha and beta as parameters ---------------
 * store: 0 = f32 A, 1 = bf16 A (uint16_t), 2 = e4m3 A (uint8_t) times row_scale[i]. The stored value, dequantised,
 * is the matrix: the oracle of a reduced-precision run is the exact GEMVER of what is stored. */
void or_gemver_f64(uint64_t n, int store, const void* A, uint64_t lda, const float* row_scale, const float* u1,
                   const float* v1, const float* u2, const float* v2, const float* y, const float* z, double alpha,
                   double beta, double* x, double* w);

/* ---- GEMVER, f32 restatement of the four components (components.hpp:85-266), fixed evaluation order ----------- */
void or_rank2_update_f32(uint64_t n, uint64_t R, const float* A, const float* u1, const float* v1, const float* u2,
                         const float* v2, float* B);
void or_gemvt_partial_f32(uint64_t n, uint64_t R, const float* B, const float* y, float* P);
void or_col_reduce_f32(uint64_t n, uint64_t strips, const float* P, const float* z, float scale, float* x);
void or_gemv_f32(uint64_t n, const float* B, const float* x, float scale, float* y);

/* the GEMVER -> BiCGK -> GEMVER chain in FP64 (chain_data.hpp:88-147); B2 (n x n) may be NULL */
void or_chain_f64(uint64_t n, const float* A0, const float* u1, const float* v1, const float* u2, const float* v2, const float* y,
                  const float* z, double alpha, double beta, double* x1, double* w1, double* q, double* s, double* x2, double* w2,
                  double* B2);
/* BiCGK's one-pass partials in this build's order: Ps (n / R x n), Pq (n x n / 256); and q from Pq */
void or_bicgk_onepass_f32(uint64_t n, uint64_t R, const float* B, const float* p, const float* r, float* Ps, float* Pq);
void or_row_reduce_f32(uint64_t n, uint64_t blocks, const float* Pq, float* q);
/* the fused two-pass form on the CPU (Â never stored): x bitwise equal to the components' x for the same R */
void or_gemver_twopass_f32(uint64_t n, uint64_t R, const float* A, const float* u1, const float* v1, const float* u2,
                           const float* v2, const float* y, const float* z, float alpha, float beta, float* P, float* x,
                           float* w);

/* ---- scans and ragged columns: cospanmdao/include/naga/cospan/device_scan.hpp:43-110, shape_hook.hpp:180-338 -- */
void     or_exclusive_scan_u32(const uint32_t* in, uint32_t* out, uint64_t n);
void     or_exclusive_scan_u64(const uint64_t* in, uint64_t* out, uint64_t n);
/* src_dtype: 0 i32, 1 u32, 2 i64, 3 u64. counts mode: src is [count_src, k] (column 0 used); offsets mode: src holds
 * count_src = N + 1 offsets. Writes extents[N], offs[N + 1], *total; returns the flag word (1 negative/invalid,
 * 2 past 32 bits or wrapped). */
uint32_t or_build_column(int src_dtype, const void* src, uint64_t count_src, uint64_t k, int offsets_mode,
                         uint32_t* extents, uint32_t* offs, uint32_t* total);

/* ---- batched ragged GEMVER sizes: N_k = round(lo * (hi / lo)^U_k), U_k from the hash (seed, k) ------------------ */
void or_ragged_sizes(uint32_t seed, uint32_t count, uint32_t lo, uint32_t hi, uint32_t* sizes);

/* ---- incompressible NS residual, FP64 (structure of examples/ns-residual/reference.hpp:30-130; physics new) --- */
/* Fields: sum of 8 Fourier modes per field, FP64 then rounded to f32. */
void or_ns_fields(uint32_t seed, int nx, int ny, float* u, float* v, float* p);
/* store: 0 f32 fields, 1 bf16 fields. Writes Ru, Rv, Rc (nx*ny each, may be NULL) and sq[3] = sum of squares. */
void or_ns_residual_f64(int nx, int ny, double h, double nu, int store, const void* u, const void* v, const void* p,
                        double* Ru, double* Rv, double* Rc, double* sq);
/* The same residual compared on the fly with a device result (no FP64 field is stored): stats[6] = rel-L2 and
 * relative inf-norm error of Ru, Rv, Rc; sq[3] = the FP64 sums of squares. */
void or_ns_compare(int nx, int ny, double h, double nu, int store, const void* u, const void* v, const void* p,
                   const float* Ru, const float* Rv, const float* Rc, double* stats, double* sq);

/* ---- compressible NS residual, FP64: examples/ns-residual/physics.hpp (numerics), reference.hpp:29-128 --------- */
/* U = (rho, rho u, rho v, E) as [4, ny, nx] f32; walls: no-slip adiabatic walls in y (else periodic); vanleer: the van
 * Leer limiter (else minmod). Writes R [4, ny, nx] and diag = (min dt, max wave speed, sum R^2 of the 4 variables). */
void   or_nsc_residual_f64(int nx, int ny, int walls, int vanleer, const float* U, double* R, double* diag);
/* The pieces, for pinning: conserved -> primitives (rho, u, v, p); the limited slope of (backward a, forward b); HLLC
 * with Davis speeds through a face with normal x (dir 0) or y (dir 1); the viscous flux from the face's two cells and
 * their gradients (ux, uy, vx, vy, Tx, Ty); a cell's local timestep and wave speed. */
void   or_nsc_prim_f64(const double* U, double* W);
double or_nsc_slope_f64(double a, double b, int vanleer);
void   or_nsc_hllc_f64(const double* wl, const double* wr, int dir, double* F);
void   or_nsc_visc_f64(const double* wl, const double* wr, const double* gl, const double* gr, int dir, double* F);
void   or_nsc_cell_diag_f64(const double* W, double* dt, double* ws);
/* Uniform flow (0.3, 0.2) at rho = p = 1 plus a perturbation of amplitude amp and wavelength 32 cells, in f32 with the
 * host's sinf / cosf (the device generator uses its own: tests read the device's field back). */
void   or_nsc_init_f32(int nx, int ny, float amp, float* U);

/* ---- threads used by the OpenMP loops here ---------------------------------------------------------------------- */
int or_threads(void);

#ifdef __cplusplus
}
#endif
#endif
