/*
 * oracle/oracle.c -- the CPU oracle of the fused-run path (TEST INFRASTRUCTURE ONLY; see oracle.h).
 *
 * Built with -ffp-contract=off: every fused multiply-add here is an explicit fmaf()/fma(), so the f32 restatements
 * fix one evaluation order that the device kernels reproduce bit for bit where the summation order is the same.
 *
 * Pinning (how each part is checked beyond restating it):
 *   hash / generators     against the reference's own chain_data.hpp, compiled from its sources (oracle/_ref)
 *   or_gemver_f64         against chain_data.hpp's reference_chain (x1, w1 at alpha = beta = 1/sqrt(n)) through
 *                         oracle/_ref; closed forms (u = v = 0 gives the textbook GEMV; y = 0 gives x = z)
 *   or_chain_f64          against reference_chain bit for bit (x1 w1 q s x2 w2) through oracle/_ref; B2 against the
 *                         closed form A0 + 2 (u1 v1^T + u2 v2^T)
 *   or_bicgk_onepass_f32, against or_chain_f64's q and s (tolerance), a brute-force scan of the DPP order at n = 256, and
 *   or_row_reduce_f32     plain sums on integer-valued data (exact in f32)
 *   f32 components        against or_gemver_f64 (tolerance) and brute force at tiny n; rank-2 update bitwise
 *                         against a plain fmaf loop
 *   scans / columns       std::exclusive_scan known answers of tests/test_device_scan.cpp:57-87 and
 *                         tests/test_shapes.cpp:656-682 (golden fixtures under tests/golden)
 *   bf16 / e4m3           exhaustive round trips over every code, round-to-nearest-even tie cases
 *   NS residual           uniform fields give exactly 0; a single Fourier mode matches the analytic derivatives to
 *                         O(h^2)
 *   or_ns_fields          its Fourier transform: all energy on integer wavenumbers in [-4, 4]^2, no mean, at most 8
 *                         conjugate pairs (7.3 on average), the 1/|k|^2 amplitude law with unit-variance draws
 *                         (Parseval), uniform phases
 *   or_relative_l2_*      against the reference's relative_l2 (chain_data.hpp:147-161) through oracle/_ref, bit for
 *                         bit, with zero vectors, NaN and infinity; closed forms
 *   or_ragged_sizes       bounds and the log-uniform spread only: parity unpinned (no reference counterpart)
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

int or_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

/* ---- hash: examples/gemver-chain/common/chain_data.hpp:34-49 -------------------------------------------------- */
uint32_t or_mix32(uint32_t x) {
    x ^= x >> 16;
    x *= 0x7feb352du;
    x ^= x >> 15;
    x *= 0x846ca68bu;
    x ^= x >> 16;
    return x;
}

float or_hash_unit(uint32_t seed, uint64_t index) {
    const uint32_t lo = (uint32_t)index, hi = (uint32_t)(index >> 32);
    const uint32_t h  = or_mix32(or_mix32(seed * 0x9e3779b9u + 0x85ebca6bu) ^ or_mix32(lo + or_mix32(hi + seed)));
    return (float)(h >> 8) * (1.0f / 8388608.0f) - 1.0f;
}

uint32_t or_array_seed(uint32_t base_seed, uint32_t id) { return base_seed * 16u + id; }

/* chain_data.hpp:52-63 (make_vector / make_matrix_A0): the flat index is i * cols + j whatever the padding */
void or_fill_hash_f32(uint32_t seed, float* out, uint64_t rows, uint64_t cols, uint64_t ld) {
#pragma omp parallel for schedule(static)
    for (long long i = 0; i < (long long)rows; ++i)
        for (uint64_t j = 0; j < cols; ++j) out[(uint64_t)i * ld + j] = or_hash_unit(seed, (uint64_t)i * cols + j);
}

/* the same with the flat index starting at `base` (instance k of a batch: base = sum of the earlier instances' sizes) */
void or_fill_hash_f32_base(uint32_t seed, float* out, uint64_t rows, uint64_t cols, uint64_t ld, uint64_t base) {
#pragma omp parallel for schedule(static)
    for (long long i = 0; i < (long long)rows; ++i)
        for (uint64_t j = 0; j < cols; ++j) out[(uint64_t)i * ld + j] = or_hash_unit(seed, base + (uint64_t)i * cols + j);
}

/* Pass 1 of the two-pass GEMVER for any n (the last strip may be shorter): P[c][j] = sum over the strip's rows of
 * fma(rank-2 updated a_ij, y_i, acc), rows ascending; the order of components.hpp:85-97 + :142-155. */
void or_strip_partials_f32(uint64_t n, uint64_t lda, uint64_t R, const float* A, const float* u1, const float* v1, const float* u2,
                           const float* v2, const float* y, float* P) {
    const uint64_t strips = (n + R - 1) / R;
#pragma omp parallel for schedule(static)
    for (long long c = 0; c < (long long)strips; ++c) {
        float* p = P + (uint64_t)c * n;
        for (uint64_t j = 0; j < n; ++j) p[j] = 0.0f;
        const uint64_t i1 = ((uint64_t)c + 1) * R < n ? ((uint64_t)c + 1) * R : n;
        for (uint64_t i = (uint64_t)c * R; i < i1; ++i)
            for (uint64_t j = 0; j < n; ++j) p[j] = fmaf(fmaf(u2[i], v2[j], fmaf(u1[i], v1[j], A[i * lda + j])), y[i], p[j]);
    }
}

/* ---- comparison: chain_data.hpp:147-161 -------------------------------------------------------------------------- */
static double rel_l2_core(const double* num_den) {
    const double num = num_den[0], den = num_den[1];
    if (isnan(num)) return INFINITY;
    if (den == 0.0) return num == 0.0 ? 0.0 : INFINITY;
    return sqrt(num / den);
}

double or_relative_l2_f32(const float* got, const double* want, uint64_t n) {
    double num = 0.0, den = 0.0;
    for (uint64_t i = 0; i < n; ++i) {
        const double d = (double)got[i] - want[i];
        num += d * d;
        den += want[i] * want[i];
    }
    const double nd[2] = {num, den};
    return rel_l2_core(nd);
}

double or_relative_l2_f64(const double* got, const double* want, uint64_t n) {
    double num = 0.0, den = 0.0;
    for (uint64_t i = 0; i < n; ++i) {
        const double d = got[i] - want[i];
        num += d * d;
        den += want[i] * want[i];
    }
    const double nd[2] = {num, den};
    return rel_l2_core(nd);
}

double or_relative_linf_f32(const float* got, const double* want, uint64_t n) {
    double md = 0.0, mw = 0.0;
    for (uint64_t i = 0; i < n; ++i) {
        const double d = fabs((double)got[i] - want[i]);
        if (isnan(d)) return INFINITY;
        if (d > md) md = d;
        if (fabs(want[i]) > mw) mw = fabs(want[i]);
    }
    if (mw == 0.0) return md == 0.0 ? 0.0 : INFINITY;
    return md / mw;
}

/* ---- storage formats --------------------------------------------------------------------------------------------- */
static uint32_t f2u(float f) {
    uint32_t u;
    memcpy(&u, &f, 4);
    return u;
}
static float u2f(uint32_t u) {
    float f;
    memcpy(&f, &u, 4);
    return f;
}

uint16_t or_f32_to_bf16(float f) {
    const uint32_t u = f2u(f);
    if ((u & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((u >> 16) | 0x0040u); /* NaN: quiet, keep the sign */
    const uint32_t rounded = u + 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(rounded >> 16);
}

float or_bf16_to_f32(uint16_t b) { return u2f((uint32_t)b << 16); }

/* OCP E4M3: bias 7, no infinities, 0x7f / 0xff are NaN, largest finite 448 = 1.75 * 2^8, smallest subnormal 2^-9. */
float or_e4m3_to_f32(uint8_t b) {
    const uint32_t s = b >> 7, e = (b >> 3) & 0xfu, m = b & 7u;
    float          v;
    if (e == 0xfu && m == 7u) return u2f(s ? 0xffc00000u : 0x7fc00000u);
    if (e == 0)
        v = ldexpf((float)m, -9);
    else
        v = ldexpf(1.0f + (float)m / 8.0f, (int)e - 7);
    return s ? -v : v;
}

uint8_t or_f32_to_e4m3(float f) {
    const uint32_t u    = f2u(f);
    const uint8_t  sign = (uint8_t)((u >> 31) << 7);
    if (isnan(f)) return (uint8_t)(sign | 0x7f);
    const float a = fabsf(f);
    if (a >= 448.0f) return (uint8_t)(sign | 0x7e); /* saturate */
    /* candidates: the value in units of the spacing of its binade, rounded to nearest even */
    int   ex;
    (void)frexpf(a, &ex); /* a = fr * 2^ex, fr in [0.5, 1) */
    int e_unb = ex - 1;   /* a in [2^e_unb, 2^(e_unb+1)) */
    if (a == 0.0f) return sign;
    if (e_unb < -6) e_unb = -6; /* subnormal range: spacing 2^-9 */
    const float spacing = ldexpf(1.0f, e_unb - 3);
    const float q       = a / spacing; /* exact: a power-of-two division */
    float       r       = nearbyintf(q); /* round to nearest even (default rounding mode) */
    /* r * spacing is the rounded magnitude; re-encode */
    const float mag = r * spacing;
    if (mag >= 448.0f) return (uint8_t)(sign | 0x7e);
    if (mag == 0.0f) return sign;
    int ex2;
    (void)frexpf(mag, &ex2);
    int e2 = ex2 - 1;
    if (e2 < -6) { /* subnormal */
        const uint32_t m = (uint32_t)(mag / ldexpf(1.0f, -9));
        return (uint8_t)(sign | m);
    }
    const uint32_t m = (uint32_t)((mag / ldexpf(1.0f, e2) - 1.0f) * 8.0f);
    return (uint8_t)(sign | (uint8_t)((e2 + 7) << 3) | (uint8_t)m);
}

void or_quantize_bf16(const float* in, uint16_t* out, uint64_t rows, uint64_t cols, uint64_t ld_in, uint64_t ld_out) {
#pragma omp parallel for schedule(static)
    for (long long i = 0; i < (long long)rows; ++i)
        for (uint64_t j = 0; j < cols; ++j) out[(uint64_t)i * ld_out + j] = or_f32_to_bf16(in[(uint64_t)i * ld_in + j]);
}

void or_quantize_e4m3_rows(const float* in, uint8_t* out, float* scale, uint64_t rows, uint64_t cols, uint64_t ld_in,
                           uint64_t ld_out) {
#pragma omp parallel for schedule(static)
    for (long long i = 0; i < (long long)rows; ++i) {
        const float* a    = in + (uint64_t)i * ld_in;
        float        amax = 0.0f;
        for (uint64_t j = 0; j < cols; ++j)
            if (fabsf(a[j]) > amax) amax = fabsf(a[j]);
        const float s = amax > 0.0f ? amax / 448.0f : 1.0f;
        scale[i]      = s;
        for (uint64_t j = 0; j < cols; ++j) out[(uint64_t)i * ld_out + j] = or_f32_to_e4m3(a[j] / s);
    }
}

/* ---- GEMVER, FP64: chain_data.hpp:93-117 ------------------------------------------------------------------------- */
static inline double stored(int store, const void* A, const float* row_scale, uint64_t i, uint64_t j, uint64_t lda) {
    switch (store) {
        case 1: return (double)or_bf16_to_f32(((const uint16_t*)A)[i * lda + j]);
        case 2: return (double)or_e4m3_to_f32(((const uint8_t*)A)[i * lda + j]) * (double)row_scale[i];
        default: return (double)((const float*)A)[i * lda + j];
    }
}

void or_gemver_f64(uint64_t n, int store, const void* A, uint64_t lda, const float* row_scale, const float* u1,
                   const float* v1, const float* u2, const float* v2, const float* y, const float* z, double alpha,
                   double beta, double* x, double* w) {
    /* B(i, j) = A(i, j) + (u1_i v1_j + u2_i v2_j), chain_data.hpp:99-102; x = z + beta * B^T y (i ascending per j);
     * w = alpha * B x (j ascending per i), chain_data.hpp:104-117. Columns are split over threads for x: each column's
     * sum keeps the reference's row order. */
    const long long JB = 256;
#pragma omp parallel for schedule(dynamic, 1)
    for (long long jb = 0; jb < (long long)n; jb += JB) {
        double         acc[256];
        const uint64_t je = (uint64_t)jb + JB < n ? (uint64_t)jb + JB : n;
        for (uint64_t j = (uint64_t)jb; j < je; ++j) acc[j - jb] = 0.0;
        for (uint64_t i = 0; i < n; ++i) {
            const double yi = y[i], a1 = u1[i], a2 = u2[i];
            for (uint64_t j = (uint64_t)jb; j < je; ++j)
                acc[j - jb] += (stored(store, A, row_scale, i, j, lda) + (a1 * v1[j] + a2 * v2[j])) * yi;
        }
        for (uint64_t j = (uint64_t)jb; j < je; ++j) x[j] = (double)z[j] + beta * acc[j - jb];
    }
#pragma omp parallel for schedule(static)
    for (long long i = 0; i < (long long)n; ++i) {
        const double a1 = u1[i], a2 = u2[i];
        double       sum = 0.0;
        for (uint64_t j = 0; j < n; ++j) sum += (stored(store, A, row_scale, (uint64_t)i, j, lda) + (a1 * v1[j] + a2 * v2[j])) * x[j];
        w[i] = alpha * sum;
    }
}

/* ---- the chain in FP64: examples/gemver-chain/common/chain_data.hpp:88-147 ------------------------------------------
 * B_t(i, j) = A0(i, j) + t (u1_i v1_j + u2_i v2_j) with t = 1 for B1 and t = 2 for B2 (G2 adds the same update to B1),
 * recomputed on the fly (chain_data.hpp:99-102). GEMVER (:104-117): acc_j = sum_i B(i, j) y_i (i ascending), x_j = z_j +
 * beta acc_j, w_i = alpha sum_j B(i, j) x_j (j ascending). BiCGK (:124-133): q_i = sum_j B1(i, j) x1_j, s_j = sum_i B1(i, j)
 * w1_i (i ascending). Threads split independent indices only, so every sum keeps the reference's order. */
static double chain_b(const float* A0, const float* u1, const float* v1, const float* u2, const float* v2, uint64_t n, uint64_t i,
                      uint64_t j, int t) {
    return (double)A0[i * n + j] + (double)t * ((double)u1[i] * (double)v1[j] + (double)u2[i] * (double)v2[j]);
}

static void chain_gemver(uint64_t n, int t, const float* A0, const float* u1, const float* v1, const float* u2, const float* v2,
                         const double* y, const double* z, double alpha, double beta, double* x, double* w) {
    const long long JB = 256;
#pragma omp parallel for schedule(dynamic, 1)
    for (long long jb = 0; jb < (long long)n; jb += JB) {
        double         acc[256];
        const uint64_t je = (uint64_t)jb + JB < n ? (uint64_t)jb + JB : n;
        for (uint64_t j = (uint64_t)jb; j < je; ++j) acc[j - jb] = 0.0;
        for (uint64_t i = 0; i < n; ++i)
            for (uint64_t j = (uint64_t)jb; j < je; ++j) acc[j - jb] += chain_b(A0, u1, v1, u2, v2, n, i, j, t) * y[i];
        for (uint64_t j = (uint64_t)jb; j < je; ++j) x[j] = z[j] + beta * acc[j - jb];
    }
#pragma omp parallel for schedule(static)
    for (long long i = 0; i < (long long)n; ++i) {
        double sum = 0.0;
        for (uint64_t j = 0; j < n; ++j) sum += chain_b(A0, u1, v1, u2, v2, n, (uint64_t)i, j, t) * x[j];
        w[i] = alpha * sum;
    }
}

void or_chain_f64(uint64_t n, const float* A0, const float* u1, const float* v1, const float* u2, const float* v2, const float* y,
                  const float* z, double alpha, double beta, double* x1, double* w1, double* q, double* s, double* x2, double* w2,
                  double* B2) {
    double* yz = (double*)malloc(2 * n * sizeof(double));
    for (uint64_t i = 0; i < n; ++i) {
        yz[i]     = y[i];
        yz[n + i] = z[i];
    }
    chain_gemver(n, 1, A0, u1, v1, u2, v2, yz, yz + n, alpha, beta, x1, w1);
#pragma omp parallel for schedule(static)
    for (long long i = 0; i < (long long)n; ++i) {
        double sum = 0.0;
        for (uint64_t j = 0; j < n; ++j) sum += chain_b(A0, u1, v1, u2, v2, n, (uint64_t)i, j, 1) * x1[j];
        q[i] = sum;
    }
    const long long JB = 256;
#pragma omp parallel for schedule(dynamic, 1)
    for (long long jb = 0; jb < (long long)n; jb += JB) {
        const uint64_t je = (uint64_t)jb + JB < n ? (uint64_t)jb + JB : n;
        for (uint64_t j = (uint64_t)jb; j < je; ++j) s[j] = 0.0;
        for (uint64_t i = 0; i < n; ++i)
            for (uint64_t j = (uint64_t)jb; j < je; ++j) s[j] += chain_b(A0, u1, v1, u2, v2, n, i, j, 1) * w1[i];
    }
    chain_gemver(n, 2, A0, u1, v1, u2, v2, q, s, alpha, beta, x2, w2);
    if (B2) {
#pragma omp parallel for schedule(static)
        for (long long i = 0; i < (long long)n; ++i)
            for (uint64_t j = 0; j < n; ++j) B2[(uint64_t)i * n + j] = chain_b(A0, u1, v1, u2, v2, n, (uint64_t)i, j, 2);
    }
    free(yz);
}

/* ---- BiCGK from one pass, f32, in this build's order (no reference order to follow: components.hpp:283-296 sums each
 * row's block with an implementation-defined group reduction) ----------------------------------------------------------
 * Ps[c][j] = fma chain of r_i B(i, j) over the strip's rows (ascending). Pq[i][b] for the 256-column block b: lane l
 * (0 .. 63) takes columns 256 b + 4 l .. + 3, d_l = fma(B3 p3, fma(B2 p2, fma(B1 p1, B0 p0))), and the block's value is
 * the total of a wave inclusive scan of d (dpp_scan_total64). */
static float dpp_scan_total64(const float* d) {
    /* Hillis-Steele inside rows of 16 lanes (shifts 1, 2, 4, 8; a lane with no source adds 0), then lane 15 of rows 0
     * and 2 added to rows 1 and 3, then lane 31 added to rows 2 and 3; lanes a step leaves out add 0. */
    float x[64], y[64];
    memcpy(x, d, sizeof x);
    for (int sh = 1; sh < 16; sh <<= 1) {
        for (int l = 0; l < 64; ++l) y[l] = x[l] + ((l % 16) >= sh ? x[l - sh] : 0.0f);
        memcpy(x, y, sizeof x);
    }
    for (int l = 0; l < 64; ++l) {
        const int row = l / 16;
        y[l]          = x[l] + ((row == 1 || row == 3) ? x[row * 16 - 1] : 0.0f);
    }
    memcpy(x, y, sizeof x);
    for (int l = 0; l < 64; ++l) y[l] = x[l] + (l >= 32 ? x[31] : 0.0f);
    return y[63];
}

void or_bicgk_onepass_f32(uint64_t n, uint64_t R, const float* B, const float* p, const float* r, float* Ps, float* Pq) {
    const uint64_t strips = n / R, blocks = n / 256;
#pragma omp parallel for schedule(static)
    for (long long c = 0; c < (long long)strips; ++c) {
        float* ps = Ps + (uint64_t)c * n;
        for (uint64_t j = 0; j < n; ++j) ps[j] = 0.0f;
        for (uint64_t k = 0; k < R; ++k) {
            const uint64_t i  = (uint64_t)c * R + k;
            const float*   b  = B + i * n;
            const float    ri = r[i];
            for (uint64_t j = 0; j < n; ++j) ps[j] = fmaf(ri, b[j], ps[j]);
            for (uint64_t blk = 0; blk < blocks; ++blk) {
                float d[64];
                for (int l = 0; l < 64; ++l) {
                    const uint64_t j = blk * 256 + 4 * (uint64_t)l;
                    d[l]             = fmaf(b[j + 3], p[j + 3], fmaf(b[j + 2], p[j + 2], fmaf(b[j + 1], p[j + 1], b[j] * p[j])));
                }
                Pq[i * blocks + blk] = dpp_scan_total64(d);
            }
        }
    }
}

/* components.hpp:326-340: q[i] = sum_b Pq[i][b], b ascending */
void or_row_reduce_f32(uint64_t n, uint64_t blocks, const float* Pq, float* q) {
#pragma omp parallel for schedule(static)
    for (long long i = 0; i < (long long)n; ++i) {
        float acc = 0.0f;
        for (uint64_t b = 0; b < blocks; ++b) acc += Pq[(uint64_t)i * blocks + b];
        q[i] = acc;
    }
}

/* ---- GEMVER, f32 restatement of the components ---------------------------------------------------------------------
 * The reference's bodies leave floating-point contraction to its compiler (components.hpp:93-94,150,200); this build
 * fixes it: b = fma(u2_i, v2_j, fma(u1_i, v1_j, a_ij)), acc = fma(b, y_i, acc), x = fma(scale, acc, z). */
static inline float rank2(float a, float u1i, float v1j, float u2i, float v2j) { return fmaf(u2i, v2j, fmaf(u1i, v1j, a)); }

/* components.hpp:85-97: strips of R rows, B[cR + k][j] */
void or_rank2_update_f32(uint64_t n, uint64_t R, const float* A, const float* u1, const float* v1, const float* u2,
                         const float* v2, float* B) {
    const uint64_t rows = (n / R) * R;
#pragma omp parallel for schedule(static)
    for (long long i = 0; i < (long long)rows; ++i)
        for (uint64_t j = 0; j < n; ++j) B[(uint64_t)i * n + j] = rank2(A[(uint64_t)i * n + j], u1[i], v1[j], u2[i], v2[j]);
}

/* components.hpp:142-155: P[c][j] = sum_{k<R} B[cR + k][j] * y[cR + k], k ascending */
void or_gemvt_partial_f32(uint64_t n, uint64_t R, const float* B, const float* y, float* P) {
    const uint64_t strips = n / R;
#pragma omp parallel for schedule(static)
    for (long long c = 0; c < (long long)strips; ++c) {
        float* p = P + (uint64_t)c * n;
        for (uint64_t j = 0; j < n; ++j) p[j] = 0.0f;
        for (uint64_t k = 0; k < R; ++k) {
            const uint64_t i  = (uint64_t)c * R + k;
            const float    yi = y[i];
            for (uint64_t j = 0; j < n; ++j) p[j] = fmaf(B[i * n + j], yi, p[j]);
        }
    }
}

/* components.hpp:192-204: x[j] = z[j] + scale * sum_c P[c][j], c ascending (plain adds) */
void or_col_reduce_f32(uint64_t n, uint64_t strips, const float* P, const float* z, float scale, float* x) {
#pragma omp parallel for schedule(static)
    for (long long j = 0; j < (long long)n; ++j) {
        float acc = 0.0f;
        for (uint64_t c = 0; c < strips; ++c) acc += P[c * n + (uint64_t)j];
        x[j] = z ? fmaf(scale, acc, z[j]) : scale * acc;
    }
}

/* The group reduction of components.hpp:247 has an implementation-defined order; this build's is a balanced pairwise
 * tree over the 256 work-items in local-id order (what a wave64 xor butterfly followed by a pairwise sum of the four
 * waves computes). */
static float tree_sum(const float* p, int m) {
    if (m == 1) return p[0];
    const int h = m / 2;
    return tree_sum(p, h) + tree_sum(p + h, h);
}

/* components.hpp:237-250: work-item l sums j = l, l + 256, ... (fma, ascending), then the group reduction */
void or_gemv_f32(uint64_t n, const float* B, const float* x, float scale, float* y) {
#pragma omp parallel for schedule(static)
    for (long long i = 0; i < (long long)n; ++i) {
        float lane[256];
        for (int l = 0; l < 256; ++l) {
            float acc = 0.0f;
            for (uint64_t j = (uint64_t)l; j < n; j += 256) acc = fmaf(B[(uint64_t)i * n + j], x[j], acc);
            lane[l] = acc;
        }
        y[i] = scale * tree_sum(lane, 256);
    }
}

/* The fused two-pass form on the CPU: pass 1 = Rank2Update + GemvTPartial with B kept in registers (same order, so the
 * strip partials and x are bitwise the components'); pass 2 = w_i = alpha * sum_j b_ij x_j with 16 interleaved
 * accumulators (j mod 16), combined pairwise. */
void or_gemver_twopass_f32(uint64_t n, uint64_t R, const float* A, const float* u1, const float* v1, const float* u2,
                           const float* v2, const float* y, const float* z, float alpha, float beta, float* P, float* x,
                           float* w) {
    const uint64_t strips = n / R;
#pragma omp parallel for schedule(static)
    for (long long c = 0; c < (long long)strips; ++c) {
        float* p = P + (uint64_t)c * n;
        for (uint64_t j = 0; j < n; ++j) p[j] = 0.0f;
        for (uint64_t k = 0; k < R; ++k) {
            const uint64_t i  = (uint64_t)c * R + k;
            const float    yi = y[i], a1 = u1[i], a2 = u2[i];
            const float*   a  = A + i * n;
            for (uint64_t j = 0; j < n; ++j) p[j] = fmaf(rank2(a[j], a1, v1[j], a2, v2[j]), yi, p[j]);
        }
    }
    or_col_reduce_f32(n, strips, P, z, beta, x);
#pragma omp parallel for schedule(static)
    for (long long i = 0; i < (long long)n; ++i) {
        float        acc[16] = {0};
        const float  a1 = u1[i], a2 = u2[i];
        const float* a  = A + (uint64_t)i * n;
        uint64_t     j  = 0;
        for (; j + 16 <= n; j += 16)
            for (int l = 0; l < 16; ++l) acc[l] = fmaf(rank2(a[j + l], a1, v1[j + l], a2, v2[j + l]), x[j + l], acc[l]);
        for (; j < n; ++j) acc[j % 16] = fmaf(rank2(a[j], a1, v1[j], a2, v2[j]), x[j], acc[j % 16]);
        w[i] = alpha * tree_sum(acc, 16);
    }
}

/* ---- scans: device_scan.hpp:43-48 (n + 1 outputs, the last is the total, running sums wrap) --------------------- */
void or_exclusive_scan_u32(const uint32_t* in, uint32_t* out, uint64_t n) {
    uint32_t s = 0;
    for (uint64_t i = 0; i < n; ++i) {
        out[i] = s;
        s += in[i];
    }
    out[n] = s;
}

void or_exclusive_scan_u64(const uint64_t* in, uint64_t* out, uint64_t n) {
    uint64_t s = 0;
    for (uint64_t i = 0; i < n; ++i) {
        out[i] = s;
        s += in[i];
    }
    out[n] = s;
}

/* ---- ragged columns: shape_hook.hpp:180-201 (flag bits, narrowing), :269-297 (counts), :303-330 (offsets) ------- */
static int load_src(int dt, const void* src, uint64_t i, int64_t* sv, uint64_t* uv) {
    switch (dt) {
        case 0: *sv = ((const int32_t*)src)[i]; return 1;
        case 1: *uv = ((const uint32_t*)src)[i]; return 0;
        case 2: *sv = ((const int64_t*)src)[i]; return 1;
        default: *uv = ((const uint64_t*)src)[i]; return 0;
    }
}

uint32_t or_build_column(int src_dtype, const void* src, uint64_t count_src, uint64_t k, int offsets_mode,
                         uint32_t* extents, uint32_t* offs, uint32_t* total) {
    uint32_t flag = 0;
    if (offsets_mode) {
        if (count_src == 0) return 0xffffffffu; /* the host refuses this before any kernel (shape_hook.hpp:245-246) */
        const uint64_t n = count_src - 1;
        for (uint64_t i = 0; i < count_src; ++i) {
            int64_t  sv = 0;
            uint64_t uv = 0;
            uint32_t x  = 0;
            int      bad = 0;
            if (load_src(src_dtype, src, i, &sv, &uv)) {
                if (sv < 0)
                    bad = 1;
                else if ((uint64_t)sv > 0xffffffffull)
                    flag |= 2u;
                else
                    x = (uint32_t)sv;
            } else {
                if (uv > 0xffffffffull)
                    flag |= 2u;
                else
                    x = (uint32_t)uv;
            }
            if (i == 0 && x != 0) bad = 1;
            offs[i] = x;
            if (bad) flag |= 1u;
        }
        for (uint64_t i = 0; i < n; ++i) {
            if (offs[i + 1] < offs[i]) {
                extents[i] = 0;
                flag |= 1u;
            } else {
                extents[i] = offs[i + 1] - offs[i];
            }
        }
        *total = offs[n];
        return flag;
    }
    const uint64_t n = count_src;
    for (uint64_t i = 0; i < n; ++i) {
        int64_t  sv = 0;
        uint64_t uv = 0;
        uint32_t x  = 0;
        if (load_src(src_dtype, src, i * k, &sv, &uv)) {
            if (sv < 0)
                flag |= 1u;
            else if ((uint64_t)sv > 0xffffffffull)
                flag |= 2u;
            else
                x = (uint32_t)sv;
        } else {
            if (uv > 0xffffffffull)
                flag |= 2u;
            else
                x = (uint32_t)uv;
        }
        extents[i] = x;
    }
    or_exclusive_scan_u32(extents, offs, n);
    for (uint64_t i = 0; i < n; ++i)
        if (offs[i + 1] < offs[i]) flag |= 2u;
    *total = offs[n];
    return flag;
}

/* ---- batched sizes ---------------------------------------------------------------------------------------------- */
void or_ragged_sizes(uint32_t seed, uint32_t count, uint32_t lo, uint32_t hi, uint32_t* sizes) {
    const double ratio = (double)hi / (double)lo;
    for (uint32_t k = 0; k < count; ++k) {
        const double u = ((double)or_hash_unit(seed, k) + 1.0) * 0.5; /* [0, 1) */
        double       v = (double)lo * pow(ratio, u);
        long long    r = llround(v);
        if (r < (long long)lo) r = lo;
        if (r > (long long)hi) r = hi;
        sizes[k] = (uint32_t)r;
    }
}

/* ---- incompressible NS ------------------------------------------------------------------------------------------
 * Fields: u, v, p each the sum of 8 modes a_m / |k_m|^2 cos(2 pi (kx x + ky y) + phi_m), x = j / nx, y = i / ny,
 * integer k in [-4, 4]^2 \ {0}, a_m ~ N(0, 1) (Box-Muller), phi_m ~ U[0, 2 pi), every draw from the hash (seed,
 * field * 64 + m * 8 + t). Evaluated as a_m (cx_j cy_i - sx_j sy_i) with per-column and per-row tables. */
#define OR_PI 3.14159265358979323846
static void ns_modes(uint32_t seed, int f, int m, int* kx, int* ky, double* amp, double* phi) {
    const uint64_t b = (uint64_t)f * 64u + (uint64_t)m * 8u;
    *kx              = (int)floorf((or_hash_unit(seed, b + 0) + 1.0f) * 4.5f) - 4;
    *ky              = (int)floorf((or_hash_unit(seed, b + 1) + 1.0f) * 4.5f) - 4;
    if (*kx == 0 && *ky == 0) *kx = 1;
    const double u1 = 1.0 - ((double)or_hash_unit(seed, b + 2) + 1.0) * 0.5; /* (0, 1] */
    const double u2 = ((double)or_hash_unit(seed, b + 3) + 1.0) * 0.5;
    const double g  = sqrt(-2.0 * log(u1)) * cos(2.0 * OR_PI * u2);
    *amp            = g / (double)((*kx) * (*kx) + (*ky) * (*ky));
    *phi            = 2.0 * OR_PI * (((double)or_hash_unit(seed, b + 4) + 1.0) * 0.5);
}

void or_ns_fields(uint32_t seed, int nx, int ny, float* u, float* v, float* p) {
    float* outs[3] = {u, v, p};
    for (int f = 0; f < 3; ++f) {
        double* cx = (double*)malloc(sizeof(double) * 8 * (size_t)nx);
        double* sx = (double*)malloc(sizeof(double) * 8 * (size_t)nx);
        double* cy = (double*)malloc(sizeof(double) * 8 * (size_t)ny);
        double* sy = (double*)malloc(sizeof(double) * 8 * (size_t)ny);
        double  amp[8];
        for (int m = 0; m < 8; ++m) {
            int    kx, ky;
            double phi;
            ns_modes(seed, f, m, &kx, &ky, &amp[m], &phi);
            for (int j = 0; j < nx; ++j) {
                const double t = 2.0 * OR_PI * (double)kx * (double)j / (double)nx;
                cx[m * nx + j] = cos(t);
                sx[m * nx + j] = sin(t);
            }
            for (int i = 0; i < ny; ++i) {
                const double t = 2.0 * OR_PI * (double)ky * (double)i / (double)ny + phi;
                cy[m * ny + i] = cos(t);
                sy[m * ny + i] = sin(t);
            }
        }
        float* o = outs[f];
#pragma omp parallel for schedule(static)
        for (int i = 0; i < ny; ++i)
            for (int j = 0; j < nx; ++j) {
                double s = 0.0;
                for (int m = 0; m < 8; ++m) s += amp[m] * (cx[m * nx + j] * cy[m * ny + i] - sx[m * nx + j] * sy[m * ny + i]);
                o[(size_t)i * nx + j] = (float)s;
            }
        free(cx);
        free(sx);
        free(cy);
        free(sy);
    }
}

static inline double fld(int store, const void* a, size_t k) {
    return store == 1 ? (double)or_bf16_to_f32(((const uint16_t*)a)[k]) : (double)((const float*)a)[k];
}

/* One cell of the residual in FP64: central differences, periodic wrap. (The reference's host residual pads the
 * domain and sweeps rows, examples/ns-residual/reference.hpp:30-130; the incompressible physics is this build's.) */
static inline void ns_cell(int nx, int ny, double inv2h, double invh2, double nu, int store, const void* u,
                           const void* v, const void* p, int i, int j, double r[3]) {
    const int    jm = j == 0 ? nx - 1 : j - 1, jp = j == nx - 1 ? 0 : j + 1;
    const int    im = i == 0 ? ny - 1 : i - 1, ip = i == ny - 1 ? 0 : i + 1;
    const size_t c = (size_t)i * nx + j, w = (size_t)i * nx + jm, e = (size_t)i * nx + jp;
    const size_t s = (size_t)im * nx + j, nn = (size_t)ip * nx + j;
    const double uc = fld(store, u, c), vc = fld(store, v, c);
    const double ue = fld(store, u, e), uw = fld(store, u, w), un = fld(store, u, nn), us = fld(store, u, s);
    const double ve = fld(store, v, e), vw = fld(store, v, w), vn = fld(store, v, nn), vs = fld(store, v, s);
    const double pe = fld(store, p, e), pw = fld(store, p, w), pn = fld(store, p, nn), ps = fld(store, p, s);
    const double ux = (ue - uw) * inv2h, uy = (un - us) * inv2h, vx = (ve - vw) * inv2h, vy = (vn - vs) * inv2h;
    const double px = (pe - pw) * inv2h, py = (pn - ps) * inv2h;
    const double lu = (ue + uw + un + us - 4.0 * uc) * invh2, lv = (ve + vw + vn + vs - 4.0 * vc) * invh2;
    r[0] = uc * ux + vc * uy + px - nu * lu;
    r[1] = uc * vx + vc * vy + py - nu * lv;
    r[2] = ux + vy;
}

void or_ns_residual_f64(int nx, int ny, double h, double nu, int store, const void* u, const void* v, const void* p,
                        double* Ru, double* Rv, double* Rc, double* sq) {
    const double inv2h = 1.0 / (2.0 * h), invh2 = 1.0 / (h * h);
    double       s0 = 0.0, s1 = 0.0, s2 = 0.0;
#pragma omp parallel for schedule(static) reduction(+ : s0, s1, s2)
    for (int i = 0; i < ny; ++i)
        for (int j = 0; j < nx; ++j) {
            double r[3];
            ns_cell(nx, ny, inv2h, invh2, nu, store, u, v, p, i, j, r);
            const size_t c = (size_t)i * nx + j;
            if (Ru) Ru[c] = r[0];
            if (Rv) Rv[c] = r[1];
            if (Rc) Rc[c] = r[2];
            s0 += r[0] * r[0];
            s1 += r[1] * r[1];
            s2 += r[2] * r[2];
        }
    if (sq) {
        sq[0] = s0;
        sq[1] = s1;
        sq[2] = s2;
    }
}

void or_ns_compare(int nx, int ny, double h, double nu, int store, const void* u, const void* v, const void* p,
                   const float* Ru, const float* Rv, const float* Rc, double* stats, double* sq) {
    const double inv2h = 1.0 / (2.0 * h), invh2 = 1.0 / (h * h);
    double       n0 = 0, n1 = 0, n2 = 0, d0 = 0, d1 = 0, d2 = 0;
    double       m0 = 0, m1 = 0, m2 = 0, r0 = 0, r1 = 0, r2 = 0;
#pragma omp parallel for schedule(static) reduction(+ : n0, n1, n2, d0, d1, d2) reduction(max : m0, m1, m2, r0, r1, r2)
    for (int i = 0; i < ny; ++i)
        for (int j = 0; j < nx; ++j) {
            double r[3];
            ns_cell(nx, ny, inv2h, invh2, nu, store, u, v, p, i, j, r);
            const size_t c  = (size_t)i * nx + j;
            const double e0 = (double)Ru[c] - r[0], e1 = (double)Rv[c] - r[1], e2 = (double)Rc[c] - r[2];
            n0 += e0 * e0;
            n1 += e1 * e1;
            n2 += e2 * e2;
            d0 += r[0] * r[0];
            d1 += r[1] * r[1];
            d2 += r[2] * r[2];
            if (fabs(e0) > m0) m0 = fabs(e0);
            if (fabs(e1) > m1) m1 = fabs(e1);
            if (fabs(e2) > m2) m2 = fabs(e2);
            if (fabs(r[0]) > r0) r0 = fabs(r[0]);
            if (fabs(r[1]) > r1) r1 = fabs(r[1]);
            if (fabs(r[2]) > r2) r2 = fabs(r[2]);
        }
    const double nd0[2] = {n0, d0}, nd1[2] = {n1, d1}, nd2[2] = {n2, d2};
    stats[0] = rel_l2_core(nd0);
    stats[1] = r0 > 0 ? m0 / r0 : (m0 == 0 ? 0 : INFINITY);
    stats[2] = rel_l2_core(nd1);
    stats[3] = r1 > 0 ? m1 / r1 : (m1 == 0 ? 0 : INFINITY);
    stats[4] = rel_l2_core(nd2);
    stats[5] = r2 > 0 ? m2 / r2 : (m2 == 0 ? 0 : INFINITY);
    if (sq) {
        sq[0] = d0;
        sq[1] = d1;
        sq[2] = d2;
    }
}

/* ---- compressible NS residual, FP64: examples/ns-residual/physics.hpp (numerics), reference.hpp:29-128 (assembly) --
 * Conserved U = (rho, rho u, rho v, E), ideal gas gamma 1.4, mu 0.01, Pr 0.72, dx = dy = 1, arrays (v, y, x).
 * R = -div F_conv + div F_visc per cell; the x-face i of row j is the west face of cell (j, i), the y-face j of column i
 * its south face. x is periodic; y is periodic or has no-slip adiabatic walls (ghost rows mirror the interior with both
 * momenta negated). MUSCL on primitives (minmod or van Leer), HLLC with Davis speeds, viscous flux with the normal
 * derivative from the face's two cells and the tangential one from the mean of their central-difference gradients.
 * Diagnostics: min over cells of 0.5 / (|u| + |v| + 2c + 4 nu_max), max of max(|u|, |v|) + c, sum of R^2 per variable. */
#define NSC_GAMMA 1.4
#define NSC_MU 0.01
#define NSC_PR 0.72
#define NSC_CFL 0.5

void or_nsc_prim_f64(const double* U, double* W) {
    const double r = U[0], u = U[1] / r, v = U[2] / r;
    W[0] = r;
    W[1] = u;
    W[2] = v;
    W[3] = (NSC_GAMMA - 1.0) * (U[3] - 0.5 * r * (u * u + v * v));
}

double or_nsc_slope_f64(double a, double b, int vanleer) {
    if (a * b <= 0.0) return 0.0;
    if (vanleer) return 2.0 * a * b / (a + b);
    return a > 0.0 ? (a < b ? a : b) : (a > b ? a : b);
}

/* The physical flux and the HLLC flux of a face with normal x (dir 0) or y (dir 1); w = (rho, u, v, p). */
static void nsc_euler_flux(const double* w, int dir, double* f) {
    const double un = dir == 0 ? w[1] : w[2], ut = dir == 0 ? w[2] : w[1];
    const double E = w[3] / (NSC_GAMMA - 1.0) + 0.5 * w[0] * (w[1] * w[1] + w[2] * w[2]);
    f[0] = w[0] * un;
    f[1] = w[0] * un * un + w[3];
    f[2] = w[0] * un * ut;
    f[3] = (E + w[3]) * un;
}

void or_nsc_hllc_f64(const double* wl, const double* wr, int dir, double* F) {
    const double g = NSC_GAMMA;
    const double ul = dir == 0 ? wl[1] : wl[2], vl = dir == 0 ? wl[2] : wl[1];
    const double ur = dir == 0 ? wr[1] : wr[2], vr = dir == 0 ? wr[2] : wr[1];
    const double cl = sqrt(g * wl[3] / wl[0]), cr = sqrt(g * wr[3] / wr[0]);
    const double sl = fmin(ul - cl, ur - cr), sr = fmax(ul + cl, ur + cr); /* Davis */
    const double ml = wl[0] * (sl - ul), mr = wr[0] * (sr - ur);          /* mass flux into each star region */
    const double sm = (wr[3] - wl[3] + ul * ml - ur * mr) / (ml - mr);
    double       f[4]; /* face frame: mass, normal momentum, tangential momentum, energy */
    if (sl >= 0.0 || sr <= 0.0) {
        const double* w = sl >= 0.0 ? wl : wr;
        double        fe[4];
        nsc_euler_flux(w, dir, fe);
        f[0] = fe[0];
        f[1] = fe[1];
        f[2] = fe[2];
        f[3] = fe[3];
    } else {
        const int     left = sm >= 0.0;
        const double* w    = left ? wl : wr;
        const double  s = left ? sl : sr, un = left ? ul : ur, ut = left ? vl : vr, m = left ? ml : mr;
        const double  E  = w[3] / (g - 1.0) + 0.5 * w[0] * (un * un + ut * ut);
        const double  k  = m / (s - sm); /* star density */
        const double  Es = k * (E / w[0] + (sm - un) * (sm + w[3] / m));
        double        fe[4];
        nsc_euler_flux(w, dir, fe);
        /* F* = F + s (U* - U); U* = (k, k sm, k ut, Es), U = (rho, rho un, rho ut, E) */
        f[0] = fe[0] + s * (k - w[0]);
        f[1] = fe[1] + s * (k * sm - w[0] * un);
        f[2] = fe[2] + s * (k * ut - w[0] * ut);
        f[3] = fe[3] + s * (Es - E);
    }
    F[0] = f[0];
    F[1] = dir == 0 ? f[1] : f[2];
    F[2] = dir == 0 ? f[2] : f[1];
    F[3] = f[3];
}

/* gl, gr: the cells' central-difference gradients (ux, uy, vx, vy, Tx, Ty); T = p / rho. */
void or_nsc_visc_f64(const double* wl, const double* wr, const double* gl, const double* gr, int dir, double* F) {
    const double Tl = wl[3] / wl[0], Tr = wr[3] / wr[0];
    double       d[6]; /* ux uy vx vy Tx Ty at the face */
    for (int q = 0; q < 6; ++q) d[q] = 0.5 * (gl[q] + gr[q]);
    const int n = dir; /* the normal derivatives come from the two cells */
    d[0 + n]     = wr[1] - wl[1];
    d[2 + n]     = wr[2] - wl[2];
    d[4 + n]     = Tr - Tl;
    const double mu = NSC_MU, kappa = NSC_MU * NSC_GAMMA / ((NSC_GAMMA - 1.0) * NSC_PR);
    const double dv  = d[0] + d[3];
    const double txx = mu * (2.0 * d[0] - 2.0 / 3.0 * dv), tyy = mu * (2.0 * d[3] - 2.0 / 3.0 * dv), txy = mu * (d[1] + d[2]);
    const double uf = 0.5 * (wl[1] + wr[1]), vf = 0.5 * (wl[2] + wr[2]);
    const double tn1 = dir == 0 ? txx : txy, tn2 = dir == 0 ? txy : tyy;
    F[0] = 0.0;
    F[1] = tn1;
    F[2] = tn2;
    F[3] = uf * tn1 + vf * tn2 + kappa * d[4 + n];
}

void or_nsc_cell_diag_f64(const double* W, double* dt, double* ws) {
    const double c = sqrt(NSC_GAMMA * W[3] / W[0]);
    const double ax = fabs(W[1]) + c, ay = fabs(W[2]) + c;
    const double nu = NSC_MU / W[0] * fmax(4.0 / 3.0, NSC_GAMMA / NSC_PR);
    *dt = NSC_CFL / (ax + ay + 4.0 * nu);
    *ws = fmax(ax, ay);
}

void or_nsc_init_f32(int nx, int ny, float amp, float* U) {
    const float  k  = (float)(2.0 * 3.14159265358979323846 / 32.0);
    const size_t nc = (size_t)nx * (size_t)ny;
    for (int j = 0; j < ny; ++j)
        for (int i = 0; i < nx; ++i) {
            const float  x = (float)i + 0.5f, y = (float)j + 0.5f;
            const float  sx = sinf(k * x), cx = cosf(k * x), sy = sinf(k * y), cy = cosf(k * y);
            const float  r = 1.0f + 0.1f * amp * sx * sy, u = 0.3f + 0.05f * amp * cx * sy, v = 0.2f + 0.05f * amp * sx * cy;
            const float  p = 1.0f + 0.05f * amp * cx * cy;
            const size_t c = (size_t)j * (size_t)nx + (size_t)i;
            U[c]          = r;
            U[nc + c]     = r * u;
            U[2 * nc + c] = r * v;
            U[3 * nc + c] = p / (1.4f - 1.0f) + 0.5f * r * (u * u + v * v);
        }
}

/* The padded primitives (rho, u, v, p) of (ny + 4) x (nx + 4) cells, index ((b * PX) + a) * 4 + var. */
static double* nsc_padded(int nx, int ny, int walls, const float* U) {
    const long    PX = nx + 4, PY = ny + 4;
    const size_t  nc = (size_t)nx * (size_t)ny;
    double* const W  = (double*)malloc(sizeof(double) * 4 * (size_t)PX * (size_t)PY);
    for (long b = 0; b < PY; ++b)
        for (long a = 0; a < PX; ++a) {
            long   j = b - 2, i = ((a - 2) % nx + nx) % nx;
            double flip = 1.0;
            if (j < 0 || j >= ny) {
                if (walls) {
                    j    = j < 0 ? -1 - j : 2L * ny - 1 - j;
                    flip = -1.0;
                } else {
                    j = (j % ny + ny) % ny;
                }
            }
            const size_t c = (size_t)j * (size_t)nx + (size_t)i;
            double       u[4] = {U[c], flip * U[nc + c], flip * U[2 * nc + c], U[3 * nc + c]};
            or_nsc_prim_f64(u, &W[((size_t)b * (size_t)PX + (size_t)a) * 4]);
        }
    return W;
}

/* The convective and viscous fluxes through the face between padded cells (bl, al) and (br, ar). */
static void nsc_face(const double* W, long PX, long bl, long al, long br, long ar, int dir, int vanleer, double* Fc, double* Fv) {
#define NSC_W(b, a) (&W[((size_t)(b) * (size_t)PX + (size_t)(a)) * 4])
    const long db = dir == 1, da = dir == 0;
    const double *pl = NSC_W(bl, al), *pr = NSC_W(br, ar);
    const double *pll = NSC_W(bl - db, al - da), *prr = NSC_W(br + db, ar + da);
    double        wl[4], wr[4], gl[6], gr[6];
    for (int q = 0; q < 4; ++q) {
        wl[q] = pl[q] + 0.5 * or_nsc_slope_f64(pl[q] - pll[q], pr[q] - pl[q], vanleer);
        wr[q] = pr[q] - 0.5 * or_nsc_slope_f64(pr[q] - pl[q], prr[q] - pr[q], vanleer);
    }
    or_nsc_hllc_f64(wl, wr, dir, Fc);
    for (int s = 0; s < 2; ++s) {
        const long    b = s ? br : bl, a = s ? ar : al;
        const double *xm = NSC_W(b, a - 1), *xp = NSC_W(b, a + 1), *ym = NSC_W(b - 1, a), *yp = NSC_W(b + 1, a);
        double* const g = s ? gr : gl;
        g[0] = 0.5 * (xp[1] - xm[1]);
        g[1] = 0.5 * (yp[1] - ym[1]);
        g[2] = 0.5 * (xp[2] - xm[2]);
        g[3] = 0.5 * (yp[2] - ym[2]);
        g[4] = 0.5 * (xp[3] / xp[0] - xm[3] / xm[0]);
        g[5] = 0.5 * (yp[3] / yp[0] - ym[3] / ym[0]);
    }
    or_nsc_visc_f64(pl, pr, gl, gr, dir, Fv);
#undef NSC_W
}

void or_nsc_residual_f64(int nx, int ny, int walls, int vanleer, const float* U, double* R, double* diag) {
    const long    PX = nx + 4;
    double* const W  = nsc_padded(nx, ny, walls, U);
    const size_t  nc = (size_t)nx * (size_t)ny;
    double        dtmin = 1e300, wsmax = -1e300, sq[4] = {0, 0, 0, 0};
#pragma omp parallel for schedule(static) reduction(min : dtmin) reduction(max : wsmax) reduction(+ : sq[:4])
    for (long j = 0; j < ny; ++j) {
        const long b = j + 2;
        for (long i = 0; i < nx; ++i) {
            const long a = i + 2;
            double     cw[4], vw[4], ce[4], ve[4], cs[4], vs[4], cn[4], vn[4];
            nsc_face(W, PX, b, a - 1, b, a, 0, vanleer, cw, vw); /* west */
            nsc_face(W, PX, b, a, b, a + 1, 0, vanleer, ce, ve); /* east */
            nsc_face(W, PX, b - 1, a, b, a, 1, vanleer, cs, vs); /* south */
            nsc_face(W, PX, b, a, b + 1, a, 1, vanleer, cn, vn); /* north */
            for (int q = 0; q < 4; ++q) {
                const double divc = (ce[q] - cw[q]) + (cn[q] - cs[q]);
                const double divv = (ve[q] - vw[q]) + (vn[q] - vs[q]);
                const double r    = -divc + divv;
                R[(size_t)q * nc + (size_t)j * (size_t)nx + (size_t)i] = r;
                sq[q] += r * r;
            }
            double dt, ws;
            or_nsc_cell_diag_f64(&W[((size_t)b * (size_t)PX + (size_t)a) * 4], &dt, &ws);
            dtmin = fmin(dtmin, dt);
            wsmax = fmax(wsmax, ws);
        }
    }
    free(W);
    diag[0] = dtmin;
    diag[1] = wsmax;
    for (int q = 0; q < 4; ++q) diag[2 + q] = sq[q];
}
