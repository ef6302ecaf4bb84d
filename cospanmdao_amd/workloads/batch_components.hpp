#pragma once

// =============================================================================
// cospanmdao_amd/workloads/batch_components.hpp — batched ragged GEMVER: K independent instances in one persistent launch
// =============================================================================
//
// Instance k has size N_k (any N >= 1): A_k (N_k x N_k f32, leading dimension ld_k = N_k rounded up to 32, so every row
// starts on a 128-byte line), vectors u1 v1 u2 v2 y z of N_k, outputs x_k = beta Â_k^T y_k + z_k and w_k = alpha Â_k x_k. Every
// instance's arrays sit back to back in one allocation per array; where each starts comes from a DEVICE exclusive scan:
//
//   BatchPlan      (index range over K) per instance nine counts, stored column-major [9][K]:
//                    0 V  = N_k                                   vector elements          (x, w, u1 .. z offsets)
//                    1 A4 = N_k * ld_k / 4                        16-byte units of A_k     (A offsets)
//                    2 P  = strips_k * N_k                        strip partials           (P offsets)
//                    3 U1 = strips_k * ceil(N_k / 4096)           pass-1 tiles (R rows x 4096 columns)
//                    4 UR = ceil(N_k / 1024)                      column-reduction blocks
//                    5 U2 = ceil(N_k / RT)                        pass-2 row groups
//                    6, 7, 8  U1, UR, U2 again, at the instance's place in the work order (rank_k) instead of at k
//                  (strips_k = ceil(N_k / R)). The work order runs the largest instances first (the host ranks them by
//                  size), so that every phase ends on small units: the data stay in instance order, only the units'
//                  places in each phase's queue follow the ranks.
//   ColumnScan     (custom op) ONE u32 exclusive scan of the 9K counts (cs_hip_build_column: decoupled look-back, wrap
//                  flag): column c's offsets are scan[cK + k] - scan[cK]; column 0 is the int32 scan of the sizes.
// then one of two forms, both reading their units from
//   UnitMap        (K work-groups) writes, for every unit of every phase, what the unit needs to start (instance size,
//                  unit index within the instance, the instance's vector / A / P offsets) at the unit's scan position:
//                  a unit is found with one load instead of a search over the scan.
//   phased (default)
//   RaggedPass1 / RaggedReduce / RaggedPass2  one plain launch per phase, work-group u runs that phase's unit u: no
//                  work queue and no waits between work-groups; the launch boundaries order the phases.
//   persistent
//   RaggedGemver   ONE launch: a grid of G work-groups takes units by ticket until none is left, first every pass-1 unit of
//                  every instance, then every column reduction, then every pass-2 row group, each in the work order (the
//                  queue protocol below). A unit waits for its own instance's units of the phase before (a per-instance
//                  ready count), so the phases overlap; a pass-1 unit or reduction publishes its stores with an L2
//                  write-back shared by its XCD's work-groups and counts itself, and the units that wait read P and x with
//                  agent-scope loads: no acquire. Waits are only ever on units with earlier tickets, held by work-groups
//                  that are running and never wait on later ones, so the queue drains without co-residency. The last
//                  work-group to leave zeroes the counters (graph replays need no memset).
// The arithmetic of a unit is the two-pass form of gemver_components.hpp (AhatTPartial / ColReduce / AhatGemv): x_k equals
// the reference components' f32 x for the same R bit for bit, in both forms. R, RT (4, 8: the work-group shares a row
// tile; 16, 32: each wave owns RT / 4 rows) and G are run-time parameters (the autotuner's knobs; G only for the
// persistent form).

#include <cstdint>
#include <memory>
#include <stdexcept>
#include <string>

#include "naga/cospan/cospan.hpp"

namespace batch {

namespace cs = naga::cospan;
using cs::DType;

inline constexpr int kWg    = 256;
inline constexpr int kCols  = kWg * 4;    // columns of one column-reduction block / one pass-2 step (4 per work-item)
inline constexpr int kQ1    = 4;          // chunks of 4 columns per work-item in a pass-1 tile
inline constexpr int kTile1 = kCols * kQ1;  // columns of one pass-1 tile: kQ1 interleaved chunks of kCols
inline constexpr int kMaxRT = 8;
inline constexpr int kMapWords = 8;       // u32 words per unit-map entry
inline constexpr int kColumns  = 9;       // count columns per instance (the plan)

__host__ __device__ inline std::uint32_t ceil_div(std::uint32_t a, std::uint32_t b) { return (a + b - 1) / b; }
/// Row pitch: N rounded up to 32 floats, one 128-byte line. With rows packed to 16 bytes, most rows share their first and
/// last lines with the rows beside them, which other waves read at other times: both passes of the 512-instance batch
/// fetched 1.09-1.12x their bytes (1.01x with whole lines per row, for 0.7% more memory) and ran 8-10% slower per byte.
__host__ __device__ inline std::uint32_t ld_of(std::uint32_t n) { return (n + 31u) & ~31u; }

/// Knobs shared between the model's owner (the autotuner) and the components that size their grids from them: read at
/// every describe. `units` = the three phases' work-unit counts for the current R and RT (what the device plan computes).
struct Knobs {
    int           groups   = 1024;
    std::uint64_t units[3] = {1, 1, 1};
};

/// The phases' unit counts over the instances: pass-1 tiles, column-reduction blocks, pass-2 row groups.
inline void count_units(const std::uint32_t* sizes, std::size_t K, std::uint32_t R, std::uint32_t RT, std::uint64_t (&u)[3]) {
    u[0] = u[1] = u[2] = 0;
    for (std::size_t k = 0; k < K; ++k) {
        const std::uint32_t n = sizes[k];
        u[0] += static_cast<std::uint64_t>(ceil_div(n, R)) * ceil_div(n, kTile1);
        u[1] += ceil_div(n, kCols);
        u[2] += ceil_div(n, RT);
    }
}

/// What a unit needs to start: its instance's size and offsets (vector elements, 16-byte units of A, P elements), its
/// index among the instance's units of its phase, and the instance (set from the unit map).
struct Unit {
    std::uint32_t n, t, vo, a4, poff;
    std::uint32_t k = 0;
};

/// The device scan's offset columns (column c of instance k = scan[cK + k] - scan[cK]).
struct Cols {
    const std::uint32_t* s;
    std::uint32_t        K;
    __device__ std::uint32_t off(int c, std::uint32_t k) const { return s[c * K + k] - s[c * K]; }
    __device__ std::uint32_t count(int c, std::uint32_t k) const { return s[c * K + k + 1] - s[c * K + k]; }
    __device__ std::uint32_t total(int c) const { return s[(c + 1) * K] - s[c * K]; }
    /// the instance whose range of column c holds unit u: the largest k with off(c, k) <= u
    __device__ std::uint32_t find(int c, std::uint32_t u) const {
        std::uint32_t lo = 0, hi = K - 1;
        while (lo < hi) {
            const std::uint32_t mid = (lo + hi + 1) / 2;
            if (off(c, mid) <= u)
                lo = mid;
            else
                hi = mid - 1;
        }
        return lo;
    }
    __device__ Unit unit(const std::uint32_t* sizes, int c, std::uint32_t k, std::uint32_t u) const {
        return Unit{sizes[k], u - off(c, k), off(0, k), off(1, k), off(2, k)};
    }
};

/// What the units read and write (raw device pointers; a phase leaves the arrays it does not touch null).
struct Arrays {
    std::uint32_t R    = 0;
    const float * A = nullptr, *u1 = nullptr, *v1 = nullptr, *u2 = nullptr, *v2 = nullptr, *y = nullptr, *z = nullptr;
    const float*  x_in = nullptr;  ///< x as pass 2 reads it
    const float*  P_in = nullptr;  ///< P as the column reduction reads it
    float *       P = nullptr, *x = nullptr, *w = nullptr;
    float         alpha = 0.0f, beta = 0.0f;
};

// ---- the units ------------------------------------------------------------------------------------------------------

namespace detail {
/// The row loop of a pass-1 tile of NQ chunks (NQ uniform over the work-group). Chunks 0 .. NQ-2 start inside the row for
/// every work-item; the last one only where `last` holds (within one wave the predicate differs only at the row's end).
template <int NQ>
__device__ inline void pass1_rows(const float* ap, std::uint32_t ld, std::uint32_t i0, std::uint32_t i1, bool last, const float* y,
                                  const float* u1, const float* u2, const float (&a1)[kQ1][4], const float (&b1)[kQ1][4],
                                  float (&acc)[kQ1][4]) {
    using F4 = float __attribute__((ext_vector_type(4)));
#pragma unroll 4
    for (std::uint32_t i = i0; i < i1; ++i, ap += ld) {
        F4 r[NQ];
#pragma unroll
        for (int q = 0; q < NQ - 1; ++q) r[q] = __builtin_nontemporal_load(reinterpret_cast<const F4*>(ap + q * kCols));
        r[NQ - 1] = last ? __builtin_nontemporal_load(reinterpret_cast<const F4*>(ap + (NQ - 1) * kCols)) : F4{0.0f, 0.0f, 0.0f, 0.0f};
        const float yi = y[i], p1 = u1[i], p2 = u2[i];
#pragma unroll
        for (int q = 0; q < NQ; ++q)
#pragma unroll
            for (int e = 0; e < 4; ++e) acc[q][e] = fmaf(fmaf(p2, b1[q][e], fmaf(p1, a1[q][e], r[q][e])), yi, acc[q][e]);
    }
}

/// The same value in every lane, held in a scalar register (the unit descriptor is one per work-group): pointers built
/// from it are uniform, so per-row values load through the scalar cache.
__device__ inline std::uint32_t uniform(std::uint32_t v) { return static_cast<std::uint32_t>(__builtin_amdgcn_readfirstlane(static_cast<int>(v))); }
}  // namespace detail

/// pass-1 tile t of an instance: strip c = t / blocks (R rows), column block b = t % blocks (kTile1 columns). Work-item
/// `tid` owns up to kQ1 chunks of 4 columns at b kTile1 + q 1024 + 4 tid (q < kQ1), so that each load instruction of a
/// wave reads one contiguous 1 KiB run; a chunk that starts inside the row is loaded whole (a row holds ld >= n columns) and
/// only its columns below n are kept. P[c][j] = the fma chain over the strip's rows, ascending.
__device__ inline void unit_pass1(const Arrays& a, const Unit& wu, std::uint32_t tid) {
    const Unit          w{detail::uniform(wu.n), detail::uniform(wu.t), detail::uniform(wu.vo), detail::uniform(wu.a4), detail::uniform(wu.poff)};
    const std::uint32_t n = w.n, ld = ld_of(n), blocks = ceil_div(n, kTile1), R = a.R;
    const std::uint32_t c = w.t / blocks, b = w.t % blocks;
    const std::uint32_t cols = n - b * kTile1 < static_cast<std::uint32_t>(kTile1) ? n - b * kTile1 : static_cast<std::uint32_t>(kTile1);
    const std::uint32_t nq   = ceil_div(cols, kCols);  // chunks of the tile (uniform)
    const std::uint32_t jb   = b * kTile1 + tid * 4;
    if (jb >= n) return;  // past the row's end in every chunk
    const std::uint32_t i0 = c * R, i1 = i0 + R < n ? i0 + R : n;
    const float*        A  = a.A + 4 * static_cast<std::size_t>(w.a4);
    const float *       u1 = a.u1 + w.vo, *v1 = a.v1 + w.vo, *u2 = a.u2 + w.vo, *v2 = a.v2 + w.vo, *y = a.y + w.vo;
    const bool          last = jb + (nq - 1) * kCols < n;  // the tile's last chunk starts inside the row
    float               a1[kQ1][4], b1[kQ1][4], acc[kQ1][4];
#pragma unroll
    for (std::uint32_t q = 0; q < kQ1; ++q) {
#pragma unroll
        for (std::uint32_t e = 0; e < 4; ++e) {
            const std::uint32_t j = jb + q * kCols + e;
            a1[q][e]              = j < n ? v1[j] : 0.0f;
            b1[q][e]              = j < n ? v2[j] : 0.0f;
            acc[q][e]             = 0.0f;
        }
    }
    const float* ap = A + static_cast<std::size_t>(i0) * ld + jb;
    static_assert(kQ1 == 2 || kQ1 == 4, "cospan batch: a pass-1 tile holds 2 or 4 chunks per work-item");
    if constexpr (kQ1 == 4) {
        switch (nq) {
            case 4: detail::pass1_rows<4>(ap, ld, i0, i1, last, y, u1, u2, a1, b1, acc); break;
            case 3: detail::pass1_rows<3>(ap, ld, i0, i1, last, y, u1, u2, a1, b1, acc); break;
            case 2: detail::pass1_rows<2>(ap, ld, i0, i1, last, y, u1, u2, a1, b1, acc); break;
            default: detail::pass1_rows<1>(ap, ld, i0, i1, last, y, u1, u2, a1, b1, acc); break;
        }
    } else {
        if (nq == 2)
            detail::pass1_rows<2>(ap, ld, i0, i1, last, y, u1, u2, a1, b1, acc);
        else
            detail::pass1_rows<1>(ap, ld, i0, i1, last, y, u1, u2, a1, b1, acc);
    }
    float* P = a.P + w.poff + static_cast<std::size_t>(c) * n;
#pragma unroll
    for (std::uint32_t q = 0; q < kQ1; ++q)
#pragma unroll
        for (std::uint32_t e = 0; e < 4; ++e)
            if (jb + q * kCols + e < n) P[jb + q * kCols + e] = acc[q][e];
}

/// column-reduction block b of an instance: x[j] = fma(beta, sum_c P[c][j], z[j]) for the columns of block b. A
/// work-item owns the block's columns tid + 256 q (q < 4) and loads 8 strips of all four at once (32 loads in flight);
/// each column's sum still runs over the strips in ascending order.
__device__ inline void unit_reduce(const Arrays& a, const Unit& wu, std::uint32_t tid) {
    constexpr int       kQ = kCols / kWg, kS = 8;
    const Unit          w{detail::uniform(wu.n), detail::uniform(wu.t), detail::uniform(wu.vo), detail::uniform(wu.a4), detail::uniform(wu.poff)};
    const std::uint32_t n = w.n, strips = ceil_div(n, a.R);
    const float*        P = a.P_in + w.poff;
    std::uint32_t       j[kQ];
    bool                ok[kQ];
    float               acc[kQ];
#pragma unroll
    for (int q = 0; q < kQ; ++q) {
        j[q]   = w.t * kCols + tid + q * kWg;
        ok[q]  = j[q] < n;
        j[q]   = ok[q] ? j[q] : 0u;
        acc[q] = 0.0f;
    }
    std::uint32_t c = 0;
    for (; c + kS <= strips; c += kS) {
        float v[kS][kQ];
#pragma unroll
        for (int s = 0; s < kS; ++s)
#pragma unroll
            for (int q = 0; q < kQ; ++q) v[s][q] = P[static_cast<std::size_t>(c + s) * n + j[q]];
#pragma unroll
        for (int s = 0; s < kS; ++s)
#pragma unroll
            for (int q = 0; q < kQ; ++q) acc[q] += v[s][q];
    }
    for (; c < strips; ++c)
#pragma unroll
        for (int q = 0; q < kQ; ++q) acc[q] += P[static_cast<std::size_t>(c) * n + j[q]];
#pragma unroll
    for (int q = 0; q < kQ; ++q)
        if (ok[q]) a.x[w.vo + j[q]] = fmaf(a.beta, acc[q], a.z[w.vo + j[q]]);
}

/// pass-2 row group g = t of an instance: rows [g RT, g RT + RT). Ends with a barrier (its LDS is reused by the next
/// unit).
template <int RT>
__device__ inline void unit_pass2(const Arrays& a, const Unit& w, std::uint32_t tid) {
    const std::uint32_t n = w.n, ld = ld_of(n);
    const std::uint32_t r0 = w.t * RT, nr = n - r0 < static_cast<std::uint32_t>(RT) ? n - r0 : static_cast<std::uint32_t>(RT);
    const float*        A  = a.A + 4 * static_cast<std::size_t>(w.a4);
    const float *       u1 = a.u1 + w.vo, *v1 = a.v1 + w.vo, *u2 = a.u2 + w.vo, *v2 = a.v2 + w.vo, *x = a.x_in + w.vo;
    float               acc[RT], p1[RT], p2[RT];
#pragma unroll
    for (int r = 0; r < RT; ++r) {
        const std::uint32_t i = r0 + (static_cast<std::uint32_t>(r) < nr ? static_cast<std::uint32_t>(r) : 0u);
        p1[r]                 = u1[i];
        p2[r]                 = u2[i];
        acc[r]                = 0.0f;
    }
    using F4         = float __attribute__((ext_vector_type(4)));
    std::uint32_t j0 = tid * 4;
    for (; j0 + 4 <= n; j0 += kCols) {
        const F4 av = F4{v1[j0], v1[j0 + 1], v1[j0 + 2], v1[j0 + 3]};
        const F4 bv = F4{v2[j0], v2[j0 + 1], v2[j0 + 2], v2[j0 + 3]};
        const F4 xv = F4{x[j0], x[j0 + 1], x[j0 + 2], x[j0 + 3]};
        F4       t[RT];
#pragma unroll
        for (int r = 0; r < RT; ++r) {
            const std::uint32_t rr = static_cast<std::uint32_t>(r) < nr ? static_cast<std::uint32_t>(r) : 0u;
            t[r]                   = __builtin_nontemporal_load(reinterpret_cast<const F4*>(A + static_cast<std::size_t>(r0 + rr) * ld + j0));
        }
#pragma unroll
        for (int r = 0; r < RT; ++r)
#pragma unroll
            for (int e = 0; e < 4; ++e) acc[r] = fmaf(fmaf(p2[r], bv[e], fmaf(p1[r], av[e], t[r][e])), xv[e], acc[r]);
    }
    if (j0 < n)
        for (int r = 0; r < RT; ++r) {
            const std::uint32_t rr = static_cast<std::uint32_t>(r) < nr ? static_cast<std::uint32_t>(r) : 0u;
            for (std::uint32_t j = j0; j < n; ++j)
                acc[r] = fmaf(fmaf(p2[r], v2[j], fmaf(p1[r], v1[j], A[static_cast<std::size_t>(r0 + rr) * ld + j])), x[j], acc[r]);
        }
    __shared__ float ws[kWg / 64][kMaxRT];
    const int        wv = static_cast<int>(tid) / 64;
#pragma unroll
    for (int r = 0; r < RT; ++r) {
        const float v = cs::wave::reduce(acc[r], cs::plus<float>{});
        if (cs::wave::lane_id() == 0) ws[wv][r] = v;
    }
    __syncthreads();
    if (tid < nr) a.w[w.vo + r0 + tid] = a.alpha * ((ws[0][tid] + ws[1][tid]) + (ws[2][tid] + ws[3][tid]));
    __syncthreads();
}

/// pass-2 row group g = t of an instance, RT = 4 RTW rows: wave v of the work-group owns rows [g RT + v RTW, + RTW) and
/// sweeps every column of them (lane l takes the 4 columns at 4 l + 256 s), then reduces each row over its lanes alone: no
/// LDS and no barrier, and a unit carries 4x the rows of unit_pass2<RTW>.
template <int RTW>
__device__ inline void unit_pass2_waves(const Arrays& a, const Unit& wu, std::uint32_t tid) {
    const Unit w{detail::uniform(wu.n), detail::uniform(wu.t), detail::uniform(wu.vo), detail::uniform(wu.a4), detail::uniform(wu.poff)};
    constexpr std::uint32_t kLanes = 64, kStep = kLanes * 4;
    const std::uint32_t     n = w.n, ld = ld_of(n), lane = tid % kLanes, wv = tid / kLanes;
    const std::uint32_t     r0 = w.t * (4 * RTW) + wv * RTW;
    if (r0 >= n) return;
    const std::uint32_t nr = n - r0 < static_cast<std::uint32_t>(RTW) ? n - r0 : static_cast<std::uint32_t>(RTW);
    const float*        A  = a.A + 4 * static_cast<std::size_t>(w.a4);
    const float *       u1 = a.u1 + w.vo, *v1 = a.v1 + w.vo, *u2 = a.u2 + w.vo, *v2 = a.v2 + w.vo, *x = a.x_in + w.vo;
    float               acc[RTW], p1[RTW], p2[RTW];
#pragma unroll
    for (int r = 0; r < RTW; ++r) {
        const std::uint32_t i = r0 + (static_cast<std::uint32_t>(r) < nr ? static_cast<std::uint32_t>(r) : 0u);
        p1[r]                 = u1[i];
        p2[r]                 = u2[i];
        acc[r]                = 0.0f;
    }
    using F4         = float __attribute__((ext_vector_type(4)));
    std::uint32_t j0 = lane * 4;
#pragma unroll 2
    for (; j0 + 4 <= n; j0 += kStep) {
        const F4 av = F4{v1[j0], v1[j0 + 1], v1[j0 + 2], v1[j0 + 3]};
        const F4 bv = F4{v2[j0], v2[j0 + 1], v2[j0 + 2], v2[j0 + 3]};
        const F4 xv = F4{x[j0], x[j0 + 1], x[j0 + 2], x[j0 + 3]};
        F4       t[RTW];
#pragma unroll
        for (int r = 0; r < RTW; ++r) {
            const std::uint32_t rr = static_cast<std::uint32_t>(r) < nr ? static_cast<std::uint32_t>(r) : 0u;
            t[r]                   = __builtin_nontemporal_load(reinterpret_cast<const F4*>(A + static_cast<std::size_t>(r0 + rr) * ld + j0));
        }
#pragma unroll
        for (int r = 0; r < RTW; ++r)
#pragma unroll
            for (int e = 0; e < 4; ++e) acc[r] = fmaf(fmaf(p2[r], bv[e], fmaf(p1[r], av[e], t[r][e])), xv[e], acc[r]);
    }
    if (j0 < n)
        for (int r = 0; r < RTW; ++r) {
            const std::uint32_t rr = static_cast<std::uint32_t>(r) < nr ? static_cast<std::uint32_t>(r) : 0u;
            for (std::uint32_t j = j0; j < n; ++j)
                acc[r] = fmaf(fmaf(p2[r], v2[j], fmaf(p1[r], v1[j], A[static_cast<std::size_t>(r0 + rr) * ld + j])), x[j], acc[r]);
        }
#pragma unroll
    for (int r = 0; r < RTW; ++r) {
        const float v = cs::wave::reduce(acc[r], cs::plus<float>{});
        if (lane == 0 && static_cast<std::uint32_t>(r) < nr) a.w[w.vo + r0 + r] = a.alpha * v;
    }
}

/// A phase's unit u from the unit map (one 32-byte entry; `total` = the phase's unit count from the scan, `map_elems` =
/// the map's size). False when there is no such unit, or when the scan's count does not fit the map (then UnitMap wrote
/// nothing for the instances past its end and no entry is used).
__device__ inline bool map_unit(const std::uint32_t* map, std::size_t map_elems, std::uint32_t u, std::uint32_t total, Unit& w) {
    using U4    = std::uint32_t __attribute__((ext_vector_type(4)));
    const U4 e0 = reinterpret_cast<const U4*>(map)[2 * static_cast<std::size_t>(u)];
    const U4 e1 = reinterpret_cast<const U4*>(map)[2 * static_cast<std::size_t>(u) + 1];
    w           = Unit{e0[0], e0[1], e0[2], e0[3], e1[0], e1[1]};
    return u < total && static_cast<std::size_t>(total) * kMapWords <= map_elems;
}

/// The persistent form's work queue, as host-and-device functions (the kernel runs them, and so does its CPU emulation,
/// which interleaves simulated work-groups adversarially). Tickets run through three phases: [0, T1) pass-1 units,
/// [T1, T1 + TR) column reductions, [T1 + TR, T1 + TR + T2) pass-2 row groups, each phase in the work order (largest
/// instances first); any later ticket means "leave". A unit waits for its own instance only: a reduction of instance k
/// until all of k's pass-1 units are published, a pass-2 group until all of k's reductions are (state word ready(p, k)
/// counts instance k's published phase-p units), so the phases overlap: the reductions and pass-2 groups of the largest
/// instances run while the last pass-1 units of the smallest still do. A unit of phase 0 or 1 publishes its stores with a
/// write-back of its XCD's L2 shared with the XCD's other work-groups, then counts itself in its instance's ready word
/// and in its phase's published word. A wait ends with one acquire by the leader; a work-group that has seen a whole phase
/// published (published(q) = its unit count) acquires once more and then neither polls nor acquires for it again. Every
/// wait targets units with earlier tickets, taken by running work-groups that finish them without waiting on later ones:
/// no co-residency needed. The counters sit 256 bytes apart (ticket, finished, the two phases' published counts, the
/// write-back counts), the K + K ready counts after them.
namespace queue {
inline constexpr int kStride    = 64;  ///< u32 words between two counters
inline constexpr int kTicket    = 0, kFinished = kStride;
CS_HD constexpr int published(int q) { return (2 + q) * kStride; }
inline constexpr int kWriteBack = 4 * kStride;                                                ///< 2 counters per XCD
inline constexpr int kReady     = kWriteBack + 2 * static_cast<int>(cs::detail::kXcds) * kStride;  ///< then ready(p, k)
CS_HD constexpr std::size_t words(std::size_t K) { return kReady + 2 * K; }
CS_HD constexpr std::size_t ready(int p, std::size_t k, std::size_t K) { return kReady + static_cast<std::size_t>(p) * K + k; }

struct Totals {
    std::uint32_t T1 = 0, TR = 0, T2 = 0;
    CS_HD std::uint32_t all() const { return T1 + TR + T2; }
    /// units of phase p
    CS_HD std::uint32_t of(int p) const { return p == 0 ? T1 : p == 1 ? TR : T2; }
    /// the first ticket of phase p
    CS_HD std::uint32_t base(int p) const { return p == 0 ? 0u : p == 1 ? T1 : T1 + TR; }
};

/// 0, 1, 2 for a ticket of that phase, 3 past the last unit.
CS_HD inline int phase_of(std::uint32_t t, const Totals& n) {
    if (t < n.T1) return 0;
    t -= n.T1;
    if (t < n.TR) return 1;
    t -= n.TR;
    return t < n.T2 ? 2 : 3;
}

/// How many phase-(p - 1) units of an instance of size n a phase-p unit of it waits for: its pass-1 units (p = 1), its
/// column reductions (p = 2); none for pass 1.
CS_HD inline std::uint32_t awaits(int p, std::uint32_t n, std::uint32_t R) {
    return p == 1 ? ceil_div(n, R) * ceil_div(n, kTile1) : p == 2 ? ceil_div(n, kCols) : 0u;
}
}  // namespace queue

CS_COMPONENTS_BEGIN

/// Per-instance counts, [9][K] column-major. params = {R, RT, G}; rank = each instance's place in the work order.
class BatchPlan : public cs::Component<BatchPlan> {
public:
    struct Inputs {
        cs::Input<DType::u32, 1> sizes;
        cs::Input<DType::u32, 1> params;
        cs::Input<DType::u32, 1> rank;
    };
    struct Outputs {
        cs::Output<DType::u32, 1> counts;
    };
    __device__ static void compute_forward(cs::id<1> k, const Inputs& in, const Outputs& out) {
        const std::size_t   K  = in.sizes.size();
        const std::uint32_t n  = in.sizes[k];
        const std::uint32_t R  = in.params[std::size_t{0}], RT = in.params[std::size_t{1}];
        const std::uint32_t st = ceil_div(n, R);
        out.counts[0 * K + k] = n;
        out.counts[1 * K + k] = n * ld_of(n) / 4u;
        out.counts[2 * K + k] = st * n;
        const std::uint32_t u[3] = {st * ceil_div(n, kTile1), ceil_div(n, kCols), ceil_div(n, RT)};
        const std::size_t   r    = in.rank[k];
        for (int p = 0; p < 3; ++p) {
            out.counts[(3 + p) * K + k] = u[p];
            if (r < K) out.counts[(6 + p) * K + r] = u[p];  // the host passes a permutation of 0 .. K-1
        }
    }

    Inputs  inputs;
    Outputs outputs;
    void    declare() override {
        declare_input(inputs.sizes, "sizes", cs::unspecified_shape<1>());
        declare_input(inputs.params, "params", 3);
        declare_input(inputs.rank, "rank", cs::unspecified_shape<1>());
        declare_output(outputs.counts, "counts", cs::unspecified_shape<1>());
    }
    void prepare_forward() override {
        if (shape_of(inputs.rank)[0] != shape_of(inputs.sizes)[0]) throw std::invalid_argument("batch: rank must hold one place per instance");
        set_shape(outputs.counts, kColumns * shape_of(inputs.sizes)[0]);
    }
    std::unique_ptr<cs::LaunchDescription> define_forward() override { return create_launch_config(cs::range_desc(inputs.sizes.shape())); }
};

/// offs[0..M] = exclusive scan of counts (u32, wraps flagged), one launch of the library's decoupled look-back scan.
class ColumnScan : public cs::CustomComponent<ColumnScan> {
public:
    struct Inputs {
        cs::Input<DType::u32, 1> counts;
    };
    struct Outputs {
        cs::Output<DType::u32, 1> offs;
        cs::Output<DType::u32, 1> extents;
        cs::Output<DType::u32, 1> result;     ///< [total, flag] of the column build
        cs::Output<DType::u8, 1>  workspace;  ///< the scan's look-back status words (scratch: left zeroed)
    };
    static void compute_forward(cs::CustomOp& op, const Inputs& in, const Outputs& out) {
        cs::detail::hip_check(cs_hip_build_column(1, op.device_ptr(in.counts), in.counts.size(), 1, 0, op.device_ptr(out.extents),
                                                  op.device_ptr(out.offs), op.device_ptr(out.result), op.device_ptr(out.workspace), op.stream),
                              "batch offsets scan");
    }

    Inputs  inputs;
    Outputs outputs;
    void    declare() override {
        declare_input(inputs.counts, "counts", cs::unspecified_shape<1>());
        declare_output(outputs.offs, "offs", cs::unspecified_shape<1>());
        declare_output(outputs.extents, "extents", cs::unspecified_shape<1>());
        declare_output(outputs.result, "result", 2);
        declare_scratch(outputs.workspace, "workspace", cs::unspecified_shape<1>());
    }
    void prepare_forward() override {
        const std::size_t m = shape_of(inputs.counts)[0];
        set_shape(outputs.offs, m + 1);
        set_shape(outputs.extents, m);
        set_shape(outputs.workspace, cs_hip_scan_workspace_bytes(m));
    }
};

/// The persistent work-queue kernel: every unit of the three phases in ONE launch of G work-groups (the queue protocol
/// above). Units come from the unit map, as in the phased form.
class RaggedGemver : public cs::Component<RaggedGemver> {
public:
    static constexpr int max_work_group_size = kWg;
    struct Inputs {
        cs::Input<DType::u32, 1> sizes, params, offs, map1, mapR, map2;
        cs::Input<DType::f32, 1> A, u1, v1, u2, v2, y, z;
        cs::Input<DType::f32, 1> alpha, beta;
    };
    struct Outputs {
        cs::Output<DType::f32, 1> P, x, w;
        cs::Output<DType::u32, 1> state;  ///< ticket, finished, write-back counts, per-instance ready counts (queue::)
    };

    __device__ static void compute_forward(cs::GroupItem<1> it, const Inputs& in, const Outputs& out) {
        const Arrays        a = arrays(in, out);
        const std::uint32_t K = static_cast<std::uint32_t>(in.sizes.size());
        const Cols          cols{cs::raw::data(in.offs), K};
        const queue::Totals n{cols.total(3), cols.total(4), cols.total(5)};
        const std::uint32_t T   = n.all();
        const std::uint32_t RT  = in.params[std::size_t{1}];
        std::uint32_t*      st  = cs::raw::data(out.state);
        const std::uint32_t tid = static_cast<std::uint32_t>(it.get_local_id(0));
        const std::uint32_t* const maps[3]  = {cs::raw::data(in.map1), cs::raw::data(in.mapR), cs::raw::data(in.map2)};
        const std::size_t          sizes[3] = {in.map1.size(), in.mapR.size(), in.map2.size()};
        __shared__ std::uint32_t s_unit;

        // Leader-only blocks sit between barriers: the first fetch; per iteration the wait and the next fetch before the
        // unit, the publication and the hand-over of the fetched ticket after it (so no two of them can be merged across
        // the loop's back edge), and the exit is decided from LDS after a barrier (uniform). The next ticket is fetched
        // after the wait and before the unit, so its latency hides under the unit. A work-group takes each unit at most
        // once: T + 1 fetches bound the loop.
        std::uint32_t next = 0, seen = 0;  // leader only: seen bit q = phase q seen wholly published (and acquired after)
        if (tid == 0) s_unit = __hip_atomic_fetch_add(&st[queue::kTicket], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __syncthreads();
        std::uint32_t u       = detail::uniform(s_unit);
        bool          drained = false;
        for (std::uint32_t fetches = 1; fetches <= T + 1; ++fetches) {
            const int p = queue::phase_of(u, n);
            Unit      w{0, 0, 0, 0, 0};
            const bool has = p < 3 && map_unit(maps[p], sizes[p], u - n.base(p), n.of(p), w);
            if (tid == 0) {
                const int q = p - 1;
                if (has && p > 0 && ((seen >> q) & 1u) == 0u) {  // the instance's units of the phase before, all published
                    if (cs::detail::poll_load(&st[queue::published(q)]) >= n.of(q)) {
                        seen |= 1u << q;  // the whole phase: no poll and no acquire for it from now on
                    } else {
                        const std::uint32_t  need = queue::awaits(p, w.n, a.R);
                        std::uint32_t* const c    = &st[queue::ready(q, w.k, K)];
                        if (!cs::bounded::poll([&] { return cs::detail::poll_load(c) >= need; }, it.spin_limit()))
                            it.report_fault(CS_HIP_FAULT_UNIT_WAIT);
                    }
                    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
                }
                if (p != 3) next = __hip_atomic_fetch_add(&st[queue::kTicket], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            }
            if (p == 3) {
                drained = true;
                break;
            }
            __syncthreads();  // the leader's wait is behind every work-item
            if (has) run_unit(a, p, w, RT, tid);
            if (p < 2) asm volatile("s_waitcnt vmcnt(0)" ::: "memory");  // this wave's stores have reached the L2
            __syncthreads();
            if (tid == 0) {
                if (p < 2) {
                    if (!cs::detail::publish_stores(&st[queue::kWriteBack], it.spin_limit())) it.report_fault(CS_HIP_FAULT_UNIT_WAIT);
                    if (has) {
                        __hip_atomic_fetch_add(&st[queue::ready(p, w.k, K)], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                        __hip_atomic_fetch_add(&st[queue::published(p)], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    }
                }
                s_unit = next;
            }
            __syncthreads();
            u = detail::uniform(s_unit);
        }
        // leaving: the last work-group to leave zeroes the state (every unit is done and nobody reads it any more); its
        // count waits for this group's own counter updates. A loop ended by the fetch bound found a state that was not
        // zero at the launch, and reports.
        __syncthreads();
        __shared__ std::uint32_t s_last;
        if (tid == 0) {
            if (!drained) it.report_fault(CS_HIP_FAULT_STATE);
            asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
            const std::uint32_t f = __hip_atomic_fetch_add(&st[queue::kFinished], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            s_last                = f + 1 == static_cast<std::uint32_t>(it.get_group_range(0)) ? 1u : 0u;
        }
        __syncthreads();
        if (s_last) {
            const std::size_t words = queue::words(K);
            for (std::size_t i = tid; i < words; i += kWg) __hip_atomic_store(&st[i], 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
    }

    /// A unit of phase p.
    __device__ static void run_unit(const Arrays& a, int p, const Unit& w, std::uint32_t RT, std::uint32_t tid) {
        if (p == 0) {
            unit_pass1(a, w, tid);
        } else if (p == 1) {
            unit_reduce(a, w, tid);
        } else {
            switch (RT) {
                case 4: unit_pass2<4>(a, w, tid); break;
                case 8: unit_pass2<8>(a, w, tid); break;
                case 16: unit_pass2_waves<4>(a, w, tid); break;
                default: unit_pass2_waves<8>(a, w, tid); break;  // 32: the host accepts no other value
            }
        }
    }

    __device__ static Arrays arrays(const Inputs& in, const Outputs& out) {
        Arrays a;
        a.R     = in.params[std::size_t{0}];
        a.A     = cs::raw::data(in.A);
        a.u1    = cs::raw::data(in.u1);
        a.v1    = cs::raw::data(in.v1);
        a.u2    = cs::raw::data(in.u2);
        a.v2    = cs::raw::data(in.v2);
        a.y     = cs::raw::data(in.y);
        a.z     = cs::raw::data(in.z);
        a.P     = cs::raw::data(out.P);
        a.P_in  = cs::raw::data(out.P);
        a.x     = cs::raw::data(out.x);
        a.x_in  = cs::raw::data(out.x);
        a.w     = cs::raw::data(out.w);
        a.alpha = in.alpha[std::size_t{0}];
        a.beta  = in.beta[std::size_t{0}];
        return a;
    }

    Inputs  inputs;
    Outputs outputs;

    void declare() override {
        declare_input(inputs.sizes, "sizes", cs::unspecified_shape<1>());
        declare_input(inputs.params, "params", 3);
        declare_input(inputs.offs, "offs", cs::unspecified_shape<1>());
        declare_input(inputs.map1, "map1", cs::unspecified_shape<1>());
        declare_input(inputs.mapR, "mapR", cs::unspecified_shape<1>());
        declare_input(inputs.map2, "map2", cs::unspecified_shape<1>());
        for (auto [w, name] : {std::pair{&inputs.A, "A"}, {&inputs.u1, "u1"}, {&inputs.v1, "v1"}, {&inputs.u2, "u2"}, {&inputs.v2, "v2"},
                               {&inputs.y, "y"}, {&inputs.z, "z"}})
            declare_input(*w, name, cs::unspecified_shape<1>());
        declare_input(inputs.alpha, "alpha", 1);
        declare_input(inputs.beta, "beta", 1);
        declare_output(outputs.P, "P", cs::unspecified_shape<1>());
        declare_output(outputs.x, "x", cs::unspecified_shape<1>());
        declare_output(outputs.w, "w", cs::unspecified_shape<1>());
        declare_scratch(outputs.state, "state", cs::unspecified_shape<1>());
    }
    void prepare_forward() override {
        set_shape(outputs.x, shape_of(inputs.u1)[0]);
        set_shape(outputs.w, shape_of(inputs.u1)[0]);
        set_shape(outputs.state, queue::words(shape_of(inputs.sizes)[0]));
    }
    /// G persistent work-groups (the grid, the autotuner's third knob); every phase's unit count must fit its map.
    std::unique_ptr<cs::LaunchDescription> define_forward() override {
        const std::size_t maps[3] = {inputs.map1.shape()[0], inputs.mapR.shape()[0], inputs.map2.shape()[0]};
        const char*       who[3]  = {"pass 1", "column reduction", "pass 2"};
        for (int p = 0; p < 3; ++p)
            if (knobs_->units[p] * kMapWords > maps[p])
                throw std::logic_error(std::string{"batch: the "} + who[p] + " units (" + std::to_string(knobs_->units[p]) +
                                       ") do not fit their unit map (" + std::to_string(maps[p] / kMapWords) + " entries)");
        return create_launch_config(cs::range_desc(cs::range<1>{static_cast<std::size_t>(knobs_->groups) * kWg}), cs::wg_desc(cs::range<1>{kWg}));
    }
    explicit RaggedGemver(std::shared_ptr<Knobs> k) : knobs_{std::move(k)} {}

private:
    std::shared_ptr<Knobs> knobs_;
};

// ---- the phased form: the unit map, then one plain launch per phase, work-group u runs unit u ----------------------

/// For every unit of the three phases, its Unit at its place in the work order (map1: pass-1 tiles, mapR: column-reduction
/// blocks, map2: pass-2 row groups; kMapWords u32 per entry: n, t, vo, a4, poff, k, 0, 0): instance k's units of phase p
/// start at offset rank_k of column 6 + p. One work-group per instance.
class UnitMap : public cs::Component<UnitMap> {
public:
    static constexpr int max_work_group_size = kWg;
    struct Inputs {
        cs::Input<DType::u32, 1> sizes, offs, rank;
    };
    struct Outputs {
        cs::Output<DType::u32, 1> map1, mapR, map2;
    };
    __device__ static void compute_forward(cs::GroupItem<1> it, const Inputs& in, const Outputs& out) {
        using U4                 = std::uint32_t __attribute__((ext_vector_type(4)));
        const Cols          cols{cs::raw::data(in.offs), static_cast<std::uint32_t>(in.sizes.size())};
        const std::uint32_t k    = static_cast<std::uint32_t>(it.get_group(0));
        const std::uint32_t tid  = static_cast<std::uint32_t>(it.get_local_id(0));
        if (k >= cols.K) return;
        const std::uint32_t n = in.sizes[std::size_t{k}], vo = cols.off(0, k), a4 = cols.off(1, k), poff = cols.off(2, k);
        std::uint32_t* const maps[3] = {cs::raw::data(out.map1), cs::raw::data(out.mapR), cs::raw::data(out.map2)};
        const std::size_t    caps[3] = {out.map1.size() / kMapWords, out.mapR.size() / kMapWords, out.map2.size() / kMapWords};
        const std::uint32_t  r       = in.rank[std::size_t{k}];
        if (r >= cols.K) return;  // not a permutation: write nothing
        for (int p = 0; p < 3; ++p) {
            const std::uint32_t o = cols.off(6 + p, r), cnt = cols.count(3 + p, k);
            if (static_cast<std::size_t>(o) + cnt > caps[p]) continue;  // a scan that disagrees with the plan: write nothing
            U4* dst = reinterpret_cast<U4*>(maps[p]) + 2 * static_cast<std::size_t>(o);
            for (std::uint32_t i = tid; i < cnt; i += kWg) {
                dst[2 * i]     = U4{n, i, vo, a4};
                dst[2 * i + 1] = U4{poff, k, 0u, 0u};
            }
        }
    }

    Inputs  inputs;
    Outputs outputs;
    void    declare() override {
        declare_input(inputs.sizes, "sizes", cs::unspecified_shape<1>());
        declare_input(inputs.offs, "offs", cs::unspecified_shape<1>());
        declare_input(inputs.rank, "rank", cs::unspecified_shape<1>());
        declare_output(outputs.map1, "map1", cs::unspecified_shape<1>());
        declare_output(outputs.mapR, "mapR", cs::unspecified_shape<1>());
        declare_output(outputs.map2, "map2", cs::unspecified_shape<1>());
    }
    std::unique_ptr<cs::LaunchDescription> define_forward() override {
        return create_launch_config(cs::range_desc(cs::range<1>{inputs.sizes.shape()[0] * kWg}), cs::wg_desc(cs::range<1>{kWg}));
    }
};

/// A phase's grid (one work-group per unit) must stay inside its unit map: the map is sized for the smallest R and RT.
inline std::size_t phase_items(std::uint64_t units, std::size_t map_elems, const char* who) {
    if (units == 0 || units * kMapWords > map_elems)
        throw std::logic_error(std::string{"batch: the "} + who + " grid of " + std::to_string(units) + " work-groups does not fit its unit map (" +
                               std::to_string(map_elems / kMapWords) + " entries)");
    return static_cast<std::size_t>(units) * kWg;
}

/// Pass 1 of every instance: P (strip partials) from A and u1 v1 u2 v2 y.
class RaggedPass1 : public cs::Component<RaggedPass1> {
public:
    static constexpr int max_work_group_size = kWg;
    struct Inputs {
        cs::Input<DType::u32, 1> params, offs, map;
        cs::Input<DType::f32, 1> A, u1, v1, u2, v2, y;
    };
    struct Outputs {
        cs::Output<DType::f32, 1> P;
    };
    __device__ static void compute_forward(cs::GroupItem<1> it, const Inputs& in, const Outputs& out) {
        const Cols cols{cs::raw::data(in.offs), static_cast<std::uint32_t>((in.offs.size() - 1) / kColumns)};
        Unit       w;
        if (!map_unit(cs::raw::data(in.map), in.map.size(), static_cast<std::uint32_t>(it.get_group(0)), cols.total(3), w)) return;
        Arrays a;
        a.R  = in.params[std::size_t{0}];
        a.A  = cs::raw::data(in.A);
        a.u1 = cs::raw::data(in.u1);
        a.v1 = cs::raw::data(in.v1);
        a.u2 = cs::raw::data(in.u2);
        a.v2 = cs::raw::data(in.v2);
        a.y  = cs::raw::data(in.y);
        a.P  = cs::raw::data(out.P);
        unit_pass1(a, w, static_cast<std::uint32_t>(it.get_local_id(0)));
    }

    Inputs  inputs;
    Outputs outputs;
    void    declare() override {
        declare_input(inputs.params, "params", 3);
        declare_input(inputs.offs, "offs", cs::unspecified_shape<1>());
        declare_input(inputs.map, "map", cs::unspecified_shape<1>());
        for (auto [w, name] : {std::pair{&inputs.A, "A"}, {&inputs.u1, "u1"}, {&inputs.v1, "v1"}, {&inputs.u2, "u2"}, {&inputs.v2, "v2"},
                               {&inputs.y, "y"}})
            declare_input(*w, name, cs::unspecified_shape<1>());
        declare_output(outputs.P, "P", cs::unspecified_shape<1>());
    }
    std::unique_ptr<cs::LaunchDescription> define_forward() override {
        return create_launch_config(cs::range_desc(cs::range<1>{phase_items(knobs_->units[0], inputs.map.shape()[0], "pass 1")}),
                                    cs::wg_desc(cs::range<1>{kWg}));
    }
    explicit RaggedPass1(std::shared_ptr<Knobs> k) : knobs_{std::move(k)} {}

private:
    std::shared_ptr<Knobs> knobs_;
};

/// The column reductions of every instance: x = beta sum_c P[c] + z.
class RaggedReduce : public cs::Component<RaggedReduce> {
public:
    static constexpr int max_work_group_size = kWg;
    struct Inputs {
        cs::Input<DType::u32, 1> params, offs, map;
        cs::Input<DType::f32, 1> P, z, beta;
    };
    struct Outputs {
        cs::Output<DType::f32, 1> x;
    };
    __device__ static void compute_forward(cs::GroupItem<1> it, const Inputs& in, const Outputs& out) {
        const Cols cols{cs::raw::data(in.offs), static_cast<std::uint32_t>((in.offs.size() - 1) / kColumns)};
        Unit       w;
        if (!map_unit(cs::raw::data(in.map), in.map.size(), static_cast<std::uint32_t>(it.get_group(0)), cols.total(4), w)) return;
        Arrays a;
        a.R    = in.params[std::size_t{0}];
        a.P_in = cs::raw::data(in.P);
        a.z    = cs::raw::data(in.z);
        a.x    = cs::raw::data(out.x);
        a.beta = in.beta[std::size_t{0}];
        unit_reduce(a, w, static_cast<std::uint32_t>(it.get_local_id(0)));
    }

    Inputs  inputs;
    Outputs outputs;
    void    declare() override {
        declare_input(inputs.params, "params", 3);
        declare_input(inputs.offs, "offs", cs::unspecified_shape<1>());
        declare_input(inputs.map, "map", cs::unspecified_shape<1>());
        declare_input(inputs.P, "P", cs::unspecified_shape<1>());
        declare_input(inputs.z, "z", cs::unspecified_shape<1>());
        declare_input(inputs.beta, "beta", 1);
        declare_output(outputs.x, "x", cs::unspecified_shape<1>());
    }
    void prepare_forward() override { set_shape(outputs.x, shape_of(inputs.z)[0]); }
    std::unique_ptr<cs::LaunchDescription> define_forward() override {
        return create_launch_config(cs::range_desc(cs::range<1>{phase_items(knobs_->units[1], inputs.map.shape()[0], "column reduction")}),
                                    cs::wg_desc(cs::range<1>{kWg}));
    }
    explicit RaggedReduce(std::shared_ptr<Knobs> k) : knobs_{std::move(k)} {}

private:
    std::shared_ptr<Knobs> knobs_;
};

/// Pass 2 of every instance: w = alpha Â x, RT rows per work-group (4 or 8: the work-group sweeps the columns of its rows
/// together; 16 or 32: each wave owns RT / 4 rows).
class RaggedPass2 : public cs::Component<RaggedPass2> {
public:
    static constexpr int max_work_group_size = kWg;
    struct Inputs {
        cs::Input<DType::u32, 1> params, offs, map;
        cs::Input<DType::f32, 1> A, u1, v1, u2, v2, x, alpha;
    };
    struct Outputs {
        cs::Output<DType::f32, 1> w;
    };
    __device__ static void compute_forward(cs::GroupItem<1> it, const Inputs& in, const Outputs& out) {
        const Cols cols{cs::raw::data(in.offs), static_cast<std::uint32_t>((in.offs.size() - 1) / kColumns)};
        Unit       w;
        if (!map_unit(cs::raw::data(in.map), in.map.size(), static_cast<std::uint32_t>(it.get_group(0)), cols.total(5), w)) return;
        Arrays a;
        a.A     = cs::raw::data(in.A);
        a.u1    = cs::raw::data(in.u1);
        a.v1    = cs::raw::data(in.v1);
        a.u2    = cs::raw::data(in.u2);
        a.v2    = cs::raw::data(in.v2);
        a.x_in  = cs::raw::data(in.x);
        a.w     = cs::raw::data(out.w);
        a.alpha = in.alpha[std::size_t{0}];
        const std::uint32_t tid = static_cast<std::uint32_t>(it.get_local_id(0));
        switch (in.params[std::size_t{1}]) {
            case 4: unit_pass2<4>(a, w, tid); break;
            case 8: unit_pass2<8>(a, w, tid); break;
            case 16: unit_pass2_waves<4>(a, w, tid); break;
            default: unit_pass2_waves<8>(a, w, tid); break;  // 32: the host accepts no other value
        }
    }

    Inputs  inputs;
    Outputs outputs;
    void    declare() override {
        declare_input(inputs.params, "params", 3);
        declare_input(inputs.offs, "offs", cs::unspecified_shape<1>());
        declare_input(inputs.map, "map", cs::unspecified_shape<1>());
        for (auto [w, name] : {std::pair{&inputs.A, "A"}, {&inputs.u1, "u1"}, {&inputs.v1, "v1"}, {&inputs.u2, "u2"}, {&inputs.v2, "v2"},
                               {&inputs.x, "x"}})
            declare_input(*w, name, cs::unspecified_shape<1>());
        declare_input(inputs.alpha, "alpha", 1);
        declare_output(outputs.w, "w", cs::unspecified_shape<1>());
    }
    void prepare_forward() override { set_shape(outputs.w, shape_of(inputs.x)[0]); }
    std::unique_ptr<cs::LaunchDescription> define_forward() override {
        return create_launch_config(cs::range_desc(cs::range<1>{phase_items(knobs_->units[2], inputs.map.shape()[0], "pass 2")}),
                                    cs::wg_desc(cs::range<1>{kWg}));
    }
    explicit RaggedPass2(std::shared_ptr<Knobs> k) : knobs_{std::move(k)} {}

private:
    std::shared_ptr<Knobs> knobs_;
};

CS_COMPONENTS_END

}  // namespace batch
