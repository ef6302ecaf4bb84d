#pragma once

// =============================================================================
// cospanmdao_amd/workloads/chain_components.hpp — the BiCGK and local-memory components of the GEMVER -> BiCGK -> GEMVER chain
// =============================================================================
//
// BiCGK on B (n x n):  q = B p ;  s = B^T r.  The reference builds it from Gemv and GemvTPartial (two passes over B), or
// from ONE pass (examples/gemver-chain/components.hpp:268-358), and its "crafted" chain keeps tiles of B in local memory
// between the stages of a segment (components.hpp:360-569). This file holds the one-pass and the local-memory components,
// laid out for CDNA4:
//
//   BicgkOnePass          work-group = strip of R rows x 1024 columns; work-item t owns the 4 columns at 4t (16-byte loads),
//                         so wave v of the group owns the 256-column block 4 blk + v. Per row i of the strip:
//                           Ps[c][j]    += r[i] B[i][j]                              (an fma chain, rows ascending)
//                           Pq[i][blk4] = wave sum of (fma chain of B[i][j] p[j] over the work-item's 4 columns)
//                         (the wave sum is a DPP inclusive scan whose last lane holds the total: six DPP adds, no LDS)
//   RowReduce             q[i] = sum over the 256-column blocks of Pq[i][b], b ascending
//   ColReduce<false>      s[j] = sum_c Ps[c][j] (gemver_components.hpp)
//
//   crafted, tiles of kRt = 16 rows x 1024 columns (64 KiB of local memory, work-item t owns the tile's 4 columns at 4t,
//   stored as one 16-byte vector per row, so each work-item only ever touches its own part of the tile):
//   Rank2UpdateLocal      B = A + u1 v1^T + u2 v2^T for the tile (Rank2Update's arithmetic), to global B AND to the tile
//   GemvTPartialLocal     P[c][j] = fma chain over the tile's rows of B y (GemvTPartial's order with R = kRt)
//   BicgkTileLocal        loads the B tile (the ONE read of B1 in the middle launch) and makes BicgkOnePass's partials
//                         from it, in BicgkOnePass's order with R = kRt
//   Rank2UpdateFromLocal  B2 = tile + u1b v1b^T + u2b v2b^T, to global B2 (no read of B1, no dependency on q or s)
// So the crafted chain equals the one-pass chain at R = kRt bit for bit, and its G1 equals the unfused G1 at R = kRt.
//
// n must be a multiple of 256 (the reference's constraint, chain_data.hpp:75-82): a wave's 256 columns are then all inside
// the matrix or all outside it.

#include <cstdint>
#include <memory>

#include "naga/cospan/cospan.hpp"
#include "gemver_components.hpp"

namespace gemver {

inline constexpr int         kBlockCols = kWg * 4;  // columns of one BiCGK / crafted work-group (4 per work-item)
inline constexpr int         kWaveCols  = 256;      // columns of one wave: one block of Pq
inline constexpr std::size_t kRt        = 16;       // rows of a crafted tile
inline constexpr std::size_t kTileElems = kRt * kBlockCols;

using F4 = float __attribute__((ext_vector_type(4)));

namespace detail {
/// The sum of x over the 64 lanes of the wave, in the last lane (a DPP inclusive scan: row shifts 1, 2, 4, 8, then the row
/// broadcasts 15 and 31). Every lane of the wave must be active.
__device__ inline float wave_total_last(float x) { return cs::wave::inclusive_scan(x, cs::plus<float>{}, 0.0f); }

/// One row of a BiCGK tile: the work-item's 4 columns b (of row i), its s partials ps, and the wave's q partial into Pq.
__device__ inline void bicgk_row(const F4& b, const F4& pj, float ri, F4& ps, float* pq_row_blk, bool last_lane) {
#pragma unroll
    for (int e = 0; e < 4; ++e) ps[e] = fmaf(ri, b[e], ps[e]);
    const float d   = fmaf(b[3], pj[3], fmaf(b[2], pj[2], fmaf(b[1], pj[1], b[0] * pj[0])));
    const float sum = wave_total_last(d);
    if (last_lane) *pq_row_blk = sum;
}

/// Strip c and column block blk of work-group g over (n / rows) x (n / 1024, rounded up) groups; the work-item's first
/// column and whether its wave's 256 columns lie inside the matrix.
struct TileAt {
    std::uint32_t c, blk, j0, wave_block;
    bool          inside;
};
__device__ inline TileAt tile_at(std::uint32_t g, std::uint32_t n, std::uint32_t tid) {
    const std::uint32_t blocks = (n + kBlockCols - 1) / kBlockCols;
    TileAt              t;
    t.c          = g / blocks;
    t.blk        = g - t.c * blocks;
    t.j0         = t.blk * kBlockCols + tid * 4;
    t.wave_block = t.blk * (kBlockCols / kWaveCols) + tid / 64;
    t.inside     = t.wave_block * kWaveCols < n;
    return t;
}
}  // namespace detail

CS_COMPONENTS_BEGIN

// =============================================================================
// BicgkOnePass: the partials of q = B p and s = B^T r from ONE pass over B
// =============================================================================
class BicgkOnePass : public cs::Component<BicgkOnePass> {
public:
    static constexpr int max_work_group_size = kWg;
    struct Inputs {
        cs::Input<DType::f32, 2> B;
        cs::Input<DType::f32, 1> p, r;
    };
    struct Outputs {
        cs::Output<DType::f32, 2> Ps, Pq;
    };
    __device__ static void compute_forward(cs::GroupItem<1> it, const Inputs& in, const Outputs& out) {
        const std::uint32_t n = u32(in.B.shape()[0]), R = u32(in.B.shape()[0] / out.Ps.shape()[0]);
        const std::uint32_t tid = u32(it.get_local_id(0));
        const auto          t   = detail::tile_at(u32(it.get_group(0)), n, tid);
        if (!t.inside) return;  // a whole wave past the last column (no barrier follows)
        const std::size_t ldb = in.B.stride()[0], ldq = out.Pq.stride()[0];
        const float*      B   = cs::raw::data(in.B) + static_cast<std::size_t>(t.c) * R * ldb + t.j0;
        const float*      r   = cs::raw::data(in.r) + static_cast<std::size_t>(t.c) * R;
        float*            pq  = cs::raw::data(out.Pq) + static_cast<std::size_t>(t.c) * R * ldq + t.wave_block;
        const F4          pj  = *reinterpret_cast<const F4*>(cs::raw::data(in.p) + t.j0);
        const bool        last = (tid & 63u) == 63u;
        F4                ps   = F4{0.0f, 0.0f, 0.0f, 0.0f};
        // four rows per step with their loads issued first (the DPP reduction is convergent: no compiler unrolling)
        std::uint32_t k = 0;
        for (; k + 4 <= R; k += 4, B += 4 * ldb, pq += 4 * ldq) {
            F4 b[4];
#pragma unroll
            for (int q = 0; q < 4; ++q) b[q] = __builtin_nontemporal_load(reinterpret_cast<const F4*>(B + q * ldb));
#pragma unroll
            for (int q = 0; q < 4; ++q) detail::bicgk_row(b[q], pj, r[k + q], ps, pq + q * ldq, last);
        }
        for (; k < R; ++k, B += ldb, pq += ldq) {
            const F4 b = __builtin_nontemporal_load(reinterpret_cast<const F4*>(B));
            detail::bicgk_row(b, pj, r[k], ps, pq, last);
        }
        *reinterpret_cast<F4*>(cs::raw::data(out.Ps) + static_cast<std::size_t>(t.c) * out.Ps.stride()[0] + t.j0) = ps;
    }

    Inputs  inputs;
    Outputs outputs;

    void declare() override {
        declare_input(inputs.B, "B", cs::unspecified_shape<2>());
        declare_input(inputs.p, "p", cs::unspecified_shape<1>());
        declare_input(inputs.r, "r", cs::unspecified_shape<1>());
        declare_output(outputs.Ps, "Ps", cs::unspecified_shape<2>());  // hooked by the owning System: (n / R, n)
        declare_output(outputs.Pq, "Pq", cs::unspecified_shape<2>());
    }
    void prepare_forward() override {
        const cs::range<2> b = shape_of(inputs.B);
        set_shape(outputs.Pq, b[0], b[0] / kWaveCols);
    }
    std::unique_ptr<cs::LaunchDescription> define_forward() override {
        const std::size_t n = inputs.B.shape()[0], blocks = (n + kBlockCols - 1) / kBlockCols;
        return create_launch_config(cs::range_desc(cs::range<1>{outputs.Ps.shape()[0] * blocks * kWg}), cs::wg_desc(cs::range<1>{kWg}));
    }
};

// =============================================================================
// RowReduce: q[i] = sum_b Pq[i][b], b ascending
// =============================================================================
class RowReduce : public cs::Component<RowReduce> {
public:
    static constexpr int max_work_group_size = kWg;
    struct Inputs {
        cs::Input<DType::f32, 2> Pq;
    };
    struct Outputs {
        cs::Output<DType::f32, 1> q;
    };
    __device__ static void compute_forward(cs::GroupItem<1> it, const Inputs& in, const Outputs& out) {
        const std::uint32_t n = u32(in.Pq.shape()[0]), blocks = u32(in.Pq.shape()[1]);
        const std::uint32_t stride = u32(it.get_group_range(0) * it.get_local_range(0));
        const std::size_t   ld     = in.Pq.stride()[0];
        for (std::uint32_t i = u32(it.get_global_id(0)); i < n; i += stride) {
            const float* row = cs::raw::data(in.Pq) + static_cast<std::size_t>(i) * ld;
            float        acc = 0.0f;
#pragma unroll 8
            for (std::uint32_t b = 0; b < blocks; ++b) acc += row[b];
            out.q[cs::id<1>{i}] = acc;
        }
    }

    Inputs  inputs;
    Outputs outputs;

    void declare() override {
        declare_input(inputs.Pq, "Pq", cs::unspecified_shape<2>());
        declare_output(outputs.q, "q", cs::unspecified_shape<1>());
    }
    void prepare_forward() override { set_shape(outputs.q, shape_of(inputs.Pq)[0]); }
    std::unique_ptr<cs::LaunchDescription> define_forward() override {
        return create_launch_config(cs::range_desc(cs::range<1>{round_up(inputs.Pq.shape()[0], kWg)}), cs::wg_desc(cs::range<1>{kWg}));
    }
};

// =============================================================================
// The crafted components: one tile of kRt x 1024 floats of B per work-group, kept in local memory
// =============================================================================
// Work-group g of a launch over (n / kRt) x ceil(n / 1024) groups owns strip c and column block blk (detail::tile_at);
// tile[k * 1024 + 4 t + e] = B[c kRt + k][blk 1024 + 4 t + e]. Each stage processes EXACTLY ONE tile per work-group: the
// stages of a segment run one after the other inside a work-group, so the next stage finds this tile in local memory.

/// Every crafted stage declares the same one-array local memory, so the stages of a local_mem::shared segment see one tile
/// (at offset 0 of the launch's block: 16-byte aligned).
using TileMem = cs::LocalMem<cs::LocalArray<DType::f32, 1>>;

namespace detail {
__device__ inline F4* tile_of(const cs::GroupItem<1>& it, TileMem lm) {
    auto [tile] = lm.get(it);
    return reinterpret_cast<F4*>(&tile[std::size_t{0}]);
}
}  // namespace detail

/// B = A + u1 v1^T + u2 v2^T for the tile; B goes to global memory (B1 is an output) and stays in the tile. B is
/// written with nontemporal stores: only a later launch reads it back, so it need not stay in the caches.
class Rank2UpdateLocal : public cs::Component<Rank2UpdateLocal> {
public:
    static constexpr int max_work_group_size = kWg;
    struct Inputs {
        cs::Input<DType::f32, 2> A;
        cs::Input<DType::f32, 1> u1, v1, u2, v2;
    };
    struct Outputs {
        cs::Output<DType::f32, 2> B;
    };
    __device__ static void compute_forward(cs::GroupItem<1> it, const Inputs& in, const Outputs& out, TileMem lm) {
        F4* const           tile = detail::tile_of(it, lm);
        const std::uint32_t n = u32(in.A.shape()[0]), tid = u32(it.get_local_id(0));
        const auto          t = detail::tile_at(u32(it.get_group(0)), n, tid);
        if (!t.inside) return;
        const std::size_t lda = in.A.stride()[0], ldb = out.B.stride()[0], i0 = static_cast<std::size_t>(t.c) * kRt;
        const F4          v1 = *reinterpret_cast<const F4*>(cs::raw::data(in.v1) + t.j0);
        const F4          v2 = *reinterpret_cast<const F4*>(cs::raw::data(in.v2) + t.j0);
        const float *     u1 = cs::raw::data(in.u1) + i0, *u2 = cs::raw::data(in.u2) + i0;
        const float*      A  = cs::raw::data(in.A) + i0 * lda + t.j0;
        float*            B  = cs::raw::data(out.B) + i0 * ldb + t.j0;
        F4                a[kRt];
#pragma unroll
        for (std::size_t k = 0; k < kRt; ++k) a[k] = __builtin_nontemporal_load(reinterpret_cast<const F4*>(A + k * lda));
#pragma unroll
        for (std::size_t k = 0; k < kRt; ++k) {
            F4 b;
#pragma unroll
            for (int e = 0; e < 4; ++e) b[e] = rank2(a[k][e], u1[k], v1[e], u2[k], v2[e]);
            __builtin_nontemporal_store(b, reinterpret_cast<F4*>(B + k * ldb));
            tile[k * kWg + tid] = b;
        }
    }

    Inputs  inputs;
    Outputs outputs;

    void declare() override {
        declare_input(inputs.A, "A", cs::unspecified_shape<2>());
        declare_input(inputs.u1, "u1", cs::unspecified_shape<1>());
        declare_input(inputs.v1, "v1", cs::unspecified_shape<1>());
        declare_input(inputs.u2, "u2", cs::unspecified_shape<1>());
        declare_input(inputs.v2, "v2", cs::unspecified_shape<1>());
        declare_output(outputs.B, "B", cs::unspecified_shape<2>());
    }
    void prepare_forward() override {
        const cs::range<2> a = shape_of(inputs.A);
        set_shape(outputs.B, a[0], a[1]);
    }
    std::unique_ptr<cs::LaunchDescription> define_forward() override {
        const std::size_t n = inputs.A.shape()[0], groups = (n / kRt) * ((n + kBlockCols - 1) / kBlockCols);
        return create_launch_config(cs::range_desc(cs::range<1>{groups * kWg}), cs::wg_desc(cs::range<1>{kWg}),
                                    cs::local_desc(cs::range<1>{kTileElems}));
    }
};

/// P[c][j] = sum_k tile[k][j] y[c kRt + k] (fma chain, k ascending), reading the tile Rank2UpdateLocal left.
class GemvTPartialLocal : public cs::Component<GemvTPartialLocal> {
public:
    static constexpr int max_work_group_size = kWg;
    struct Inputs {
        cs::Input<DType::f32, 1> y;
    };
    struct Outputs {
        cs::Output<DType::f32, 2> P;
    };
    __device__ static void compute_forward(cs::GroupItem<1> it, const Inputs& in, const Outputs& out, TileMem lm) {
        const F4* const     tile = detail::tile_of(it, lm);
        const std::uint32_t n = u32(in.y.shape()[0]), tid = u32(it.get_local_id(0));
        const auto          t = detail::tile_at(u32(it.get_group(0)), n, tid);
        cs::group_barrier(it.get_group());  // the tile is complete
        if (!t.inside) return;
        const float* y   = cs::raw::data(in.y) + static_cast<std::size_t>(t.c) * kRt;
        F4           acc = F4{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
        for (std::size_t k = 0; k < kRt; ++k) {
            const F4 b = tile[k * kWg + tid];
#pragma unroll
            for (int e = 0; e < 4; ++e) acc[e] = fmaf(b[e], y[k], acc[e]);
        }
        *reinterpret_cast<F4*>(cs::raw::data(out.P) + static_cast<std::size_t>(t.c) * out.P.stride()[0] + t.j0) = acc;
    }

    Inputs  inputs;
    Outputs outputs;

    void declare() override {
        declare_input(inputs.y, "y", cs::unspecified_shape<1>());
        declare_output(outputs.P, "P", cs::unspecified_shape<2>());  // hooked by the owning System: (n / kRt, n)
    }
    std::unique_ptr<cs::LaunchDescription> define_forward() override {
        const std::size_t n = inputs.y.shape()[0], groups = (n / kRt) * ((n + kBlockCols - 1) / kBlockCols);
        return create_launch_config(cs::range_desc(cs::range<1>{groups * kWg}), cs::wg_desc(cs::range<1>{kWg}),
                                    cs::local_desc(cs::range<1>{kTileElems}));
    }
};

/// Loads the B tile into local memory and makes BicgkOnePass's partials (R = kRt) from it.
class BicgkTileLocal : public cs::Component<BicgkTileLocal> {
public:
    static constexpr int max_work_group_size = kWg;
    struct Inputs {
        cs::Input<DType::f32, 2> B;
        cs::Input<DType::f32, 1> p, r;
    };
    struct Outputs {
        cs::Output<DType::f32, 2> Ps, Pq;
    };
    __device__ static void compute_forward(cs::GroupItem<1> it, const Inputs& in, const Outputs& out, TileMem lm) {
        F4* const           tile = detail::tile_of(it, lm);
        const std::uint32_t n = u32(in.B.shape()[0]), tid = u32(it.get_local_id(0));
        const auto          t = detail::tile_at(u32(it.get_group(0)), n, tid);
        if (!t.inside) return;
        const std::size_t ldb = in.B.stride()[0], ldq = out.Pq.stride()[0], i0 = static_cast<std::size_t>(t.c) * kRt;
        const float*      B   = cs::raw::data(in.B) + i0 * ldb + t.j0;
        const float*      r   = cs::raw::data(in.r) + i0;
        float*            pq  = cs::raw::data(out.Pq) + i0 * ldq + t.wave_block;
        const F4          pj  = *reinterpret_cast<const F4*>(cs::raw::data(in.p) + t.j0);
        const bool        last = (tid & 63u) == 63u;
        F4                b[kRt];
#pragma unroll
        for (std::size_t k = 0; k < kRt; ++k) b[k] = __builtin_nontemporal_load(reinterpret_cast<const F4*>(B + k * ldb));
        F4 ps = F4{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
        for (std::size_t k = 0; k < kRt; ++k) {
            tile[k * kWg + tid] = b[k];
            detail::bicgk_row(b[k], pj, r[k], ps, pq + k * ldq, last);
        }
        *reinterpret_cast<F4*>(cs::raw::data(out.Ps) + static_cast<std::size_t>(t.c) * out.Ps.stride()[0] + t.j0) = ps;
    }

    Inputs  inputs;
    Outputs outputs;

    void declare() override {
        declare_input(inputs.B, "B", cs::unspecified_shape<2>());
        declare_input(inputs.p, "p", cs::unspecified_shape<1>());
        declare_input(inputs.r, "r", cs::unspecified_shape<1>());
        declare_output(outputs.Ps, "Ps", cs::unspecified_shape<2>());  // hooked by the owning System: (n / kRt, n)
        declare_output(outputs.Pq, "Pq", cs::unspecified_shape<2>());
    }
    void prepare_forward() override {
        const cs::range<2> b = shape_of(inputs.B);
        set_shape(outputs.Pq, b[0], b[0] / kWaveCols);
    }
    std::unique_ptr<cs::LaunchDescription> define_forward() override {
        const std::size_t n = inputs.B.shape()[0], groups = (n / kRt) * ((n + kBlockCols - 1) / kBlockCols);
        return create_launch_config(cs::range_desc(cs::range<1>{groups * kWg}), cs::wg_desc(cs::range<1>{kWg}),
                                    cs::local_desc(cs::range<1>{kTileElems}));
    }
};

/// G2's B2 = B1 + u1b v1b^T + u2b v2b^T with B1 from the tile BicgkTileLocal loaded: no read of B1 here.
class Rank2UpdateFromLocal : public cs::Component<Rank2UpdateFromLocal> {
public:
    static constexpr int max_work_group_size = kWg;
    struct Inputs {
        cs::Input<DType::f32, 1> u1, v1, u2, v2;
    };
    struct Outputs {
        cs::Output<DType::f32, 2> B;
    };
    __device__ static void compute_forward(cs::GroupItem<1> it, const Inputs& in, const Outputs& out, TileMem lm) {
        const F4* const     tile = detail::tile_of(it, lm);
        const std::uint32_t n = u32(in.u1.shape()[0]), tid = u32(it.get_local_id(0));
        const auto          t = detail::tile_at(u32(it.get_group(0)), n, tid);
        cs::group_barrier(it.get_group());  // the tile is complete
        if (!t.inside) return;
        const std::size_t ldb = out.B.stride()[0], i0 = static_cast<std::size_t>(t.c) * kRt;
        const F4          v1 = *reinterpret_cast<const F4*>(cs::raw::data(in.v1) + t.j0);
        const F4          v2 = *reinterpret_cast<const F4*>(cs::raw::data(in.v2) + t.j0);
        const float *     u1 = cs::raw::data(in.u1) + i0, *u2 = cs::raw::data(in.u2) + i0;
        float*            B  = cs::raw::data(out.B) + i0 * ldb + t.j0;
#pragma unroll
        for (std::size_t k = 0; k < kRt; ++k) {
            const F4 a = tile[k * kWg + tid];
            F4       b;
#pragma unroll
            for (int e = 0; e < 4; ++e) b[e] = rank2(a[e], u1[k], v1[e], u2[k], v2[e]);
            __builtin_nontemporal_store(b, reinterpret_cast<F4*>(B + k * ldb));
        }
    }

    Inputs  inputs;
    Outputs outputs;

    void declare() override {
        declare_input(inputs.u1, "u1", cs::unspecified_shape<1>());
        declare_input(inputs.v1, "v1", cs::unspecified_shape<1>());
        declare_input(inputs.u2, "u2", cs::unspecified_shape<1>());
        declare_input(inputs.v2, "v2", cs::unspecified_shape<1>());
        declare_output(outputs.B, "B", cs::unspecified_shape<2>());
    }
    void prepare_forward() override {
        const std::size_t n = shape_of(inputs.u1)[0];
        set_shape(outputs.B, n, n);
    }
    std::unique_ptr<cs::LaunchDescription> define_forward() override {
        const std::size_t n = inputs.u1.shape()[0], groups = (n / kRt) * ((n + kBlockCols - 1) / kBlockCols);
        return create_launch_config(cs::range_desc(cs::range<1>{groups * kWg}), cs::wg_desc(cs::range<1>{kWg}),
                                    cs::local_desc(cs::range<1>{kTileElems}));
    }
};

CS_COMPONENTS_END

}  // namespace gemver
