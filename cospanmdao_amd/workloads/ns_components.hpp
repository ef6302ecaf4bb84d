#pragma once

// =============================================================================
// cospanmdao_amd/workloads/ns_components.hpp — the incompressible Navier-Stokes residual components
// =============================================================================
//
// On a periodic nx x ny grid (row-major (y, x), spacing h, viscosity nu), second-order central differences:
//   Ru = u u_x + v u_y + p_x - nu lap(u),   Rv = u v_x + v v_y + p_y - nu lap(v),   Rc = u_x + v_y
//   f_x = (f[i][j+1] - f[i][j-1]) / 2h,  f_y = (f[i+1][j] - f[i-1][j]) / 2h,
//   lap(f) = ((f_e - f_c) + (f_w - f_c)) + ((f_n - f_c) + (f_s - f_c))  / h^2   (differences first: exact for smooth fields)
// followed by the global sums of squares of Ru, Rv, Rc in FP64 and their L2 norms.
//
// The physics is this build's (the reference's ns-residual is compressible, examples/ns-residual/physics.hpp); the
// STRUCTURE follows the reference's fused example (examples/ns-residual/systems.hpp:88-140, README.md:40-57): a tile
// stage writing the residual and a per-tile partial of the diagnostics, a global_barrier, then a deterministic reduction
// of the partials in a fixed order (tile.hpp:212-229), all in one fused group (its barrier is a launch boundary:
// the stencil, then the reduction, as two launches of one kernel); and an unfused variant of separate kernels with
// intermediates in global memory (components.hpp:636-852).
//
// NsTile<S>: a work-group owns TY rows x (256 * V) columns; each work-item owns V = 4 consecutive columns (one 16-byte
// load per field and row for f32 fields, 8 bytes for bf16) and marches down its rows with a ring of five rows in
// registers (i - 1, i, i + 1 in use, i + 2 in flight, i + 3 loading while row i is computed), so every field element is
// read from HBM once (plus two halo rows per tile) and a row's loads have two iterations to arrive. The x neighbours of a work-item's V columns come from the adjacent
// lanes by DPP wave shifts; only lanes 0 and 63 load their halo column from memory. Tiles are mapped to XCDs in
// contiguous runs and sweep down in even strips, up in odd ones, so the two tiles on a strip boundary read the rows they
// share at the same time, from the same L2. S is the storage of u, v, p: f32 or bf16.

#include <cstdint>
#include <memory>
#include <type_traits>

#include "naga/cospan/cospan.hpp"

namespace ns {

namespace cs = naga::cospan;
using cs::DType;

inline constexpr int kWg = 256;

__host__ __device__ inline std::uint32_t u32(std::size_t v) { return static_cast<std::uint32_t>(v); }

CS_COMPONENTS_BEGIN
template <DType S>
struct Field;
template <>
struct Field<DType::f32> {
    using T                   = float;
    static constexpr int kVec = 4;
    using Raw                 = float __attribute__((ext_vector_type(4)));
    __device__ static void widen(const Raw& r, float (&a)[kVec]) {
#pragma unroll
        for (int e = 0; e < kVec; ++e) a[e] = r[e];
    }
    __device__ static float one(const T& v) { return v; }
};
/// bf16 fields keep the f32 layout of work (4 columns per work-item: one 8-byte load per field and row), so the tile
/// kernel has the same register footprint and occupancy as with f32 fields and simply reads half the bytes.
template <>
struct Field<DType::bf16> {
    using T                   = cs::bf16;
    static constexpr int kVec = 4;
    using Raw                 = std::uint32_t __attribute__((ext_vector_type(2)));
    __device__ static void widen(const Raw& r, float (&a)[kVec]) {
#pragma unroll
        for (int q = 0; q < 2; ++q) {
            a[2 * q]     = __uint_as_float(r[q] << 16);
            a[2 * q + 1] = __uint_as_float(r[q] & 0xffff0000u);
        }
    }
    __device__ static float one(const T& v) { return v.to_float(); }
};

/// The coefficients every stage reads: {1 / 2h, 1 / h^2, nu} as f32.
struct Coef {
    float inv2h, invh2, nu;
};
__device__ inline Coef coef_of(const cs::Input<DType::f32, 1>& c) {
    return Coef{c[std::size_t{0}], c[std::size_t{1}], c[std::size_t{2}]};
}

/// One cell of the residual in f32 (the shared formula of every variant).
__device__ inline void cell(const Coef& k, float uc, float ue, float uw, float un, float us, float vc, float ve, float vw, float vn,
                            float vs, float pe, float pw, float pn, float ps, float& ru, float& rv, float& rc) {
#pragma clang fp contract(off)  // the only fused multiply-adds are the explicit ones: every variant rounds alike
    const float ux = (ue - uw) * k.inv2h, uy = (un - us) * k.inv2h;
    const float vx = (ve - vw) * k.inv2h, vy = (vn - vs) * k.inv2h;
    const float px = (pe - pw) * k.inv2h, py = (pn - ps) * k.inv2h;
    const float lu = (((ue - uc) + (uw - uc)) + ((un - uc) + (us - uc))) * k.invh2;
    const float lv = (((ve - vc) + (vw - vc)) + ((vn - vc) + (vs - vc))) * k.invh2;
    ru             = fmaf(-k.nu, lu, fmaf(uc, ux, fmaf(vc, uy, px)));
    rv             = fmaf(-k.nu, lv, fmaf(uc, vx, fmaf(vc, vy, py)));
    rc             = ux + vy;
}

/// Lane l gets lane l-1's value (DPP wave_shr:1) / lane l+1's value (wave_shl:1); the edge lane keeps `x`.
__device__ inline float from_left_lane(float x) {
    return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(__builtin_bit_cast(int, x), __builtin_bit_cast(int, x), 0x138, 0xf, 0xf, false));
}
__device__ inline float from_right_lane(float x) {
    return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(__builtin_bit_cast(int, x), __builtin_bit_cast(int, x), 0x130, 0xf, 0xf, false));
}

/// The logical tile of linear work-group `lin` of `total`: runs of total / 8 consecutive tiles per XCD (work-groups are
/// dispatched to the 8 XCDs round-robin), so vertically adjacent tiles share an L2.
__device__ inline std::uint32_t xcd_tile(std::uint32_t lin, std::uint32_t total) {
    if (total % 8 != 0) return lin;
    return (lin % 8) * (total / 8) + lin / 8;
}

// =============================================================================
// NsTile<S>: the fused stencil stage
// =============================================================================
template <DType S>
class NsTile : public cs::Component<NsTile<S>> {
public:
    static constexpr int max_work_group_size = kWg;
    static constexpr int min_waves_per_eu    = 4;  ///< the five-row ring fits 128 VGPRs: four work-groups per CU
    using F                                  = Field<S>;
    static constexpr int kVec                = F::kVec;
    static constexpr int kCols               = kWg * kVec;

    struct Inputs {
        cs::Input<S, 2>          u, v, p;
        cs::Input<DType::f32, 1> coef;  ///< {1 / 2h, 1 / h^2, nu}
        cs::Input<DType::u32, 1> ty;    ///< rows per tile
    };
    struct Outputs {
        cs::Output<DType::f32, 2> Ru, Rv, Rc;
        cs::Output<DType::f64, 2> part;  ///< [tiles, 3]: per tile, the sums of squares of Ru, Rv, Rc
    };

    __device__ static void compute_forward(cs::GroupItem<2> it, const Inputs& in, const Outputs& out) {
        const std::uint32_t ny = u32(in.u.shape()[0]), nx = u32(in.u.shape()[1]);
        const std::uint32_t TY = in.ty[std::size_t{0}];
        const std::uint32_t nblk = u32(it.get_group_range(1)), ntile = u32(it.get_group_range(0)) * nblk;
        const std::uint32_t tile = xcd_tile(u32(it.get_group_linear_id()), ntile);
        const std::uint32_t r = tile / nblk, cb = tile % nblk;
        const std::uint32_t lane = static_cast<std::uint32_t>(cs::wave::lane_id());
        const std::uint32_t tid  = u32(it.get_local_id(1));
        const std::uint32_t j0r  = (cb * kWg + tid) * kVec;
        const bool          live = j0r < nx;
        const std::uint32_t j0   = live ? j0r : nx - kVec;  // idle lanes shadow the last vector: DPP reads stay defined
        const std::uint32_t i0 = r * TY, i1 = i0 + TY < ny ? i0 + TY : ny;
        const Coef          k  = coef_of(in.coef);
        const auto*         U  = cs::raw::data(in.u);
        const auto*         V  = cs::raw::data(in.v);
        const auto*         P  = cs::raw::data(in.p);
        const std::size_t   ld = in.u.stride()[0];
        // The x neighbours a lane cannot take from the adjacent lane: lane 0's left neighbour and the right neighbour of
        // the wave's last live lane (with the periodic wrap). Both columns are the same for the whole wave.
        const std::uint32_t jw0  = j0r - lane * kVec < nx ? j0r - lane * kVec : nx - kVec;  // the wave's first column
        const std::uint32_t last = (nx - kVec - jw0) / kVec < 63 ? (nx - kVec - jw0) / kVec : 63;
        const std::uint32_t jwl  = jw0 + last * kVec;  // the last live lane's first column
        const std::uint32_t jl = jw0 == 0 ? nx - 1 : jw0 - 1, jr = jwl + kVec >= nx ? 0 : jwl + kVec;
        const bool left_mem = lane == 0, right_mem = lane == last;

        // A row of the ring: the work-item's kVec columns of u, v, p as loaded (widened to f32 where they are used).
        struct Row {
            typename F::Raw u, v, p;
        };
        // The wave's two halo columns of u, v, p (wave-uniform addresses, loaded by every lane alike: no divergent
        // branch in the loop). Only the middle row of the window needs them.
        struct Halo {
            float ul, ur, vl, vr, pl, pr;
        };
        auto load_row = [&](std::uint32_t i, Row& w) __attribute__((always_inline)) {
            // cached loads: the halo rows a tile shares with its neighbours are then served by the L2
            const std::size_t off = static_cast<std::size_t>(i) * ld;
            w.u = *reinterpret_cast<const typename F::Raw*>(U + off + j0);
            w.v = *reinterpret_cast<const typename F::Raw*>(V + off + j0);
            w.p = *reinterpret_cast<const typename F::Raw*>(P + off + j0);
        };
        auto load_halo = [&](std::uint32_t i, Halo& h) __attribute__((always_inline)) {
            const std::size_t off = static_cast<std::size_t>(i) * ld;
            h.ul = F::one(U[off + jl]);
            h.vl = F::one(V[off + jl]);
            h.pl = F::one(P[off + jl]);
            h.ur = F::one(U[off + jr]);
            h.vr = F::one(V[off + jr]);
            h.pr = F::one(P[off + jr]);
        };
        double s0 = 0.0, s1 = 0.0, s2 = 0.0;
        // Row i of the tile from the rows i - 1, i, i + 1 and row i's halo: the residual stored, its squares summed.
        auto emit = [&](std::uint32_t i, const Row& wm, const Row& wc, const Row& wn, const Halo& h) __attribute__((always_inline)) {
            float mu[kVec], mv[kVec], mp[kVec], cu[kVec], cv[kVec], cp[kVec], nu_[kVec], nv[kVec], np[kVec];
            F::widen(wm.u, mu);
            F::widen(wm.v, mv);
            F::widen(wm.p, mp);
            F::widen(wc.u, cu);
            F::widen(wc.v, cv);
            F::widen(wc.p, cp);
            F::widen(wn.u, nu_);
            F::widen(wn.v, nv);
            F::widen(wn.p, np);
            // x neighbours of this work-item's columns: from the adjacent lanes, else from memory (the halo)
            float ul = from_left_lane(cu[kVec - 1]), ur = from_right_lane(cu[0]);
            float vl = from_left_lane(cv[kVec - 1]), vr = from_right_lane(cv[0]);
            float pl = from_left_lane(cp[kVec - 1]), pr = from_right_lane(cp[0]);
            if (left_mem) {
                ul = h.ul;
                vl = h.vl;
                pl = h.pl;
            }
            if (right_mem) {
                ur = h.ur;
                vr = h.vr;
                pr = h.pr;
            }
            float ru[kVec], rv[kVec], rc[kVec];
#pragma unroll
            for (int e = 0; e < kVec; ++e) {
                const float uw = e == 0 ? ul : cu[e - 1], ue = e == kVec - 1 ? ur : cu[e + 1];
                const float vw = e == 0 ? vl : cv[e - 1], ve = e == kVec - 1 ? vr : cv[e + 1];
                const float pw = e == 0 ? pl : cp[e - 1], pe = e == kVec - 1 ? pr : cp[e + 1];
                cell(k, cu[e], ue, uw, nu_[e], mu[e], cv[e], ve, vw, nv[e], mv[e], pe, pw, np[e], mp[e], ru[e], rv[e], rc[e]);
            }
            if (live) {
                using F4 = float __attribute__((ext_vector_type(4)));
                float* oru = &out.Ru[cs::id<2>{i, j0}];
                float* orv = &out.Rv[cs::id<2>{i, j0}];
                float* orc = &out.Rc[cs::id<2>{i, j0}];
#pragma unroll
                for (int q = 0; q < kVec / 4; ++q) {
                    __builtin_nontemporal_store(F4{ru[4 * q], ru[4 * q + 1], ru[4 * q + 2], ru[4 * q + 3]}, reinterpret_cast<F4*>(oru) + q);
                    __builtin_nontemporal_store(F4{rv[4 * q], rv[4 * q + 1], rv[4 * q + 2], rv[4 * q + 3]}, reinterpret_cast<F4*>(orv) + q);
                    __builtin_nontemporal_store(F4{rc[4 * q], rc[4 * q + 1], rc[4 * q + 2], rc[4 * q + 3]}, reinterpret_cast<F4*>(orc) + q);
                }
#pragma unroll
                for (int e = 0; e < kVec; ++e) {
                    s0 = fma(static_cast<double>(ru[e]), static_cast<double>(ru[e]), s0);
                    s1 = fma(static_cast<double>(rv[e]), static_cast<double>(rv[e]), s1);
                    s2 = fma(static_cast<double>(rc[e]), static_cast<double>(rc[e]), s2);
                }
            }
        };
        // The sweep over the tile's rows, top-down (kDown) or bottom-up. Tiles in even strips sweep down and tiles in odd
        // strips up, so the two tiles on either side of a strip boundary reach the rows they share (each one's edge row is
        // the other's halo row) at the same time, while those rows are in their XCD's L2: the halo rows are read from HBM
        // once, not twice. Position s = 0 .. len - 1 is the s-th row swept; s = -1 and s = len are the halo rows behind
        // and ahead of the sweep. A ring of five rows: while row s is computed from rows s - 1, s, s + 1, row s + 2 is in
        // flight and row s + 3 starts loading, so each row has two iterations to arrive. The halo columns of row s + 1
        // are issued one iteration ahead, BEFORE row s + 3: loads complete in order, so waiting for the halo never waits
        // for the newer row. Invariant at the top of position s: r0 = s - 1, r1 = s, r2 = s + 1, r3 = s + 2, hn = halo
        // columns of s.
        auto sweep = [&](auto down_tag) __attribute__((always_inline)) {
            constexpr bool     kDown = decltype(down_tag)::value;
            const int          len   = static_cast<int>(i1 - i0);
            auto               at    = [&](int q) {
                const std::int64_t v = kDown ? static_cast<std::int64_t>(i0) + q : static_cast<std::int64_t>(i1) - 1 - q;
                return static_cast<std::uint32_t>(v < 0 ? v + ny : v >= static_cast<std::int64_t>(ny) ? v - ny : v);
            };
            // the row behind the sweep is i - 1 going down and i + 1 going up
            auto put = [&](int q, const Row& behind, const Row& c, const Row& ahead, const Halo& h) __attribute__((always_inline)) {
                if constexpr (kDown)
                    emit(at(q), behind, c, ahead, h);
                else
                    emit(at(q), ahead, c, behind, h);
            };
            Row  r0{}, r1{}, r2{}, r3{}, r4{};
            Halo hc{}, hn{};
            load_row(at(-1), r0);
            load_row(at(0), r1);
            load_row(at(1), r2);
            load_halo(at(0), hn);
            if (len > 1) load_row(at(2), r3);
            // Steady state: five rows per pass, every load in range, no exit inside (the wait counts stay exact). The loop
            // is unrolled by the ring's length, so the rows never move between registers.
            auto step = [&](int q, const Row& m, const Row& c, const Row& n, Row& next) __attribute__((always_inline)) {
                hc = hn;
                load_halo(at(q + 1), hn);
                load_row(at(q + 3), next);  // the middle row two iterations from now
                put(q, m, c, n, hc);
            };
            int q = 0;
            for (; q + 7 <= len; q += 5) {
                step(q, r0, r1, r2, r4);
                step(q + 1, r1, r2, r3, r0);
                step(q + 2, r2, r3, r4, r1);
                step(q + 3, r3, r4, r0, r2);
                step(q + 4, r4, r0, r1, r3);
            }
            // The last rows (at most six): the ring shifts by register moves.
            for (; q < len; ++q) {
                hc = hn;
                if (q + 1 < len) load_halo(at(q + 1), hn);
                if (q + 2 < len) load_row(at(q + 3), r4);
                put(q, r0, r1, r2, hc);
                r0 = r1;
                r1 = r2;
                r2 = r3;
                r3 = r4;
            }
        };
        if ((r & 1u) == 0u)
            sweep(std::true_type{});
        else
            sweep(std::false_type{});
        // the tile's partial: a balanced tree over the work-group (deterministic)
        const auto g = it.get_group();
        s0           = cs::reduce_over_group(g, s0, cs::plus<double>());
        s1           = cs::reduce_over_group(g, s1, cs::plus<double>());
        s2           = cs::reduce_over_group(g, s2, cs::plus<double>());
        if (tid == 0) {
            out.part[cs::id<2>{tile, 0}] = s0;
            out.part[cs::id<2>{tile, 1}] = s1;
            out.part[cs::id<2>{tile, 2}] = s2;
        }
    }

    Inputs  inputs;
    Outputs outputs;

    explicit NsTile(int ty) : ty_{ty} {}
    void declare() override {
        this->declare_input(inputs.u, "u", cs::unspecified_shape<2>());
        this->declare_input(inputs.v, "v", cs::unspecified_shape<2>());
        this->declare_input(inputs.p, "p", cs::unspecified_shape<2>());
        this->declare_input(inputs.coef, "coef", 3);
        this->declare_input(inputs.ty, "ty", 1);
        this->init_input("ty", cs::init::fill(static_cast<std::uint32_t>(ty_)));
        this->declare_output(outputs.Ru, "Ru", cs::unspecified_shape<2>());
        this->declare_output(outputs.Rv, "Rv", cs::unspecified_shape<2>());
        this->declare_output(outputs.Rc, "Rc", cs::unspecified_shape<2>());
        this->declare_output(outputs.part, "part", cs::unspecified_shape<2>());
    }
    void prepare_forward() override {
        const cs::range<2> s = this->shape_of(inputs.u);
        if (s[1] % kVec != 0 || s[1] < static_cast<std::size_t>(kVec) || s[0] < 2)
            throw std::invalid_argument("ns: NsTile needs nx a multiple of " + std::to_string(kVec) + " and ny >= 2");
        this->set_shape(outputs.Ru, s[0], s[1]);
        this->set_shape(outputs.Rv, s[0], s[1]);
        this->set_shape(outputs.Rc, s[0], s[1]);
        this->set_shape(outputs.part, tiles(s), 3);
    }
    std::unique_ptr<cs::LaunchDescription> define_forward() override {
        const cs::range<2> s = inputs.u.shape();
        return this->create_launch_config(cs::range_desc(cs::range<2>{strips(s), blocks(s) * kWg}), cs::wg_desc(cs::range<2>{1, kWg}));
    }

private:
    std::size_t strips(const cs::range<2>& s) const { return (s[0] + ty_ - 1) / ty_; }
    std::size_t blocks(const cs::range<2>& s) const { return (s[1] + kCols - 1) / kCols; }
    std::size_t tiles(const cs::range<2>& s) const { return strips(s) * blocks(s); }
    int         ty_;
};

// =============================================================================
// NsNorm: the fixed-order reduction of per-tile partials (one wave)
// =============================================================================
class NsNorm : public cs::Component<NsNorm> {
public:
    static constexpr int max_work_group_size = kWg;
    struct Inputs {
        cs::Input<DType::f64, 2> part;  ///< [tiles, 3]
    };
    struct Outputs {
        cs::Output<DType::f64, 1> sq;    ///< the three sums of squares
        cs::Output<DType::f64, 1> norm;  ///< ||Ru||, ||Rv||, ||Rc||, and the norm of all three
    };
    /// Up to kOneWave partials (the fused tile's), one wave does it all and the other waves return at once: lane l sums
    /// partials l, l + 64, ... in order, then a butterfly over the lanes; no LDS, no barrier. More partials (the unfused
    /// variant's one per 4096 cells): every work-item sums a stride of 256, each wave reduces by butterfly, and the four
    /// wave results meet through LDS after one barrier, summed pairwise in wave order. Both orders are fixed.
    static constexpr std::uint32_t kOneWave = 1024;
    __device__ static void compute_forward(cs::GroupItem<1> it, const Inputs& in, const Outputs& out) {
        const std::uint32_t tiles = u32(in.part.shape()[0]);
        const std::uint32_t tid   = u32(it.get_local_id(0));
        const std::uint32_t nt    = tiles <= kOneWave ? static_cast<std::uint32_t>(cs::wave::kSize) : static_cast<std::uint32_t>(kWg);
        if (tid >= nt) return;  // uniform: the branch depends on the partial count only
        double s[3] = {0.0, 0.0, 0.0};
        for (std::uint32_t t = tid; t < tiles; t += nt)
            for (int c = 0; c < 3; ++c) s[c] += in.part[cs::id<2>{t, static_cast<std::size_t>(c)}];
        for (int c = 0; c < 3; ++c) s[c] = cs::wave::reduce(s[c], cs::plus<double>());
        if (nt == static_cast<std::uint32_t>(kWg)) {
            constexpr int     kWaves = kWg / cs::wave::kSize;
            __shared__ double wave_sums[kWaves][3];
            const int         wv = static_cast<int>(tid) / cs::wave::kSize;
            if (cs::wave::lane_id() == 0)
                for (int c = 0; c < 3; ++c) wave_sums[wv][c] = s[c];
            __syncthreads();
            static_assert(kWaves == 4, "cospan ns: the norm's work-group is four waves");
            for (int c = 0; c < 3; ++c) s[c] = (wave_sums[0][c] + wave_sums[1][c]) + (wave_sums[2][c] + wave_sums[3][c]);
        }
        if (tid == 0) {
            for (int c = 0; c < 3; ++c) {
                out.sq[cs::id<1>{static_cast<std::size_t>(c)}]   = s[c];
                out.norm[cs::id<1>{static_cast<std::size_t>(c)}] = sqrt(s[c]);
            }
            out.norm[cs::id<1>{3}] = sqrt((s[0] + s[1]) + s[2]);
        }
    }

    Inputs  inputs;
    Outputs outputs;

    void declare() override {
        declare_input(inputs.part, "part", cs::unspecified_shape<2>());
        declare_output(outputs.sq, "sq", 3);
        declare_output(outputs.norm, "norm", 4);
    }
    std::unique_ptr<cs::LaunchDescription> define_forward() override {
        return create_launch_config(cs::range_desc(cs::range<1>{kWg}), cs::wg_desc(cs::range<1>{kWg}));
    }
};

// =============================================================================
// The unfused variant: gradients -> intermediates in global memory -> combine -> squares -> NsNorm
// =============================================================================

/// One index per cell: the eight derivative fields (u_x, u_y, v_x, v_y, p_x, p_y, lap u, lap v) as [8, ny, nx].
/// Written with nontemporal stores, as are NsCombine's residuals: each is read back only by a later launch.
template <DType S>
class NsGradients : public cs::Component<NsGradients<S>> {
public:
    struct Inputs {
        cs::Input<S, 2>          u, v, p;
        cs::Input<DType::f32, 1> coef;
    };
    struct Outputs {
        cs::Output<DType::f32, 3> d;
    };
    __device__ static void compute_forward(cs::id<2> c, const Inputs& in, const Outputs& out) {
#pragma clang fp contract(off)
        const std::size_t ny = in.u.shape()[0], nx = in.u.shape()[1], i = c[0], j = c[1];
        const Coef        k  = coef_of(in.coef);
        const std::size_t jm = j == 0 ? nx - 1 : j - 1, jp = j + 1 == nx ? 0 : j + 1;
        const std::size_t im = i == 0 ? ny - 1 : i - 1, ip = i + 1 == ny ? 0 : i + 1;
        auto              f  = [&](const cs::Input<S, 2>& a, std::size_t y, std::size_t x) { return Field<S>::one(a[cs::id<2>{y, x}]); };
        const float uc = f(in.u, i, j), ue = f(in.u, i, jp), uw = f(in.u, i, jm), un = f(in.u, ip, j), us = f(in.u, im, j);
        const float vc = f(in.v, i, j), ve = f(in.v, i, jp), vw = f(in.v, i, jm), vn = f(in.v, ip, j), vs = f(in.v, im, j);
        const float pe = f(in.p, i, jp), pw = f(in.p, i, jm), pn = f(in.p, ip, j), ps = f(in.p, im, j);
        __builtin_nontemporal_store((ue - uw) * k.inv2h, &out.d[cs::id<3>{0, i, j}]);
        __builtin_nontemporal_store((un - us) * k.inv2h, &out.d[cs::id<3>{1, i, j}]);
        __builtin_nontemporal_store((ve - vw) * k.inv2h, &out.d[cs::id<3>{2, i, j}]);
        __builtin_nontemporal_store((vn - vs) * k.inv2h, &out.d[cs::id<3>{3, i, j}]);
        __builtin_nontemporal_store((pe - pw) * k.inv2h, &out.d[cs::id<3>{4, i, j}]);
        __builtin_nontemporal_store((pn - ps) * k.inv2h, &out.d[cs::id<3>{5, i, j}]);
        __builtin_nontemporal_store((((ue - uc) + (uw - uc)) + ((un - uc) + (us - uc))) * k.invh2, &out.d[cs::id<3>{6, i, j}]);
        __builtin_nontemporal_store((((ve - vc) + (vw - vc)) + ((vn - vc) + (vs - vc))) * k.invh2, &out.d[cs::id<3>{7, i, j}]);
    }

    Inputs  inputs;
    Outputs outputs;
    void    declare() override {
        this->declare_input(inputs.u, "u", cs::unspecified_shape<2>());
        this->declare_input(inputs.v, "v", cs::unspecified_shape<2>());
        this->declare_input(inputs.p, "p", cs::unspecified_shape<2>());
        this->declare_input(inputs.coef, "coef", 3);
        this->declare_output(outputs.d, "d", cs::unspecified_shape<3>());
    }
    void prepare_forward() override {
        const cs::range<2> s = this->shape_of(inputs.u);
        this->set_shape(outputs.d, 8, s[0], s[1]);
    }
    std::unique_ptr<cs::LaunchDescription> define_forward() override { return this->create_launch_config(cs::range_desc(inputs.u.shape())); }
};

template <DType S>
class NsCombine : public cs::Component<NsCombine<S>> {
public:
    struct Inputs {
        cs::Input<S, 2>          u, v;
        cs::Input<DType::f32, 3> d;
        cs::Input<DType::f32, 1> coef;
    };
    struct Outputs {
        cs::Output<DType::f32, 2> Ru, Rv, Rc;
    };
    __device__ static void compute_forward(cs::id<2> c, const Inputs& in, const Outputs& out) {
#pragma clang fp contract(off)
        const std::size_t i = c[0], j = c[1];
        const Coef        k  = coef_of(in.coef);
        const float       uc = Field<S>::one(in.u[c]), vc = Field<S>::one(in.v[c]);
        auto              d  = [&](std::size_t q) { return in.d[cs::id<3>{q, i, j}]; };
        __builtin_nontemporal_store(fmaf(-k.nu, d(6), fmaf(uc, d(0), fmaf(vc, d(1), d(4)))), &out.Ru[c]);
        __builtin_nontemporal_store(fmaf(-k.nu, d(7), fmaf(uc, d(2), fmaf(vc, d(3), d(5)))), &out.Rv[c]);
        __builtin_nontemporal_store(d(0) + d(3), &out.Rc[c]);
    }

    Inputs  inputs;
    Outputs outputs;
    void    declare() override {
        this->declare_input(inputs.u, "u", cs::unspecified_shape<2>());
        this->declare_input(inputs.v, "v", cs::unspecified_shape<2>());
        this->declare_input(inputs.d, "d", cs::unspecified_shape<3>());
        this->declare_input(inputs.coef, "coef", 3);
        this->declare_output(outputs.Ru, "Ru", cs::unspecified_shape<2>());
        this->declare_output(outputs.Rv, "Rv", cs::unspecified_shape<2>());
        this->declare_output(outputs.Rc, "Rc", cs::unspecified_shape<2>());
    }
    void prepare_forward() override {
        const cs::range<2> s = this->shape_of(inputs.u);
        this->set_shape(outputs.Ru, s[0], s[1]);
        this->set_shape(outputs.Rv, s[0], s[1]);
        this->set_shape(outputs.Rc, s[0], s[1]);
    }
    std::unique_ptr<cs::LaunchDescription> define_forward() override { return this->create_launch_config(cs::range_desc(inputs.u.shape())); }
};

/// Per work-group of kCells cells: the FP64 sums of squares of Ru, Rv, Rc (the partials NsNorm reduces). Work-item l of
/// group g sums cells g kCells + k kWg + l for k = 0 .. kPerItem-1 in that order (each step one coalesced run of the
/// group), then the group reduces: one reduction per 4096 cells instead of per 256 (0.19 -> 0.05 ms at 4096^2).
class NsSquares : public cs::Component<NsSquares> {
public:
    static constexpr int         max_work_group_size = kWg;
    static constexpr std::size_t kPerItem            = 16;
    static constexpr std::size_t kCells              = kPerItem * kWg;
    struct Inputs {
        cs::Input<DType::f32, 2> Ru, Rv, Rc;
    };
    struct Outputs {
        cs::Output<DType::f64, 2> part;
    };
    __device__ static void compute_forward(cs::GroupItem<1> it, const Inputs& in, const Outputs& out) {
        const std::size_t n = in.Ru.size(), t0 = it.get_group(0) * kCells + it.get_local_id(0);
        const float*      a = cs::raw::data(in.Ru);
        const float*      b = cs::raw::data(in.Rv);
        const float*      c = cs::raw::data(in.Rc);
        double            s0 = 0.0, s1 = 0.0, s2 = 0.0;
#pragma unroll
        for (std::size_t k = 0; k < kPerItem; ++k) {
            const std::size_t t = t0 + k * kWg;
            if (t < n) {
                const float x = __builtin_nontemporal_load(a + t), y = __builtin_nontemporal_load(b + t), z = __builtin_nontemporal_load(c + t);
                s0 += static_cast<double>(x) * x;
                s1 += static_cast<double>(y) * y;
                s2 += static_cast<double>(z) * z;
            }
        }
        const auto g = it.get_group();
        s0           = cs::reduce_over_group(g, s0, cs::plus<double>());
        s1           = cs::reduce_over_group(g, s1, cs::plus<double>());
        s2           = cs::reduce_over_group(g, s2, cs::plus<double>());
        if (it.get_local_id(0) == 0) {
            const std::size_t grp = it.get_group(0);
            out.part[cs::id<2>{grp, 0}] = s0;
            out.part[cs::id<2>{grp, 1}] = s1;
            out.part[cs::id<2>{grp, 2}] = s2;
        }
    }

    Inputs  inputs;
    Outputs outputs;
    void    declare() override {
        declare_input(inputs.Ru, "Ru", cs::unspecified_shape<2>());
        declare_input(inputs.Rv, "Rv", cs::unspecified_shape<2>());
        declare_input(inputs.Rc, "Rc", cs::unspecified_shape<2>());
        declare_output(outputs.part, "part", cs::unspecified_shape<2>());
    }
    void prepare_forward() override { set_shape(outputs.part, groups(shape_of(inputs.Ru).size()), 3); }
    std::unique_ptr<cs::LaunchDescription> define_forward() override {
        return create_launch_config(cs::range_desc(cs::range<1>{groups(inputs.Ru.size()) * kWg}), cs::wg_desc(cs::range<1>{kWg}));
    }

private:
    static std::size_t groups(std::size_t n) { return (n + kCells - 1) / kCells; }
};

CS_COMPONENTS_END

}  // namespace ns
