#pragma once

// =============================================================================
// cospanmdao_amd/workloads/gemver_components.hpp — the GEMVER components (kernel bodies)
// =============================================================================
//
// GEMVER:  B = A + u1 v1^T + u2 v2^T ;  x = beta * B^T y + z ;  w = alpha * B x.
//
// The reference's four components (examples/gemver-chain/components.hpp:75-266), with the floating-point contraction
// their bodies leave to the compiler fixed here (so the CPU oracle restates it bit for bit):
//   Rank2Update   B[i][j] = fma(u2_i, v2_j, fma(u1_i, v1_j, A[i][j])), strips of R rows       (components.hpp:85-97)
//   GemvTPartial  P[c][j] = sum_{k<R} B[cR+k][j] y[cR+k] as an fma chain, k ascending         (components.hpp:142-155)
//                 (both: a work-item owns 4 adjacent columns of a strip, 16-byte accesses; the reference's owns one)
//   ColReduce<Z>  x[j] = fma(scale, sum_c P[c][j], z[j]) (Z) or scale * sum (c ascending)      (components.hpp:192-204)
//   Gemv          out[i] = scale * group_reduce(fma chain over j = l, l + wg, ...)            (components.hpp:237-250)
//
// The CDNA4 two-pass form (Â = A + u1 v1^T + u2 v2^T is recomputed in registers on both passes and never stored; A is read
// twice, with 16-byte vector loads):
//   AhatTPartial<S>  pass 1: work-group (strip c, column block) computes P[c][j] = sum_{k<R} Â[cR+k][j] y[cR+k] with the
//                    SAME evaluation order as Rank2Update + GemvTPartial, so x from it equals the components' x bit for bit.
//   ColReduce<true>  x, as above.
//   AhatGemv<S, RT>  pass 2: work-group of RT rows; each work-item keeps RT running dots over its columns (column chunks
//                    of 256 * V, V elements per 16-byte load), then per row a wave butterfly and a pairwise sum of the
//                    four wave results (RT = 4, 8). RT = 16, 32: each wave owns RT/4 of the rows alone (columns of
//                    64 * V per step, one wave butterfly per row, no LDS). Reduced-precision storage accumulates each
//                    chunk's partial into a compensated (Neumaier) running sum.
// S is the storage dtype of A: f32, bf16 (exact widening), or e4m3 with a per-row f32 scale (a_ij = e4m3(q_ij) * s_i,
// widened by the gfx950 conversion instruction). Vectors and all arithmetic are f32; the rank-2 term is never quantised.

#include <cstdint>
#include <memory>

#include "naga/cospan/cospan.hpp"

namespace gemver {

namespace cs = naga::cospan;
using cs::DType;

inline constexpr int kWg = 256;  // the work-group size of every GEMVER kernel

/// The smallest multiple of m not below v (launch ranges of group-range components cover whole work-groups).
inline std::size_t round_up(std::size_t v, std::size_t m) { return (v + m - 1) / m * m; }

__host__ __device__ inline std::uint32_t u32(std::size_t v) { return static_cast<std::uint32_t>(v); }

__device__ inline float rank2(float a, float u1i, float v1j, float u2i, float v2j) { return fmaf(u2i, v2j, fmaf(u1i, v1j, a)); }

CS_COMPONENTS_BEGIN
// =============================================================================
// The reference's four components
// =============================================================================

class Rank2Update : public cs::Component<Rank2Update> {
public:
    static constexpr int max_work_group_size = kWg;
    struct Inputs {
        cs::Input<DType::f32, 2> A;
        cs::Input<DType::f32, 1> u1, v1, u2, v2;
        cs::Input<DType::u32, 1> R;
    };
    struct Outputs {
        cs::Output<DType::f32, 2> B;
    };
    /// Item t = (strip c, columns 4 q .. 4 q + 3) over (n / R) x (n / 4) items (16-byte loads and stores; n % 4 == 0).
    /// GemvTPartial uses the same mapping, so in a fused segment a work-item reads back only its own writes.
    __device__ static void compute_forward(cs::GroupItem<1> it, const Inputs& in, const Outputs& out) {
        using F4            = float __attribute__((ext_vector_type(4)));
        const std::uint32_t n = u32(in.A.shape()[1]), R = in.R[std::size_t{0}], quads = n / 4, total = (n / R) * quads;
        const std::uint32_t stride = u32(it.get_group_range(0) * it.get_local_range(0));
        const std::size_t   lda = in.A.stride()[0], ldb = out.B.stride()[0];
        for (std::uint32_t t = u32(it.get_global_id(0)); t < total; t += stride) {
            const std::uint32_t c = t / quads, j = 4 * (t - c * quads);
            const F4            v1 = *reinterpret_cast<const F4*>(cs::raw::data(in.v1) + j);
            const F4            v2 = *reinterpret_cast<const F4*>(cs::raw::data(in.v2) + j);
            const float*        a  = cs::raw::data(in.A) + static_cast<std::size_t>(c) * R * lda + j;
            float*              b  = cs::raw::data(out.B) + static_cast<std::size_t>(c) * R * ldb + j;
            const float *       u1 = cs::raw::data(in.u1) + static_cast<std::size_t>(c) * R, *u2 = cs::raw::data(in.u2) + static_cast<std::size_t>(c) * R;
            // streaming: A read once and B written once per launch, neither kept in the caches (a fused reader of B comes
            // back to it after the whole stage, long after the lines left the L2): 0.464 -> 0.427 ms at n = 16384
#pragma unroll 4
            for (std::uint32_t k = 0; k < R; ++k) {
                const F4 av = __builtin_nontemporal_load(reinterpret_cast<const F4*>(a + k * lda));
                F4       bv;
#pragma unroll
                for (int e = 0; e < 4; ++e) bv[e] = rank2(av[e], u1[k], v1[e], u2[k], v2[e]);
                __builtin_nontemporal_store(bv, reinterpret_cast<F4*>(b + k * ldb));
            }
        }
    }

    Inputs  inputs;
    Outputs outputs;

    explicit Rank2Update(int R) : R_{R} {}
    void declare() override {
        declare_input(inputs.A, "A", cs::unspecified_shape<2>());
        declare_input(inputs.u1, "u1", cs::unspecified_shape<1>());
        declare_input(inputs.v1, "v1", cs::unspecified_shape<1>());
        declare_input(inputs.u2, "u2", cs::unspecified_shape<1>());
        declare_input(inputs.v2, "v2", cs::unspecified_shape<1>());
        declare_input(inputs.R, "R", 1);
        init_input("R", cs::init::fill(u32(static_cast<std::size_t>(R_))));
        declare_output(outputs.B, "B", cs::unspecified_shape<2>());
    }
    void prepare_forward() override {
        const cs::range<2> a = shape_of(inputs.A);
        set_shape(outputs.B, a[0], a[1]);
    }
    std::unique_ptr<cs::LaunchDescription> define_forward() override {
        const std::size_t n    = inputs.A.shape()[1];
        const std::size_t work = (n / static_cast<std::size_t>(R_)) * (n / 4);
        return create_launch_config(cs::range_desc(cs::range<1>{round_up(work, kWg)}), cs::wg_desc(cs::range<1>{kWg}));
    }

private:
    int R_;
};

class GemvTPartial : public cs::Component<GemvTPartial> {
public:
    static constexpr int max_work_group_size = kWg;
    struct Inputs {
        cs::Input<DType::f32, 2> B;
        cs::Input<DType::f32, 1> y;
    };
    struct Outputs {
        cs::Output<DType::f32, 2> P;
    };
    /// Rank2Update's item mapping (strip c, 4 columns); each column's sum is one fma chain, k ascending.
    __device__ static void compute_forward(cs::GroupItem<1> it, const Inputs& in, const Outputs& out) {
        using F4            = float __attribute__((ext_vector_type(4)));
        const std::uint32_t n = u32(in.B.shape()[1]), R = u32(in.B.shape()[0] / out.P.shape()[0]), quads = n / 4, total = (n / R) * quads;
        const std::uint32_t stride = u32(it.get_group_range(0) * it.get_local_range(0));
        const std::size_t   ldb    = in.B.stride()[0];
        for (std::uint32_t t = u32(it.get_global_id(0)); t < total; t += stride) {
            const std::uint32_t c = t / quads, j = 4 * (t - c * quads);
            const float*        b = cs::raw::data(in.B) + static_cast<std::size_t>(c) * R * ldb + j;
            const float*        y = cs::raw::data(in.y) + static_cast<std::size_t>(c) * R;
            F4                  acc = F4{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll 4
            for (std::uint32_t k = 0; k < R; ++k) {
                const F4 bv = __builtin_nontemporal_load(reinterpret_cast<const F4*>(b + k * ldb));  // streamed: 0.215 -> 0.201 ms
#pragma unroll
                for (int e = 0; e < 4; ++e) acc[e] = fmaf(bv[e], y[k], acc[e]);
            }
            *reinterpret_cast<F4*>(cs::raw::data(out.P) + static_cast<std::size_t>(c) * out.P.stride()[0] + j) = acc;
        }
    }

    Inputs  inputs;
    Outputs outputs;

    void declare() override {
        declare_input(inputs.B, "B", cs::unspecified_shape<2>());
        declare_input(inputs.y, "y", cs::unspecified_shape<1>());
        declare_output(outputs.P, "P", cs::unspecified_shape<2>());  // hooked by the owning System: (n / R, n)
    }
    std::unique_ptr<cs::LaunchDescription> define_forward() override {
        const std::size_t range = outputs.P.shape()[0] * (outputs.P.shape()[1] / 4);
        return create_launch_config(cs::range_desc(cs::range<1>{round_up(range, kWg)}), cs::wg_desc(cs::range<1>{kWg}));
    }
};

/// Dots: also the partial dot products of x with v1 and v2 over each work-group's columns, dots[0][g] and dots[1][g]
/// (a fixed-order group reduction), for the factored pass 2 of reduced-precision storage.
template <bool Z, bool Dots = false>
class ColReduce : public cs::Component<ColReduce<Z, Dots>> {
public:
    static constexpr int max_work_group_size = kWg;
    struct InputsZ {
        cs::Input<DType::f32, 2> P;
        cs::Input<DType::f32, 1> z;
        cs::Input<DType::f32, 1> scale;
        cs::Input<DType::f32, 1> v1, v2;  ///< Dots only
    };
    struct InputsNoZ {
        cs::Input<DType::f32, 2> P;
        cs::Input<DType::f32, 1> scale;
        cs::Input<DType::f32, 1> v1, v2;  ///< Dots only
    };
    using Inputs = std::conditional_t<Z, InputsZ, InputsNoZ>;
    struct Outputs {
        cs::Output<DType::f32, 1> x;
        cs::Output<DType::f32, 2> dots;  ///< Dots only: [2, work-groups]
    };
    __device__ static void compute_forward(cs::GroupItem<1> it, const Inputs& in, const Outputs& out) {
        const std::uint32_t n = u32(in.P.shape()[1]), strips = u32(in.P.shape()[0]);
        const float         scale  = in.scale[std::size_t{0}];
        const std::uint32_t stride = u32(it.get_group_range(0) * it.get_local_range(0));
        const float*        P      = cs::raw::data(in.P);
        const std::size_t   ld     = in.P.stride()[0];
        if constexpr (Dots) {
            // one column per work-item (the launch covers n exactly once)
            const std::uint32_t j = u32(it.get_global_id(0));
            float               p1 = 0.0f, p2 = 0.0f;
            if (j < n) {
                float        acc = 0.0f;
                const float* pj  = P + j;
#pragma unroll 16
                for (std::uint32_t c = 0; c < strips; ++c) acc += pj[static_cast<std::size_t>(c) * ld];
                float xj;
                if constexpr (Z)
                    xj = fmaf(scale, acc, in.z[cs::id<1>{j}]);
                else
                    xj = scale * acc;
                out.x[cs::id<1>{j}] = xj;
                p1                  = in.v1[cs::id<1>{j}] * xj;
                p2                  = in.v2[cs::id<1>{j}] * xj;
            }
            const auto grp = it.get_group();
            p1             = cs::reduce_over_group(grp, p1, cs::plus<float>());
            p2             = cs::reduce_over_group(grp, p2, cs::plus<float>());
            if (it.get_local_id(0) == 0) {
                const std::size_t g = it.get_group(0);
                out.dots[cs::id<2>{0, g}] = p1;
                out.dots[cs::id<2>{1, g}] = p2;
            }
            return;
        }
        for (std::uint32_t j = u32(it.get_global_id(0)); j < n; j += stride) {
            float        acc = 0.0f;
            const float* pj  = P + j;
            // the sum stays in ascending c; unrolling keeps 64 independent loads in flight per work-item (the hardware's
            // limit of outstanding loads per wave is 63): a column of 1024 strips (16-row strips at n = 16384) is 16
            // rounds of loads
#pragma unroll 64
            for (std::uint32_t c = 0; c < strips; ++c) acc += pj[static_cast<std::size_t>(c) * ld];
            if constexpr (Z)
                out.x[cs::id<1>{j}] = fmaf(scale, acc, in.z[cs::id<1>{j}]);
            else
                out.x[cs::id<1>{j}] = scale * acc;
        }
    }

    Inputs  inputs;
    Outputs outputs;

    /// `wg`: work-items per work-group without the dots (kWg inside a fused group, whose stages share one shape; one
    /// wave as a launch of its own, which spreads the columns over every CU: the reduction is latency-bound)
    explicit ColReduce(int wg = kWg) : wg_{wg} {}

    void declare() override {
        this->declare_input(inputs.P, "P", cs::unspecified_shape<2>());
        if constexpr (Z) this->declare_input(inputs.z, "z", cs::unspecified_shape<1>());
        this->declare_input(inputs.scale, "scale", cs::unspecified_shape<1>());
        this->declare_output(outputs.x, "x", cs::unspecified_shape<1>());
        if constexpr (Dots) {
            this->declare_input(inputs.v1, "v1", cs::unspecified_shape<1>());
            this->declare_input(inputs.v2, "v2", cs::unspecified_shape<1>());
            this->declare_output(outputs.dots, "dots", cs::unspecified_shape<2>());
        }
    }
    void prepare_forward() override {
        const std::size_t n = this->shape_of(inputs.P)[1];
        this->set_shape(outputs.x, n);
        if constexpr (Dots) this->set_shape(outputs.dots, 2, round_up(n, kWg) / kWg);
    }
    std::unique_ptr<cs::LaunchDescription> define_forward() override {
        // one column per work-item; the dots' partials are per work-group of kWg
        const std::size_t range = inputs.P.shape()[1], wg = Dots ? kWg : static_cast<std::size_t>(wg_);
        return this->create_launch_config(cs::range_desc(cs::range<1>{round_up(range, wg)}), cs::wg_desc(cs::range<1>{wg}));
    }

private:
    int wg_;
};

class Gemv : public cs::Component<Gemv> {
public:
    static constexpr int max_work_group_size = kWg;
    struct Inputs {
        cs::Input<DType::f32, 2> B;
        cs::Input<DType::f32, 1> x;
        cs::Input<DType::f32, 1> scale;
    };
    struct Outputs {
        cs::Output<DType::f32, 1> y;
    };
    /// Rows per work-group: each row keeps the reference's order (work-item l: one fma chain over j = l, l + 256, ...; then
    /// the group's tree), and x[j] is loaded once for the work-group's rows instead of once per row.
    static constexpr int kRows = 4;
    __device__ static void compute_forward(cs::GroupItem<1> it, const Inputs& in, const Outputs& out) {
        const std::uint32_t n     = u32(in.B.shape()[0]);
        const float         scale = in.scale[std::size_t{0}];
        const std::uint32_t l = u32(it.get_local_id(0)), wg = u32(it.get_local_range(0));
        const std::uint32_t g = u32(it.get_group(0)), ng = u32(it.get_group_range(0));
        const std::size_t ldb = in.B.stride()[0];
        const float*      x   = cs::raw::data(in.x);
        for (std::uint32_t i0 = g * kRows; i0 < n; i0 += ng * kRows) {
            const float* row[kRows];
            float        acc[kRows];
#pragma unroll
            for (int r = 0; r < kRows; ++r) {
                const std::uint32_t i = i0 + static_cast<std::uint32_t>(r) < n ? i0 + static_cast<std::uint32_t>(r) : i0;
                row[r]                = cs::raw::data(in.B) + static_cast<std::size_t>(i) * ldb;
                acc[r]                = 0.0f;
            }
            // the loads of 4 steps of the rows in flight; each row's sum stays one fma chain over j = l, l + wg, ...
#pragma unroll 4
            for (std::uint32_t j = l; j < n; j += wg) {
                const float xj = x[j];
#pragma unroll
                for (int r = 0; r < kRows; ++r) acc[r] = fmaf(__builtin_nontemporal_load(row[r] + j), xj, acc[r]);  // streamed: 0.193 -> 0.177 ms
            }
#pragma unroll
            for (int r = 0; r < kRows; ++r) {
                const float sum = cs::reduce_over_group(it.get_group(), acc[r], cs::plus<float>());
                if (l == 0 && i0 + static_cast<std::uint32_t>(r) < n) out.y[cs::id<1>{i0 + static_cast<std::uint32_t>(r)}] = scale * sum;
            }
        }
    }

    Inputs  inputs;
    Outputs outputs;

    void declare() override {
        declare_input(inputs.B, "B", cs::unspecified_shape<2>());
        declare_input(inputs.x, "x", cs::unspecified_shape<1>());
        declare_input(inputs.scale, "scale", cs::unspecified_shape<1>());
        declare_output(outputs.y, "y", cs::unspecified_shape<1>());
    }
    void prepare_forward() override { set_shape(outputs.y, shape_of(inputs.B)[0]); }
    std::unique_ptr<cs::LaunchDescription> define_forward() override {
        const std::size_t range = (inputs.B.shape()[0] + kRows - 1) / kRows * kWg;  // a group per kRows rows
        return create_launch_config(cs::range_desc(cs::range<1>{range}), cs::wg_desc(cs::range<1>{kWg}));
    }
};

// =============================================================================
// Storage of A: 16-byte vector loads, widening to f32
// =============================================================================

template <DType S>
struct Storage;

template <>
struct Storage<DType::f32> {
    using T                    = float;
    static constexpr int kVec  = 4;
    using Raw                  = float __attribute__((ext_vector_type(4)));
    __device__ static void widen(const Raw& r, float s, float (&a)[kVec]) {
        (void)s;
        a[0] = r[0];
        a[1] = r[1];
        a[2] = r[2];
        a[3] = r[3];
    }
    __device__ static float widen1(const T& v, float s) {
        (void)s;
        return v;
    }
};

template <>
struct Storage<DType::bf16> {
    using T                   = cs::bf16;
    static constexpr int kVec = 8;
    using Raw                 = std::uint32_t __attribute__((ext_vector_type(4)));
    __device__ static void widen(const Raw& r, float s, float (&a)[kVec]) {
        (void)s;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            a[2 * q]     = __uint_as_float(r[q] << 16);
            a[2 * q + 1] = __uint_as_float(r[q] & 0xffff0000u);
        }
    }
    __device__ static float widen1(const T& v, float s) {
        (void)s;
        return v.to_float();
    }
};

template <>
struct Storage<DType::e4m3> {
    using T                   = cs::e4m3;
    static constexpr int kVec = 16;
    using Raw                 = std::uint32_t __attribute__((ext_vector_type(4)));
    __device__ static void widen(const Raw& r, float s, float (&a)[kVec]) {
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const auto lo = __builtin_amdgcn_cvt_pk_f32_fp8(static_cast<int>(r[q]), false);
            const auto hi = __builtin_amdgcn_cvt_pk_f32_fp8(static_cast<int>(r[q]), true);
            a[4 * q]      = lo[0] * s;
            a[4 * q + 1]  = lo[1] * s;
            a[4 * q + 2]  = hi[0] * s;
            a[4 * q + 3]  = hi[1] * s;
        }
    }
    __device__ static float widen1(const T& v, float s) {
        const auto p = __builtin_amdgcn_cvt_pk_f32_fp8(static_cast<int>(v.bits), false);
        return p[0] * s;
    }
};

/// A streaming 16-byte load (A is read once per pass: nontemporal, so it does not evict the reused vectors).
template <class Raw>
__device__ inline Raw load_stream(const void* p) {
    return __builtin_nontemporal_load(static_cast<const Raw*>(p));
}

/// N floats from an f32 vector port at element j (16-byte aligned when N % 4 == 0 and j % 4 == 0).
template <int N>
__device__ inline void load_f32(const float* p, float (&v)[N]) {
    static_assert(N % 4 == 0, "cospan: f32 vectors load 4 at a time");
    using F4 = float __attribute__((ext_vector_type(4)));
#pragma unroll
    for (int q = 0; q < N / 4; ++q) {
        const F4 t   = *reinterpret_cast<const F4*>(p + 4 * q);
        v[4 * q]     = t[0];
        v[4 * q + 1] = t[1];
        v[4 * q + 2] = t[2];
        v[4 * q + 3] = t[3];
    }
}

/// The order in which a row sweep of pass 2 takes its `count` full column steps: starting at step `first` and wrapping
/// around. Rows of A lie 4 n bytes apart, so sweeps that all start at column 0 and advance together keep every concurrent
/// row stream at the same column offset, on the same few HBM channels; each wave (or work-group) starts at a step of its
/// own instead (pass 2 at N = 16384: 157.5 -> 153.5 us).
struct Rotation {
    std::uint32_t          first = 0, count = 0;
    __device__ static Rotation of(std::uint32_t who, std::uint32_t count) { return Rotation{count ? who * 5u % count : 0u, count}; }
    __device__ std::uint32_t at(std::uint32_t t) const {
        const std::uint32_t s = first + t;
        return s >= count ? s - count : s;
    }
};

// =============================================================================
// Pass 1: P[c][j] = sum_{k<R} Â[cR + k][j] y[cR + k]
// =============================================================================

template <DType S>
class AhatTPartial : public cs::Component<AhatTPartial<S>> {
public:
    static constexpr int max_work_group_size = kWg;
    using St                                 = Storage<S>;
    /// elements of one 16-byte load
    static constexpr int kEl = St::kVec;
    /// 16-byte chunks per work-item: 16 f32 (64 bytes), 32 bf16 (64 bytes) or 32 e4m3 columns (32 bytes; 64 were no
    /// faster) of each row. With 16 columns per work-item whatever the storage, a work-group read 8 KiB (bf16) or 4 KiB
    /// (e4m3) of a row per row and pass 1 ran 7% and 14% slower at N = 32768; f32 gains nothing from 32 columns.
    static constexpr int kV = S == DType::e4m3 ? 2 : 4;
    /// columns of one chunk row across the work-group (one load instruction covers 64 * kEl contiguous columns per wave)
    static constexpr int kChunk = kWg * kEl;
    /// columns of one work-group
    static constexpr int kCols = kChunk * kV;
    /// Reduced-precision storage sums the stored values against s_i y_i (s_i = 1 for bf16) and adds the rank-2 terms once
    /// per column: P[c][j] = sum_i q_ij (s_i y_i) + v1_j sum_i u1_i y_i + v2_j sum_i u2_i y_i (one fma per element
    /// instead of three and a scale). f32 keeps the components' order (bit-exact).
    static constexpr bool kFactored = S != DType::f32;
    /// rows in flight per work-item: 16 loads of 16 bytes (f32 4 rows x 4 chunks, bf16 8 x 2, e4m3 16 x 1)
    static constexpr int kRowsInFlight = 16 / kV;

    struct Inputs {
        cs::Input<S, 2>          A;
        cs::Input<DType::f32, 1> row_scale;  ///< e4m3 storage only (extent 0 otherwise)
        cs::Input<DType::f32, 1> u1, v1, u2, v2, y;
        cs::Input<DType::u32, 1> R;
    };
    struct Outputs {
        cs::Output<DType::f32, 2> P;
    };

    /// group (c, cb): strip c of R rows, column block cb of kCols columns; work-item t owns the kV chunks of kEl columns
    /// at cb * kCols + q * kChunk + t * kEl (q < kV), interleaved so that every load instruction of a wave reads one
    /// contiguous run. Each column's sum over the strip is one fma chain in k order (the components' order).
    __device__ static void compute_forward(cs::GroupItem<2> it, const Inputs& in, const Outputs& out) {
        const std::uint32_t rows = u32(in.A.shape()[0]), n = u32(in.A.shape()[1]);
        const std::size_t   lda  = in.A.stride()[0];
        // R in a scalar register: the strip's row indices are then uniform and the per-row values (y, u1, u2, the row
        // scale) load through the scalar cache instead of as vector loads of one address (four per row next to one 16-byte
        // load of A for e4m3)
        const std::uint32_t R = static_cast<std::uint32_t>(__builtin_amdgcn_readfirstlane(static_cast<int>(in.R[std::size_t{0}])));
        const std::uint32_t c = u32(it.get_group(0)), cb = u32(it.get_group(1));
        const std::uint32_t jb = cb * kCols + u32(it.get_local_id(1)) * kEl;
        if constexpr (kFactored) {
            factored(it, in, out, R, c, jb);
            return;
        }
        if (jb >= n) return;
        const std::uint32_t i0 = c * R, i1 = (i0 + R < rows) ? i0 + R : rows;
        const auto*         A  = cs::raw::data(in.A);
        const float*        u1 = cs::raw::data(in.u1);
        const float*        u2 = cs::raw::data(in.u2);
        const float*        y  = cs::raw::data(in.y);
        const float*        sc = cs::raw::data(in.row_scale);
        const float*        V1 = cs::raw::data(in.v1);
        const float*        V2 = cs::raw::data(in.v2);

        // chunks wholly inside the row take the vector path; at most one chunk per work-item is partial
        bool          full[kV];
        std::uint32_t part_q = kV, part_m = 0;
        float         acc[kV][kEl], v1[kV][kEl], v2[kV][kEl];
#pragma unroll
        for (int q = 0; q < kV; ++q) {
            const std::uint32_t j = jb + static_cast<std::uint32_t>(q) * kChunk;
            full[q]               = j + kEl <= n;
            if (!full[q] && j < n) {
                part_q = static_cast<std::uint32_t>(q);
                part_m = n - j;
            }
#pragma unroll
            for (int e = 0; e < kEl; ++e) acc[q][e] = 0.0f;
            if (full[q]) {
                load_f32<kEl>(V1 + j, v1[q]);
                load_f32<kEl>(V2 + j, v2[q]);
            }
        }
        const typename St::T* a = A + static_cast<std::size_t>(i0) * lda + jb;
#pragma unroll 4
        for (std::uint32_t i = i0; i < i1; ++i, a += lda) {
            typename St::Raw r[kV];
#pragma unroll
            for (int q = 0; q < kV; ++q)
                if (full[q]) r[q] = load_stream<typename St::Raw>(a + q * kChunk);
            const float s = S == DType::e4m3 ? sc[i] : 1.0f, yi = y[i], a1 = u1[i], a2 = u2[i];
#pragma unroll
            for (int q = 0; q < kV; ++q) {
                if (!full[q]) continue;
                float av[kEl];
                St::widen(r[q], s, av);
#pragma unroll
                for (int e = 0; e < kEl; ++e) acc[q][e] = fmaf(rank2(av[e], a1, v1[q][e], a2, v2[q][e]), yi, acc[q][e]);
            }
        }
        float* p = &out.P[cs::id<2>{c, 0}];
#pragma unroll
        for (int q = 0; q < kV; ++q) {
            if (!full[q]) continue;
            using F4 = float __attribute__((ext_vector_type(4)));
            float* pq = p + jb + q * kChunk;
#pragma unroll
            for (int k4 = 0; k4 < kEl / 4; ++k4)
                reinterpret_cast<F4*>(pq)[k4] = F4{acc[q][4 * k4], acc[q][4 * k4 + 1], acc[q][4 * k4 + 2], acc[q][4 * k4 + 3]};
        }
        if (part_q < kV) {
            // the ragged end of the row: element by element, the same order
            const std::uint32_t j = jb + part_q * kChunk;
            float               pacc[kEl];
            for (std::uint32_t e = 0; e < part_m; ++e) pacc[e] = 0.0f;
            for (std::uint32_t i = i0; i < i1; ++i) {
                const float s = S == DType::e4m3 ? sc[i] : 1.0f, yi = y[i], a1 = u1[i], a2 = u2[i];
                for (std::uint32_t e = 0; e < part_m; ++e) {
                    const float av = St::widen1(A[static_cast<std::size_t>(i) * lda + j + e], s);
                    pacc[e]        = fmaf(rank2(av, a1, V1[j + e], a2, V2[j + e]), yi, pacc[e]);
                }
            }
            for (std::uint32_t e = 0; e < part_m; ++e) p[j + e] = pacc[e];
        }
    }

    /// Reduced-precision storage: the strip's s_i y_i staged in LDS by the work-group (one coalesced load of each per-row
    /// value instead of a load of one address per work-item and row), sum u1_i y_i and sum u2_i y_i reduced over the group
    /// in a fixed order, then P[c][j] = sum_i q_ij (s_i y_i) + v1_j sum u1_i y_i + v2_j sum u2_i y_i. Every work-item stays
    /// to the end (the barriers), those past the row only skip the loads and stores.
    __device__ static void factored(cs::GroupItem<2> it, const Inputs& in, const Outputs& out, std::uint32_t R, std::uint32_t c,
                                    std::uint32_t jb) {
        const std::uint32_t rows = u32(in.A.shape()[0]), n = u32(in.A.shape()[1]);
        const std::size_t   lda  = in.A.stride()[0];
        const std::uint32_t i0 = c * R, i1 = (i0 + R < rows) ? i0 + R : rows;
        const std::uint32_t tid = u32(it.get_local_id(1));
        const auto*         A   = cs::raw::data(in.A);
        const float *       y = cs::raw::data(in.y), *u1 = cs::raw::data(in.u1), *u2 = cs::raw::data(in.u2);
        const float *       sc = cs::raw::data(in.row_scale), *V1 = cs::raw::data(in.v1), *V2 = cs::raw::data(in.v2);
        __shared__ float    ys[kWg];

        bool          full[kV];
        std::uint32_t part_q = kV, part_m = 0;
        float         acc[kV][kEl];
#pragma unroll
        for (int q = 0; q < kV; ++q) {
            const std::uint32_t j = jb + static_cast<std::uint32_t>(q) * kChunk;
            full[q]               = j + kEl <= n;
            if (!full[q] && j < n) {
                part_q = static_cast<std::uint32_t>(q);
                part_m = n - j;
            }
#pragma unroll
            for (int e = 0; e < kEl; ++e) acc[q][e] = 0.0f;
        }
        float pacc[kEl];
#pragma unroll
        for (int e = 0; e < kEl; ++e) pacc[e] = 0.0f;
        float t1 = 0.0f, t2 = 0.0f;  // this work-item's rows of the strip: u1_i y_i, u2_i y_i
        for (std::uint32_t b0 = i0; b0 < i1; b0 += kWg) {
            const std::uint32_t nb = i1 - b0 < static_cast<std::uint32_t>(kWg) ? i1 - b0 : static_cast<std::uint32_t>(kWg);
            if (tid < nb) {
                const std::uint32_t i  = b0 + tid;
                const float         yi = y[i];
                ys[tid]                = (S == DType::e4m3 ? sc[i] : 1.0f) * yi;
                t1                     = fmaf(u1[i], yi, t1);
                t2                     = fmaf(u2[i], yi, t2);
            }
            __syncthreads();
            const typename St::T* a = A + static_cast<std::size_t>(b0) * lda + jb;
#pragma unroll kRowsInFlight
            for (std::uint32_t k = 0; k < nb; ++k, a += lda) {
                typename St::Raw r[kV];
#pragma unroll
                for (int q = 0; q < kV; ++q)
                    if (full[q]) r[q] = load_stream<typename St::Raw>(a + q * kChunk);
                const float yk = ys[k];
#pragma unroll
                for (int q = 0; q < kV; ++q) {
                    if (!full[q]) continue;
                    float av[kEl];
                    St::widen(r[q], 1.0f, av);
#pragma unroll
                    for (int e = 0; e < kEl; ++e) acc[q][e] = fmaf(av[e], yk, acc[q][e]);
                }
            }
            if (part_q < kV) {  // the ragged end of the row: element by element
                const std::uint32_t j = jb + part_q * kChunk;
                for (std::uint32_t k = 0; k < nb; ++k)
                    for (std::uint32_t e = 0; e < part_m; ++e)
                        pacc[e] = fmaf(St::widen1(A[static_cast<std::size_t>(b0 + k) * lda + j + e], 1.0f), ys[k], pacc[e]);
            }
            __syncthreads();
        }
        t1 = cs::reduce_over_group(it.get_group(), t1, cs::plus<float>());
        t2 = cs::reduce_over_group(it.get_group(), t2, cs::plus<float>());
        if (jb >= n) return;
        float* p = &out.P[cs::id<2>{c, 0}];
#pragma unroll
        for (int q = 0; q < kV; ++q) {
            if (!full[q]) continue;
            const std::uint32_t j = jb + static_cast<std::uint32_t>(q) * kChunk;
            float               v1[kEl], v2[kEl];
            load_f32<kEl>(V1 + j, v1);
            load_f32<kEl>(V2 + j, v2);
            using F4  = float __attribute__((ext_vector_type(4)));
            float* pq = p + j;
#pragma unroll
            for (int k4 = 0; k4 < kEl / 4; ++k4) {
                F4 o;
#pragma unroll
                for (int e = 0; e < 4; ++e) o[e] = fmaf(v2[4 * k4 + e], t2, fmaf(v1[4 * k4 + e], t1, acc[q][4 * k4 + e]));
                reinterpret_cast<F4*>(pq)[k4] = o;
            }
        }
        if (part_q < kV) {
            const std::uint32_t j = jb + part_q * kChunk;
            for (std::uint32_t e = 0; e < part_m; ++e) p[j + e] = fmaf(V2[j + e], t2, fmaf(V1[j + e], t1, pacc[e]));
        }
    }

    Inputs  inputs;
    Outputs outputs;

    explicit AhatTPartial(int R) : R_{R} {}
    void declare() override {
        this->declare_input(inputs.A, "A", cs::unspecified_shape<2>());
        this->declare_input(inputs.row_scale, "row_scale", cs::unspecified_shape<1>());
        this->declare_input(inputs.u1, "u1", cs::unspecified_shape<1>());
        this->declare_input(inputs.v1, "v1", cs::unspecified_shape<1>());
        this->declare_input(inputs.u2, "u2", cs::unspecified_shape<1>());
        this->declare_input(inputs.v2, "v2", cs::unspecified_shape<1>());
        this->declare_input(inputs.y, "y", cs::unspecified_shape<1>());
        this->declare_input(inputs.R, "R", 1);
        this->init_input("R", cs::init::fill(u32(static_cast<std::size_t>(R_))));
        this->declare_output(outputs.P, "P", cs::unspecified_shape<2>());
    }
    void prepare_forward() override {
        const cs::range<2> a = this->shape_of(inputs.A);
        this->set_shape(outputs.P, (a[0] + R_ - 1) / R_, a[1]);
    }
    /// a work-group per (strip of R rows, block of kCols columns)
    std::unique_ptr<cs::LaunchDescription> define_forward() override {
        const cs::range<2> a = inputs.A.shape();
        const std::size_t  strips = (a[0] + R_ - 1) / R_, blocks = (a[1] + kCols - 1) / kCols;
        return this->create_launch_config(cs::range_desc(cs::range<2>{strips, blocks * kWg}), cs::wg_desc(cs::range<2>{1, kWg}));
    }

private:
    int R_;
};

// =============================================================================
// Pass 2: w[i] = alpha * sum_j Â[i][j] x[j], RT rows per work-group
// =============================================================================

template <DType S, int RT>
class AhatGemv : public cs::Component<AhatGemv<S, RT>> {
public:
    static constexpr int max_work_group_size = kWg;
    using St                                 = Storage<S>;
    static constexpr int kVec                = St::kVec;
    static constexpr int kRows               = RT;
    static constexpr bool kCompensated       = S != DType::f32;
    /// Reduced-precision storage: w_i = alpha (s_i sum_j q_ij x_j + u1_i (v1 . x) + u2_i (v2 . x)) (s_i = 1 for bf16)
    /// with the two dots from the column reduction: one fma per stored element instead of three and a scale, and no v1,
    /// v2 re-reads per row step, which at one or two bytes per element outweighed A (bf16 pass 2 with wave-owned rows
    /// 0.353 -> 0.324 ms at N = 32768)
    static constexpr bool kFactored = S != DType::f32;
    /// column chunks in flight per work-item for reduced-precision storage in the shared layout: 16 loads of 16 bytes of A
    /// (RT rows each); f32 as tuned
    static constexpr int kChunksInFlight = (kFactored && RT < 16) ? 16 / RT : 1;

    struct Inputs {
        cs::Input<S, 2>          A;
        cs::Input<DType::f32, 1> row_scale;
        cs::Input<DType::f32, 1> u1, v1, u2, v2, x;
        cs::Input<DType::f32, 1> scale;
        cs::Input<DType::f32, 2> dots;  ///< factored only: the column reduction's partial v1 . x, v2 . x
    };
    struct Outputs {
        cs::Output<DType::f32, 1> w;
    };

    /// RT >= 16: the wave layout (wave v of the work-group owns rows RT g + v RT/4 .. + RT/4 and sweeps all their columns,
    /// 64 lanes x kVec columns per step, then reduces each row within the wave: no LDS, no barrier). Measured faster than
    /// the block layout below for the f32 headline (155.5 against 161 us per 1 GiB pass at 16 rows).
    static constexpr int kWaveRows = RT >= 16 ? RT / 4 : 0;
    /// f32 storage in the wave layout: the staged step (wave_rows)
    static constexpr bool kStaged = S == DType::f32 && kWaveRows > 0;

    __device__ static void compute_forward(cs::GroupItem<1> it, const Inputs& in, const Outputs& out) {
        if constexpr (kWaveRows > 0) {
            wave_rows(it, in, out);
            return;
        }
        const std::uint32_t rows = u32(in.A.shape()[0]), n = u32(in.A.shape()[1]);
        const std::size_t   lda  = in.A.stride()[0];
        const std::uint32_t r0   = u32(it.get_group(0)) * RT;
        if (r0 >= rows) return;  // uniform over the work-group
        const std::uint32_t nr  = rows - r0 < static_cast<std::uint32_t>(RT) ? rows - r0 : static_cast<std::uint32_t>(RT);
        const std::uint32_t tid = u32(it.get_local_id(0));
        const auto*         A   = cs::raw::data(in.A);
        const float*        sc  = cs::raw::data(in.row_scale);
        const float*        xv  = cs::raw::data(in.x);
        const float*        v1p = cs::raw::data(in.v1);
        const float*        v2p = cs::raw::data(in.v2);

        float a1[RT], a2[RT], s[RT], acc[RT], comp[RT];
#pragma unroll
        for (int r = 0; r < RT; ++r) {
            const std::uint32_t i = r0 + (static_cast<std::uint32_t>(r) < nr ? static_cast<std::uint32_t>(r) : 0u);
            a1[r]                 = in.u1[cs::id<1>{i}];
            a2[r]                 = in.u2[cs::id<1>{i}];
            s[r]                  = S == DType::e4m3 ? sc[i] : 1.0f;
            acc[r]                = 0.0f;
            comp[r]               = 0.0f;
        }

        // Neumaier: acc + comp is the running sum of the partials
        const auto add_comp = [](float& acc_, float& comp_, float part) {
            const float t = acc_ + part;
            comp_ += fabsf(acc_) >= fabsf(part) ? (acc_ - t) + part : (part - t) + acc_;
            acc_ = t;
        };
        // factored: v1 . x and v2 . x, the column reduction's partials summed in a fixed order (each work-item a stride of
        // them, then the group's tree): the chunk loop then reads only x next to A
        float D1 = 0.0f, D2 = 0.0f;
        if constexpr (kFactored) {
            const std::uint32_t g = u32(in.dots.shape()[1]);
            const float*        dp = cs::raw::data(in.dots);
            float               p1 = 0.0f, p2 = 0.0f;
            for (std::uint32_t k = tid; k < g; k += kWg) {
                p1 += dp[k];
                p2 += dp[g + k];
            }
            D1 = cs::reduce_over_group(it.get_group(), p1, cs::plus<float>());
            D2 = cs::reduce_over_group(it.get_group(), p2, cs::plus<float>());
        }
        // the full column steps in this work-group's rotated order (chunks unrolled so that 16 loads of 16 bytes of A are
        // in flight per work-item); then the work-item's columns of a partial last step
        const std::uint32_t step  = kWg * kVec;
        const std::uint32_t nfull = n / step;
        const Rotation      rot   = Rotation::of(r0 / RT, nfull);
#pragma unroll kChunksInFlight
        for (std::uint32_t t = 0; t < nfull; ++t) {
            const std::uint32_t j0 = rot.at(t) * step + tid * kVec;
            float               v1[kVec], v2[kVec], x[kVec];
            if constexpr (!kFactored) {
                load_f32<kVec>(v1p + j0, v1);
                load_f32<kVec>(v2p + j0, v2);
            }
            load_f32<kVec>(xv + j0, x);
            typename St::Raw raw[RT];
#pragma unroll
            for (int r = 0; r < RT; ++r) {
                const std::uint32_t rr = static_cast<std::uint32_t>(r) < nr ? static_cast<std::uint32_t>(r) : 0u;
                raw[r]                 = load_stream<typename St::Raw>(A + static_cast<std::size_t>(r0 + rr) * lda + j0);
            }
            if constexpr (kFactored) {
#pragma unroll
                for (int r = 0; r < RT; ++r) {
                    float av[kVec];
                    St::widen(raw[r], 1.0f, av);
                    float part = 0.0f;
#pragma unroll
                    for (int v = 0; v < kVec; ++v) part = fmaf(av[v], x[v], part);
                    add_comp(acc[r], comp[r], part);
                }
                continue;
            }
#pragma unroll
            for (int r = 0; r < RT; ++r) {
                float av[kVec];
                St::widen(raw[r], s[r], av);
                if constexpr (kCompensated) {
                    float part = 0.0f;
#pragma unroll
                    for (int v = 0; v < kVec; ++v) part = fmaf(rank2(av[v], a1[r], v1[v], a2[r], v2[v]), x[v], part);
                    // Neumaier: acc + comp is the running sum of the chunk partials
                    const float t = acc[r] + part;
                    comp[r] += fabsf(acc[r]) >= fabsf(part) ? (acc[r] - t) + part : (part - t) + acc[r];
                    acc[r] = t;
                } else {
#pragma unroll
                    for (int v = 0; v < kVec; ++v) acc[r] = fmaf(rank2(av[v], a1[r], v1[v], a2[r], v2[v]), x[v], acc[r]);
                }
            }
        }
        const std::uint32_t j0 = nfull * step + tid * kVec;
        const std::uint32_t j1 = j0 + kVec < n ? j0 + kVec : n;
        if (j0 < n && kFactored) {  // the partial last step: this work-item's columns in it, one at a time
            for (int r = 0; r < RT; ++r) {
                const std::uint32_t rr   = static_cast<std::uint32_t>(r) < nr ? static_cast<std::uint32_t>(r) : 0u;
                float               part = 0.0f;
                for (std::uint32_t j = j0; j < j1; ++j) part = fmaf(St::widen1(A[static_cast<std::size_t>(r0 + rr) * lda + j], 1.0f), xv[j], part);
                add_comp(acc[r], comp[r], part);
            }
        } else if (j0 < n) {  // the partial last step: this work-item's columns in it, one at a time
            for (int r = 0; r < RT; ++r) {
                const std::uint32_t rr   = static_cast<std::uint32_t>(r) < nr ? static_cast<std::uint32_t>(r) : 0u;
                float               part = 0.0f;
                for (std::uint32_t j = j0; j < j1; ++j) {
                    const float av = St::widen1(A[static_cast<std::size_t>(r0 + rr) * lda + j], s[r]);
                    part           = fmaf(rank2(av, a1[r], v1p[j], a2[r], v2p[j]), xv[j], part);
                }
                if constexpr (kCompensated) {
                    const float t = acc[r] + part;
                    comp[r] += fabsf(acc[r]) >= fabsf(part) ? (acc[r] - t) + part : (part - t) + acc[r];
                    acc[r] = t;
                } else {
                    acc[r] += part;
                }
            }
        }

        // per row: wave butterfly, then the pairwise sum of the four wave results
        constexpr int    kSums = RT;
        __shared__ float wave_sums[kWg / cs::wave::kSize][kSums];
        const int        wv = static_cast<int>(tid) / cs::wave::kSize;
#pragma unroll
        for (int r = 0; r < RT; ++r) {
            float v = kCompensated ? acc[r] + comp[r] : acc[r];
            v       = cs::wave::reduce(v, cs::plus<float>{});
            if (cs::wave::lane_id() == 0) wave_sums[wv][r] = v;
        }
        __syncthreads();
        if (tid < nr) {
            const float alpha = in.scale[std::size_t{0}];
            const auto  total = [&](int k) { return (wave_sums[0][k] + wave_sums[1][k]) + (wave_sums[2][k] + wave_sums[3][k]); };
            const float sum   = total(static_cast<int>(tid));
            if constexpr (kFactored) {
                const std::uint32_t i  = r0 + tid;
                const float         si = S == DType::e4m3 ? sc[i] : 1.0f;
                out.w[cs::id<1>{i}]    = alpha * fmaf(in.u2[cs::id<1>{i}], D2, fmaf(in.u1[cs::id<1>{i}], D1, si * sum));
            } else {
                out.w[cs::id<1>{r0 + tid}] = alpha * sum;
            }
        }
    }

    /// The wave layout: each wave keeps kWaveRows running dots over its own rows, lane l takes the kVec columns at
    /// kVec (l + 64 s); each row's dot is one chain per lane (compensated for reduced-precision storage) and a wave
    /// butterfly. The factored terms of reduced-precision storage (v1 . x, v2 . x) are reduced over the whole work-group
    /// first, before any wave leaves.
    __device__ static void wave_rows(cs::GroupItem<1> it, const Inputs& in, const Outputs& out) {
        constexpr int       W    = kWaveRows;
        const std::uint32_t rows = u32(in.A.shape()[0]), n = u32(in.A.shape()[1]);
        const std::size_t   lda  = in.A.stride()[0];
        const std::uint32_t tid  = u32(it.get_local_id(0));
        const std::uint32_t lane = tid % static_cast<std::uint32_t>(cs::wave::kSize), wv = tid / static_cast<std::uint32_t>(cs::wave::kSize);
        const std::uint32_t r0   = u32(it.get_group(0)) * RT + wv * static_cast<std::uint32_t>(W);
        const auto*         A    = cs::raw::data(in.A);
        const float*        sc   = cs::raw::data(in.row_scale);
        const float*        xv   = cs::raw::data(in.x);
        const float*        v1p  = cs::raw::data(in.v1);
        const float*        v2p  = cs::raw::data(in.v2);
        float               D1 = 0.0f, D2 = 0.0f;
        if constexpr (kFactored) {  // every wave takes part in the group reduction, so no wave leaves before it
            const std::uint32_t g  = u32(in.dots.shape()[1]);
            const float*        dp = cs::raw::data(in.dots);
            float               p1 = 0.0f, p2 = 0.0f;
            for (std::uint32_t k = tid; k < g; k += kWg) {
                p1 += dp[k];
                p2 += dp[g + k];
            }
            D1 = cs::reduce_over_group(it.get_group(), p1, cs::plus<float>());
            D2 = cs::reduce_over_group(it.get_group(), p2, cs::plus<float>());
        }
        if (r0 >= rows) return;  // uniform over the wave; no barrier follows
        const std::uint32_t nr = rows - r0 < static_cast<std::uint32_t>(W) ? rows - r0 : static_cast<std::uint32_t>(W);
        float               a1[W], a2[W], s[W], acc[W], comp[W];
#pragma unroll
        for (int r = 0; r < W; ++r) {
            const std::uint32_t i = r0 + (static_cast<std::uint32_t>(r) < nr ? static_cast<std::uint32_t>(r) : 0u);
            a1[r]                 = in.u1[cs::id<1>{i}];
            a2[r]                 = in.u2[cs::id<1>{i}];
            s[r]                  = S == DType::e4m3 ? sc[i] : 1.0f;
            acc[r]                = 0.0f;
            comp[r]               = 0.0f;
        }
        const auto add = [](float& acc_, float& comp_, float part) {
            if constexpr (kCompensated) {  // Neumaier: acc + comp is the running sum of the chunk partials
                const float t = acc_ + part;
                comp_ += fabsf(acc_) >= fabsf(part) ? (acc_ - t) + part : (part - t) + acc_;
                acc_ = t;
            } else {
                acc_ += part;
            }
        };
        // the full column steps in this wave's rotated order; then the lane's columns of a partial last step
        const std::uint32_t step  = static_cast<std::uint32_t>(cs::wave::kSize) * kVec;
        std::uint32_t       j0    = lane * kVec;
        const std::uint32_t nfull = n / step;
        const std::uint32_t s0    = Rotation::of(r0 / static_cast<std::uint32_t>(W), nfull).first;
        for (std::uint32_t t = 0; t < nfull && kStaged; ++t) {
            // f32: v1 and the rows' chunks are issued first and fma(u1, v1, A) is taken row by row as the rows arrive;
            // then v2 and x, then the rest of the rank-2 term and the dot (the same operations in the same order per
            // element as the body below). This schedule measured 3% faster than the compiler's own, which waited for
            // the rows before issuing v1, v2 and x one round trip at a time.
            std::uint32_t sidx = s0 + t;
            if (sidx >= nfull) sidx -= nfull;
            j0 = sidx * step + lane * kVec;
            float            v1[kVec], v2[kVec], x[kVec], b[W][kVec];
            typename St::Raw raw[W];
            load_f32<kVec>(v1p + j0, v1);
#pragma unroll
            for (int r = 0; r < W; ++r) {
                const std::uint32_t rr = static_cast<std::uint32_t>(r) < nr ? static_cast<std::uint32_t>(r) : 0u;
                raw[r]                 = load_stream<typename St::Raw>(A + static_cast<std::size_t>(r0 + rr) * lda + j0);
            }
            __builtin_amdgcn_sched_barrier(0);
#pragma unroll
            for (int r = 0; r < W; ++r) {
                float av[kVec];
                St::widen(raw[r], 1.0f, av);
#pragma unroll
                for (int v = 0; v < kVec; ++v) b[r][v] = fmaf(a1[r], v1[v], av[v]);
            }
            __builtin_amdgcn_sched_barrier(0);
            load_f32<kVec>(v2p + j0, v2);
            load_f32<kVec>(xv + j0, x);
            __builtin_amdgcn_sched_barrier(0);
#pragma unroll
            for (int r = 0; r < W; ++r)
#pragma unroll
                for (int v = 0; v < kVec; ++v) acc[r] = fmaf(fmaf(a2[r], v2[v], b[r][v]), x[v], acc[r]);
        }
        for (std::uint32_t t = 0; t < nfull && !kStaged; ++t) {
            std::uint32_t sidx = s0 + t;
            if (sidx >= nfull) sidx -= nfull;
            j0 = sidx * step + lane * kVec;
            float x[kVec], v1[kVec], v2[kVec];
            load_f32<kVec>(xv + j0, x);
            if constexpr (!kFactored) {
                load_f32<kVec>(v1p + j0, v1);
                load_f32<kVec>(v2p + j0, v2);
            }
            typename St::Raw raw[W];
#pragma unroll
            for (int r = 0; r < W; ++r) {
                const std::uint32_t rr = static_cast<std::uint32_t>(r) < nr ? static_cast<std::uint32_t>(r) : 0u;
                raw[r]                 = load_stream<typename St::Raw>(A + static_cast<std::size_t>(r0 + rr) * lda + j0);
            }
#pragma unroll
            for (int r = 0; r < W; ++r) {
                float av[kVec];
                St::widen(raw[r], kFactored ? 1.0f : s[r], av);
                if constexpr (kFactored) {
                    float part = 0.0f;
#pragma unroll
                    for (int v = 0; v < kVec; ++v) part = fmaf(av[v], x[v], part);
                    add(acc[r], comp[r], part);
                } else if constexpr (kCompensated) {
                    float part = 0.0f;
#pragma unroll
                    for (int v = 0; v < kVec; ++v) part = fmaf(rank2(av[v], a1[r], v1[v], a2[r], v2[v]), x[v], part);
                    add(acc[r], comp[r], part);
                } else {
#pragma unroll
                    for (int v = 0; v < kVec; ++v) acc[r] = fmaf(rank2(av[v], a1[r], v1[v], a2[r], v2[v]), x[v], acc[r]);
                }
            }
        }
        j0 = nfull * step + lane * kVec;
        if (j0 < n) {  // the partial last step: this lane's columns in it, one at a time
            const std::uint32_t j1 = j0 + kVec < n ? j0 + kVec : n;
            for (int r = 0; r < W; ++r) {
                const std::uint32_t rr   = static_cast<std::uint32_t>(r) < nr ? static_cast<std::uint32_t>(r) : 0u;
                float               part = 0.0f;
                for (std::uint32_t j = j0; j < j1; ++j) {
                    const float av = St::widen1(A[static_cast<std::size_t>(r0 + rr) * lda + j], kFactored ? 1.0f : s[r]);
                    part           = kFactored ? fmaf(av, xv[j], part) : fmaf(rank2(av, a1[r], v1p[j], a2[r], v2p[j]), xv[j], part);
                }
                add(acc[r], comp[r], part);
            }
        }
        const float alpha = in.scale[std::size_t{0}];
#pragma unroll
        for (int r = 0; r < W; ++r) {
            const float sum = cs::wave::reduce(kCompensated ? acc[r] + comp[r] : acc[r], cs::plus<float>{});
            if (lane == 0 && static_cast<std::uint32_t>(r) < nr) {
                const std::uint32_t i = r0 + static_cast<std::uint32_t>(r);
                if constexpr (kFactored)
                    out.w[cs::id<1>{i}] = alpha * fmaf(in.u2[cs::id<1>{i}], D2, fmaf(in.u1[cs::id<1>{i}], D1, s[r] * sum));
                else
                    out.w[cs::id<1>{i}] = alpha * sum;
            }
        }
    }

    Inputs  inputs;
    Outputs outputs;

    void declare() override {
        this->declare_input(inputs.A, "A", cs::unspecified_shape<2>());
        this->declare_input(inputs.row_scale, "row_scale", cs::unspecified_shape<1>());
        this->declare_input(inputs.u1, "u1", cs::unspecified_shape<1>());
        this->declare_input(inputs.v1, "v1", cs::unspecified_shape<1>());
        this->declare_input(inputs.u2, "u2", cs::unspecified_shape<1>());
        this->declare_input(inputs.v2, "v2", cs::unspecified_shape<1>());
        this->declare_input(inputs.x, "x", cs::unspecified_shape<1>());
        this->declare_input(inputs.scale, "scale", cs::unspecified_shape<1>());
        if constexpr (kFactored) this->declare_input(inputs.dots, "dots", cs::unspecified_shape<2>());
        this->declare_output(outputs.w, "w", cs::unspecified_shape<1>());
    }
    void prepare_forward() override { this->set_shape(outputs.w, this->shape_of(inputs.A)[0]); }
    /// a work-group of kWg work-items per RT rows (the body needs exactly kWg: four waves)
    std::unique_ptr<cs::LaunchDescription> define_forward() override {
        const std::size_t groups = (inputs.A.shape()[0] + RT - 1) / RT;
        return this->create_launch_config(cs::range_desc(cs::range<1>{groups * kWg}), cs::wg_desc(cs::range<1>{kWg}));
    }
};

CS_COMPONENTS_END

}  // namespace gemver
