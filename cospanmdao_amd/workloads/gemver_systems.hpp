#pragma once

// =============================================================================
// cospanmdao_amd/workloads/gemver_systems.hpp — GEMVER as a System, in the reference's fusion styles and this build's
// =============================================================================
//
// Ports (promoted or tied, so every style exposes one set of names, as examples/gemver-chain/systems.hpp:7-16 does):
//   inputs   A u1 v1 u2 v2 y z, row_scale (the e4m3 storage's per-row scale), alpha, beta
//   outputs  x w, and B (the styles that materialise it), P (the strip partials)
//
//   unfused    Rank2Update, GemvTPartial, ColReduce, Gemv: four launches (reference systems.hpp:112-122)
//   fused      fuse_components(connects, segment<range_mode::whole>(r2u, gtp)), cred, gemv: three launches (:105-111)
//   naive      fuse_components(segment<whole>(r2u, gtp), global_barrier(), cred, global_barrier(), gemv): one fused group (:82-104)
//   twopass    AhatTPartial, ColReduce, AhatGemv: Â never stored, A read twice: three launches
//   twopass_1  the same three stages as one fused group with global barriers
//
// The reference ties one `scale` into cred and gemv (alpha = beta = 1/sqrt(n), systems.hpp:79,136-137); this System takes
// the two GEMVER scalars separately (x = beta Â^T y + z, w = alpha Â x).

#include <cstdint>
#include <string>

#include "naga/cospan/cospan.hpp"
#include "gemver_components.hpp"

namespace gemver {

enum class Style { unfused = 0, fused = 1, naive = 2, twopass = 3, twopass_1 = 4 };

/// The inputs' initial values: the reference's hash generator on the device (chain_data.hpp:34-72, seeds A = 1 .. z = 7
/// for base seed 0), A stored as f32, bf16 or e4m3 with its per-row scale (quantised on the device from the f32 values).
struct InputSeeds {
    std::uint32_t base = 0;
    std::uint32_t of(std::uint32_t id) const { return base * 16u + id; }
};

inline cs::InitSpec hash_vector(std::uint32_t seed, std::size_t n) {
    return cs::init::custom<DType::f32, 1>(
        [seed, n](cs_hip_stream s, cs::Initializer<DType::f32, 1> v) {
            cs::detail::hip_check(cs_hip_fill_hash_f32(seed, cs::raw::data(v), 1, n, n, s), "hash initializer");
        },
        cs::Shape{n});
}

/// A (n x n) from the hash, stored as S; e4m3 also needs the row scales (hash_scales below).
template <DType S>
cs::InitSpec hash_matrix(std::uint32_t seed, std::size_t n) {
    return cs::init::custom<S, 2>(
        [seed, n](cs_hip_stream s, cs::Initializer<S, 2> a) {
            if constexpr (S == DType::f32) {
                cs::detail::hip_check(cs_hip_fill_hash_f32(seed, cs::raw::data(a), n, n, n, s), "hash initializer");
            } else {
                void* tmp = nullptr;
                cs::detail::hip_check(cs_hip_malloc(&tmp, n * n * 4 + n * 4), "initializer scratch");
                float* f = static_cast<float*>(tmp);
                cs::detail::hip_check(cs_hip_fill_hash_f32(seed, f, n, n, n, s), "hash initializer");
                if constexpr (S == DType::bf16)
                    cs::detail::hip_check(cs_hip_quantize_bf16(f, reinterpret_cast<std::uint16_t*>(cs::raw::data(a)), n, n, n, n, s), "quantize");
                else
                    cs::detail::hip_check(cs_hip_quantize_e4m3_rows(f, reinterpret_cast<std::uint8_t*>(cs::raw::data(a)), f + n * n, n, n, n, n, s),
                                          "quantize");
                cs::detail::hip_check(cs_hip_stream_synchronize(s), "initializer");
                cs::detail::hip_check(cs_hip_free(tmp), "initializer scratch");
            }
        },
        cs::Shape{n, n});
}

/// The per-row scales of the e4m3 storage of hash_matrix (the same deterministic quantiser, scales kept).
inline cs::InitSpec hash_scales(std::uint32_t seed, std::size_t n) {
    return cs::init::custom<DType::f32, 1>(
        [seed, n](cs_hip_stream s, cs::Initializer<DType::f32, 1> sc) {
            void* tmp = nullptr;
            cs::detail::hip_check(cs_hip_malloc(&tmp, n * n * 5), "initializer scratch");
            float* f = static_cast<float*>(tmp);
            cs::detail::hip_check(cs_hip_fill_hash_f32(seed, f, n, n, n, s), "hash initializer");
            cs::detail::hip_check(
                cs_hip_quantize_e4m3_rows(f, reinterpret_cast<std::uint8_t*>(f + n * n), cs::raw::data(sc), n, n, n, n, s), "quantize");
            cs::detail::hip_check(cs_hip_stream_synchronize(s), "initializer");
            cs::detail::hip_check(cs_hip_free(tmp), "initializer scratch");
        },
        cs::Shape{n});
}

template <DType S = DType::f32, int RT = 8>
class Gemver : public cs::System {
public:
    Gemver(Style style, int n, int R, float alpha, float beta, InputSeeds seeds)
        : style_{style}, n_{n}, R_{R}, alpha_{alpha}, beta_{beta}, seeds_{seeds} {}

    void setup() override {
        const std::size_t n = static_cast<std::size_t>(n_), strips = (n + R_ - 1) / R_;
        std::string       pre;
        if (style_ == Style::twopass || style_ == Style::twopass_1) {
            // reduced-precision storage: the column reduction also gives v1 . x and v2 . x to the factored pass 2
            constexpr bool dots = AhatGemv<S, RT>::kFactored;
            using Cred          = ColReduce<true, dots>;
            if (style_ == Style::twopass_1) {
                if constexpr (dots)
                    add_subsystem(cs::fuse_components("fw", cs::connects_desc({{"p1.P", "cred.P"}, {"cred.x", "p2.x"}, {"cred.dots", "p2.dots"}}),
                                                      cs::subsystem_desc<AhatTPartial<S>>("p1", R_), cs::global_barrier(),
                                                      cs::subsystem_desc<Cred>("cred"), cs::global_barrier(),
                                                      cs::subsystem_desc<AhatGemv<S, RT>>("p2")));
                else
                    add_subsystem(cs::fuse_components("fw", cs::connects_desc({{"p1.P", "cred.P"}, {"cred.x", "p2.x"}}),
                                                      cs::subsystem_desc<AhatTPartial<S>>("p1", R_), cs::global_barrier(),
                                                      cs::subsystem_desc<Cred>("cred"), cs::global_barrier(),
                                                      cs::subsystem_desc<AhatGemv<S, RT>>("p2")));
                pre = "fw.";
            } else {
                add_subsystem(cs::subsystem_desc<AhatTPartial<S>>("p1", R_));
                add_subsystem(cs::subsystem_desc<Cred>("cred", cs::wave::kSize));  // a launch of its own: one-wave work-groups
                add_subsystem(cs::subsystem_desc<AhatGemv<S, RT>>("p2"));
                connect("p1.P", "cred.P");
                connect("cred.x", "p2.x");
                if constexpr (dots) connect("cred.dots", "p2.dots");
            }
            // both passes read the same A and rank-2 vectors: one variable each
            tie_inputs<S, 2>({pre + "p1.A", pre + "p2.A"}, "A");
            tie_inputs<DType::f32, 1>({pre + "p1.row_scale", pre + "p2.row_scale"}, "row_scale");
            tie_inputs<DType::f32, 1>({pre + "p1.u1", pre + "p2.u1"}, "u1");
            if constexpr (dots) {
                tie_inputs<DType::f32, 1>({pre + "p1.v1", pre + "p2.v1", pre + "cred.v1"}, "v1");
                tie_inputs<DType::f32, 1>({pre + "p1.v2", pre + "p2.v2", pre + "cred.v2"}, "v2");
            } else {
                tie_inputs<DType::f32, 1>({pre + "p1.v1", pre + "p2.v1"}, "v1");
                tie_inputs<DType::f32, 1>({pre + "p1.v2", pre + "p2.v2"}, "v2");
            }
            tie_inputs<DType::f32, 1>({pre + "p1.u2", pre + "p2.u2"}, "u2");
            promote<DType::f32, 1>(pre + "p1.y", "y");
            promote<DType::f32, 1>(pre + "cred.z", "z");
            promote<DType::f32, 1>(pre + "cred.scale", "beta");
            promote<DType::f32, 1>(pre + "p2.scale", "alpha");
            promote<DType::f32, 2>(pre + "p1.P", "P");
            promote<DType::f32, 1>(pre + "cred.x", "x");
            promote<DType::f32, 1>(pre + "p2.w", "w");
            shape_hook(pre + "p1.P", cs::dim::fixed_at_run(strips), cs::dim::fixed_at_run(n));
            if constexpr (S == DType::e4m3)
                init_input("row_scale", hash_scales(seeds_.of(1), n));
            else
                init_input("row_scale", cs::init::fill(1.0f, {1}));
            init_input("A", hash_matrix<S>(seeds_.of(1), n));
        } else {
            if (style_ == Style::naive) {
                add_subsystem(cs::fuse_components(
                    "fw", cs::connects_desc({{"r2u.B", "gtp.B"}, {"gtp.P", "cred.P"}, {"r2u.B", "gemv.B"}, {"cred.x", "gemv.x"}}),
                    cs::segment<cs::range_mode::whole>(cs::subsystem_desc<Rank2Update>("r2u", R_), cs::subsystem_desc<GemvTPartial>("gtp")),
                    cs::global_barrier(), cs::subsystem_desc<ColReduce<true>>("cred"), cs::global_barrier(),
                    cs::subsystem_desc<Gemv>("gemv")));
                pre = "fw.";
            } else {
                if (style_ == Style::fused) {
                    // r2u and gtp share their (c, j) mapping, so the same work-item reads its own B writes
                    add_subsystem(cs::fuse_components("fw", cs::connects_desc({{"r2u.B", "gtp.B"}}),
                                                      cs::segment<cs::range_mode::whole>(cs::subsystem_desc<Rank2Update>("r2u", R_),
                                                                                         cs::subsystem_desc<GemvTPartial>("gtp"))));
                    pre = "fw.";
                } else {
                    add_subsystem(cs::subsystem_desc<Rank2Update>("r2u", R_));
                    add_subsystem(cs::subsystem_desc<GemvTPartial>("gtp"));
                    connect("r2u.B", "gtp.B");
                }
                add_subsystem(cs::subsystem_desc<ColReduce<true>>("cred"));
                add_subsystem(cs::subsystem_desc<Gemv>("gemv"));
                connect(pre + "gtp.P", "cred.P");
                connect(pre + "r2u.B", "gemv.B");
                connect("cred.x", "gemv.x");
                pre = "";
            }
            const std::string g = style_ == Style::naive ? "fw." : "";
            const std::string r = style_ == Style::unfused ? "" : "fw.";
            promote<DType::f32, 2>(r + "r2u.A", "A");
            promote<DType::f32, 1>(r + "r2u.u1", "u1");
            promote<DType::f32, 1>(r + "r2u.v1", "v1");
            promote<DType::f32, 1>(r + "r2u.u2", "u2");
            promote<DType::f32, 1>(r + "r2u.v2", "v2");
            promote<DType::f32, 1>(r + "gtp.y", "y");
            promote<DType::f32, 1>(g + "cred.z", "z");
            promote<DType::f32, 1>(g + "cred.scale", "beta");
            promote<DType::f32, 1>(g + "gemv.scale", "alpha");
            promote<DType::f32, 2>(r + "r2u.B", "B");
            promote<DType::f32, 2>(r + "gtp.P", "P");
            promote<DType::f32, 1>(g + "cred.x", "x");
            promote<DType::f32, 1>(g + "gemv.y", "w");
            shape_hook(r + "gtp.P", cs::dim::fixed_at_run(strips), cs::dim::fixed_at_run(n));
            init_input("A", hash_matrix<DType::f32>(seeds_.of(1), n));
        }
        init_input("u1", hash_vector(seeds_.of(2), n));
        init_input("v1", hash_vector(seeds_.of(3), n));
        init_input("u2", hash_vector(seeds_.of(4), n));
        init_input("v2", hash_vector(seeds_.of(5), n));
        init_input("y", hash_vector(seeds_.of(6), n));
        init_input("z", hash_vector(seeds_.of(7), n));
        init_input("alpha", cs::init::fill(alpha_, {1}));
        init_input("beta", cs::init::fill(beta_, {1}));
    }

private:
    Style      style_;
    int        n_, R_;
    float      alpha_, beta_;
    InputSeeds seeds_;
};

}  // namespace gemver
