#pragma once

// =============================================================================
// cospanmdao_amd/workloads/chain_systems.hpp — the GEMVER -> BiCGK -> GEMVER chain as Systems, in the reference's styles
// =============================================================================
//
// What the chain computes: G1 = GEMVER(A; y, z) giving B1, x1, w1; BiCGK on B1 with p = x1, r = w1 giving q = B1 p and
// s = B1^T r; G2 = GEMVER(B1; y = q, z = s) giving B2, x2, w2. Inside each GEMVER alpha = beta = 1/sqrt(n), held by one
// `scale` input that the column reduction and the row GEMV share through a tie (the reference's convention,
// examples/gemver-chain/systems.hpp:79 and chain_data.hpp:13); BiCGK runs at scale 1.
//
// Port names at the root, the same in every style so that cs_chain_* reads one set of paths:
//   in   A; u1a v1a u2a v2a ya za (G1); u1b v1b u2b v2b (G2's own copies of the rank-2 vectors)
//   out  B1 x1 w1; q s; B2 x2 w2
//
// Launches per chain and what shares a launch (the reference's six styles, systems.hpp:38-53; components from
// gemver_components.hpp and chain_components.hpp):
//   unfused   11  one launch per component
//   parallel  11  the same launches; BiCGK's two products are independent leaves, so the executor puts them on two streams
//   fused      8  per GEMVER one launch for rank-2 update + column partials (one segment over the whole range), then the
//                 column reduction and the row GEMV; BiCGK's two products in one launch (a segment each)
//   onepass    9  fused GEMVERs; BiCGK reads B1 once (BicgkOnePass) and reduces its row and column partials
//   naive      8  three fused groups whose global barriers become launch boundaries: G1, BiCGK, G2
//   crafted    8  three fused groups over the LDS-tile components; the middle one lets BiCGK and G2's rank-2 update share
//                 one read of B1 (strips of kRt rows)
// Results: unfused, parallel and fused add in one order, onepass and naive in another, so each class agrees bit for bit;
// crafted equals onepass at R = kRt.
//
// Inputs come from the reference's counter hash on the device (chain_data.hpp:34-72, seeds A = 1 .. z = 7); G2's u1b ..
// v2b hold the same values as G1's vectors (main.cpp:85-93).

#include <cmath>
#include <cstdint>
#include <string>

#include "naga/cospan/cospan.hpp"
#include "chain_components.hpp"
#include "gemver_systems.hpp"

namespace chain {

namespace cs = naga::cospan;
using cs::DType;
using gemver::kWg;

enum class Style { unfused = 0, parallel = 1, fused = 2, onepass = 3, naive = 4, crafted = 5 };

// =============================================================================
// GEMVER: ports A u1 v1 u2 v2 y z -> B x w (systems.hpp:71-139)
// =============================================================================
class Gemver : public cs::System {
public:
    Gemver(Style style, int n, int R) : style_{style}, n_{n}, R_{R} {}

    void setup() override {
        using namespace gemver;
        const float       ab = 1.0f / std::sqrt(static_cast<float>(n_));
        const std::size_t n = static_cast<std::size_t>(n_), strips = n / static_cast<std::size_t>(R_);
        if (style_ == Style::naive) {
            add_subsystem(cs::fuse_components(
                "fw", cs::connects_desc({{"r2u.B", "gtp.B"}, {"gtp.P", "cred.P"}, {"r2u.B", "gemv.B"}, {"cred.x", "gemv.x"}}),
                cs::segment<cs::range_mode::whole>(cs::subsystem_desc<Rank2Update>("r2u", R_), cs::subsystem_desc<GemvTPartial>("gtp")),
                cs::global_barrier(), cs::subsystem_desc<ColReduce<true>>("cred"), cs::global_barrier(), cs::subsystem_desc<Gemv>("gemv")));
            promote_ports("fw.", "fw.");
            shape_hook("fw.gtp.P", cs::dim::fixed_at_run(strips), cs::dim::fixed_at_run(n));
            tie_inputs<DType::f32, 1>({"fw.cred.scale", "fw.gemv.scale"}, "scale", cs::shape(1));
            init_input("scale", cs::init::fill(ab));
            return;
        }
        std::string pre;
        if (style_ == Style::fused || style_ == Style::onepass) {
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
        promote_ports(pre, "");
        shape_hook(pre + "gtp.P", cs::dim::fixed_at_run(strips), cs::dim::fixed_at_run(n));
        tie_inputs<DType::f32, 1>({"cred.scale", "gemv.scale"}, "scale", cs::shape(1));
        init_input("scale", cs::init::fill(ab));
    }

private:
    /// r: the prefix of r2u and gtp, g: of cred and gemv
    void promote_ports(const std::string& r, const std::string& g) {
        promote<DType::f32, 2>(r + "r2u.A", "A");
        promote<DType::f32, 1>(r + "r2u.u1", "u1");
        promote<DType::f32, 1>(r + "r2u.v1", "v1");
        promote<DType::f32, 1>(r + "r2u.u2", "u2");
        promote<DType::f32, 1>(r + "r2u.v2", "v2");
        promote<DType::f32, 1>(r + "gtp.y", "y");
        promote<DType::f32, 1>(g + "cred.z", "z");
        promote<DType::f32, 2>(r + "r2u.B", "B");
        promote<DType::f32, 1>(g + "cred.x", "x");
        promote<DType::f32, 1>(g + "gemv.y", "w");
    }

    Style style_;
    int   n_, R_;
};

// =============================================================================
// BiCGK: ports B p r -> q s (systems.hpp:144-221)
// =============================================================================
class Bicgk : public cs::System {
public:
    Bicgk(Style style, int n, int R) : style_{style}, n_{n}, R_{R} {}

    void setup() override {
        using namespace gemver;
        const std::size_t n = static_cast<std::size_t>(n_), strips = n / static_cast<std::size_t>(R_);
        if (style_ == Style::naive || style_ == Style::onepass) {
            std::string pre;
            if (style_ == Style::naive) {
                add_subsystem(cs::fuse_components(
                    "fw", cs::connects_desc({{"onepass.Pq", "rowred.Pq"}, {"onepass.Ps", "cred_s.P"}}),
                    cs::subsystem_desc<BicgkOnePass>("onepass"), cs::global_barrier(),
                    cs::segment(cs::subsystem_desc<RowReduce>("rowred"), cs::subsystem_desc<ColReduce<false>>("cred_s"))));
                pre = "fw.";
            } else {
                add_subsystem(cs::subsystem_desc<BicgkOnePass>("onepass"));
                add_subsystem(cs::subsystem_desc<RowReduce>("rowred"));
                add_subsystem(cs::subsystem_desc<ColReduce<false>>("cred_s"));
                connect("onepass.Pq", "rowred.Pq");
                connect("onepass.Ps", "cred_s.P");
            }
            promote<DType::f32, 2>(pre + "onepass.B", "B");
            promote<DType::f32, 1>(pre + "onepass.p", "p");
            promote<DType::f32, 1>(pre + "onepass.r", "r");
            promote<DType::f32, 1>(pre + "rowred.q", "q");
            promote<DType::f32, 1>(pre + "cred_s.x", "s");
            shape_hook(pre + "onepass.Ps", cs::dim::fixed_at_run(strips), cs::dim::fixed_at_run(n));
            promote<DType::f32, 1>(pre + "cred_s.scale", "scale");  // one user: a promotion (a tie needs two members)
            init_input("scale", cs::init::fill(1.0f, {1}));
            return;
        }
        std::string pre;
        if (style_ == Style::parallel) {
            add_subsystem(cs::system_from("par", cs::subsystem_desc<Gemv>("gemv_q"), cs::subsystem_desc<GemvTPartial>("gtp_s")));
            pre = "par.";
        } else if (style_ == Style::fused) {
            add_subsystem(cs::fuse_components("fw", cs::subsystem_desc<Gemv>("gemv_q"), cs::subsystem_desc<GemvTPartial>("gtp_s")));
            pre = "fw.";
        } else {
            add_subsystem(cs::subsystem_desc<Gemv>("gemv_q"));
            add_subsystem(cs::subsystem_desc<GemvTPartial>("gtp_s"));
        }
        add_subsystem(cs::subsystem_desc<ColReduce<false>>("cred_s"));
        connect(pre + "gtp_s.P", "cred_s.P");
        tie_inputs<DType::f32, 2>({pre + "gemv_q.B", pre + "gtp_s.B"}, "B");
        promote<DType::f32, 1>(pre + "gemv_q.x", "p");
        promote<DType::f32, 1>(pre + "gtp_s.y", "r");
        promote<DType::f32, 1>(pre + "gemv_q.y", "q");
        promote<DType::f32, 1>("cred_s.x", "s");
        shape_hook(pre + "gtp_s.P", cs::dim::fixed_at_run(strips), cs::dim::fixed_at_run(n));
        tie_inputs<DType::f32, 1>({pre + "gemv_q.scale", "cred_s.scale"}, "scale", cs::shape(1));
        init_input("scale", cs::init::fill(1.0f));
    }

private:
    Style style_;
    int   n_, R_;
};

// =============================================================================
// The chain: G1 -> BiCGK -> G2 (systems.hpp:226-346)
// =============================================================================
class Chain : public cs::System {
public:
    /// n: the size; R: the strip height (the crafted style's strips are kRt rows whatever R is); seed: the inputs' base seed
    Chain(Style style, int n, int R, std::uint32_t seed) : style_{style}, n_{n}, R_{R}, seeds_{seed} {}

    void setup() override {
        if (style_ == Style::crafted)
            setup_crafted();
        else
            setup_children();
        const std::size_t n = static_cast<std::size_t>(n_);
        init_input("A", gemver::hash_matrix<DType::f32>(seeds_.of(1), n));
        const char* a[6] = {"u1a", "v1a", "u2a", "v2a", "ya", "za"};
        const char* b[4] = {"u1b", "v1b", "u2b", "v2b"};
        for (std::uint32_t i = 0; i < 6; ++i) init_input(a[i], gemver::hash_vector(seeds_.of(2 + i), n));
        for (std::uint32_t i = 0; i < 4; ++i) init_input(b[i], gemver::hash_vector(seeds_.of(2 + i), n));
    }

private:
    void setup_children() {
        add_subsystem(cs::subsystem_desc<Gemver>("g1", style_, n_, R_));
        add_subsystem(cs::subsystem_desc<Bicgk>("bicgk", style_, n_, R_));
        add_subsystem(cs::subsystem_desc<Gemver>("g2", style_, n_, R_));
        connect("g1.B", "bicgk.B");
        connect("g1.x", "bicgk.p");  // p = x1
        connect("g1.w", "bicgk.r");  // r = w1
        connect("g1.B", "g2.A");     // G2's A is B1
        connect("bicgk.q", "g2.y");  // y = q
        connect("bicgk.s", "g2.z");  // z = s
        promote<DType::f32, 2>("g1.A", "A");
        promote<DType::f32, 1>("g1.u1", "u1a");
        promote<DType::f32, 1>("g1.v1", "v1a");
        promote<DType::f32, 1>("g1.u2", "u2a");
        promote<DType::f32, 1>("g1.v2", "v2a");
        promote<DType::f32, 1>("g1.y", "ya");
        promote<DType::f32, 1>("g1.z", "za");
        promote<DType::f32, 1>("g2.u1", "u1b");
        promote<DType::f32, 1>("g2.v1", "v1b");
        promote<DType::f32, 1>("g2.u2", "u2b");
        promote<DType::f32, 1>("g2.v2", "v2b");
        promote<DType::f32, 2>("g1.B", "B1");
        promote<DType::f32, 1>("g1.x", "x1");
        promote<DType::f32, 1>("g1.w", "w1");
        promote<DType::f32, 1>("bicgk.q", "q");
        promote<DType::f32, 1>("bicgk.s", "s");
        promote<DType::f32, 2>("g2.B", "B2");
        promote<DType::f32, 1>("g2.x", "x2");
        promote<DType::f32, 1>("g2.w", "w2");
    }

    /// Three fused launches over the local-memory components (systems.hpp:270-344).
    void setup_crafted() {
        using namespace gemver;
        const float       ab = 1.0f / std::sqrt(static_cast<float>(n_));
        const std::size_t n = static_cast<std::size_t>(n_), strips = n / kRt;

        // G1: the B tile stays in local memory from the rank-2 update to the partial column sums; then x needs every strip's
        // partials (barrier), and w needs the whole x (barrier).
        add_subsystem(cs::fuse_components(
            "g1", cs::connects_desc({{"gtp.P", "cred.P"}, {"r2u.B", "gemv.B"}, {"cred.x", "gemv.x"}}),
            cs::segment<cs::local_mem::shared>(cs::subsystem_desc<Rank2UpdateLocal>("r2u"), cs::subsystem_desc<GemvTPartialLocal>("gtp")),
            cs::global_barrier(), cs::subsystem_desc<ColReduce<true>>("cred"), cs::global_barrier(), cs::subsystem_desc<Gemv>("gemv")));
        // The middle launch: BiCGK and G2's rank-2 update share ONE read of B1 (G2's update needs neither q nor s).
        add_subsystem(cs::fuse_components(
            "mid", cs::connects_desc({{"bk.Pq", "rowred.Pq"}, {"bk.Ps", "cred_s.P"}}),
            cs::segment<cs::local_mem::shared>(cs::subsystem_desc<BicgkTileLocal>("bk"), cs::subsystem_desc<Rank2UpdateFromLocal>("r2u")),
            cs::global_barrier(), cs::segment(cs::subsystem_desc<RowReduce>("rowred"), cs::subsystem_desc<ColReduce<false>>("cred_s"))));
        // G2: the partial column sums need y = q, so they read B2 from global memory: the two-pass form's pass 1 with zero
        // rank-2 vectors (fma(0, 0, b) = b: GemvTPartial's fma chain over each strip, bit for bit, at the streaming rate).
        add_subsystem(cs::fuse_components("g2", cs::connects_desc({{"gtp.P", "cred.P"}, {"cred.x", "gemv.x"}}),
                                          cs::subsystem_desc<AhatTPartial<DType::f32>>("gtp", kRt), cs::global_barrier(),
                                          cs::subsystem_desc<ColReduce<true>>("cred"), cs::global_barrier(), cs::subsystem_desc<Gemv>("gemv")));
        tie_inputs<DType::f32, 1>({"g1.cred.scale", "g1.gemv.scale"}, "g1_scale", cs::shape(1));
        tie_inputs<DType::f32, 1>({"g2.cred.scale", "g2.gemv.scale"}, "g2_scale", cs::shape(1));

        connect("g1.r2u.B", "mid.bk.B");      // B1
        connect("g1.cred.x", "mid.bk.p");     // p = x1
        connect("g1.gemv.y", "mid.bk.r");     // r = w1
        connect("mid.r2u.B", "g2.gtp.A");     // B2 feeds both of G2's readers
        connect("mid.r2u.B", "g2.gemv.B");
        connect("mid.rowred.q", "g2.gtp.y");  // y = q
        connect("mid.cred_s.x", "g2.cred.z"); // z = s

        promote<DType::f32, 2>("g1.r2u.A", "A");
        promote<DType::f32, 1>("g1.r2u.u1", "u1a");
        promote<DType::f32, 1>("g1.r2u.v1", "v1a");
        promote<DType::f32, 1>("g1.r2u.u2", "u2a");
        promote<DType::f32, 1>("g1.r2u.v2", "v2a");
        promote<DType::f32, 1>("g1.gtp.y", "ya");
        promote<DType::f32, 1>("g1.cred.z", "za");
        promote<DType::f32, 1>("mid.r2u.u1", "u1b");
        promote<DType::f32, 1>("mid.r2u.v1", "v1b");
        promote<DType::f32, 1>("mid.r2u.u2", "u2b");
        promote<DType::f32, 1>("mid.r2u.v2", "v2b");
        promote<DType::f32, 2>("g1.r2u.B", "B1");
        promote<DType::f32, 1>("g1.cred.x", "x1");
        promote<DType::f32, 1>("g1.gemv.y", "w1");
        promote<DType::f32, 1>("mid.rowred.q", "q");
        promote<DType::f32, 1>("mid.cred_s.x", "s");
        promote<DType::f32, 1>("mid.cred_s.scale", "mid_scale");  // one user: a promotion (a tie needs two members)
        promote<DType::f32, 2>("mid.r2u.B", "B2");
        promote<DType::f32, 1>("g2.cred.x", "x2");
        promote<DType::f32, 1>("g2.gemv.y", "w2");

        const auto ds = cs::dim::fixed_at_run(strips), dn = cs::dim::fixed_at_run(n);
        shape_hook("g1.gtp.P", ds, dn);
        shape_hook("mid.bk.Ps", ds, dn);
        shape_hook("g2.gtp.P", ds, dn);
        init_input("g1_scale", cs::init::fill(ab));
        init_input("mid_scale", cs::init::fill(1.0f, {1}));
        init_input("g2_scale", cs::init::fill(ab));
        init_input("g2.gtp.row_scale", cs::init::fill(1.0f, {1}));
        for (const char* v : {"g2.gtp.u1", "g2.gtp.v1", "g2.gtp.u2", "g2.gtp.v2"}) init_input(v, cs::init::fill(0.0f, {n}));
    }

    Style              style_;
    int                n_, R_;
    gemver::InputSeeds seeds_;
};

}  // namespace chain
