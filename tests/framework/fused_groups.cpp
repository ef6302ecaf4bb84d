// tests/framework/fused_groups.cpp — the fused-group machinery through cs::Model on the GPU (a test program).
//
// The cases and their closed-form expectations follow the reference's tests/test_fused_groups.cpp:254-517 (segments,
// range modes, local memory, work-group shape errors, bare components, one component type in two groups, group waits,
// global barriers, wait_group / wait_earlier_segments across groups of an earlier segment, several forwards with new
// inputs). The components are this file's own. Prints one line per case and "OK" at the end when every check holds;
// exits 1 otherwise.
#include <algorithm>
#include <cstdio>
#include <functional>
#include <string>
#include <vector>

#include <naga/cospan/cospan.hpp>

namespace cs = naga::cospan;
using cs::DType;

static int g_fail = 0;
#define CHECK(cond, ...)                                                  \
    do {                                                                  \
        if (!(cond)) {                                                    \
            ++g_fail;                                                     \
            std::printf("  FAIL line %d: ", __LINE__);                    \
            std::printf(__VA_ARGS__);                                     \
            std::printf("\n");                                            \
        }                                                                 \
    } while (0)

namespace tfg {

constexpr std::size_t kWg = 64;

using Build = std::function<void(cs::System&)>;
class Top : public cs::System {
    Build build_;

public:
    explicit Top(Build b) : build_{std::move(b)} {}
    void setup() override { build_(*this); }
};
cs::Model make(Build b) { return cs::create_model<Top>(std::move(b)); }

std::vector<float> out_of(cs::Model& m, const char* path) {
    auto v = m.view_output<DType::f32, 1>(path);
    v.sync();
    return std::vector<float>(v.data(), v.data() + v.size());
}
std::vector<float> filled(std::size_t n, float v) { return std::vector<float>(n, v); }
bool same(const std::vector<float>& a, const std::vector<float>& b) { return a == b; }

CS_COMPONENTS_BEGIN

/// y[i] = i + 1 (index range)
class Src : public cs::Component<Src> {
public:
    struct Inputs {};
    struct Outputs {
        cs::Output<DType::f32, 1> y;
    };
    static void compute_forward(cs::id<1> i, const Inputs&, const Outputs& out) { out.y[i] = static_cast<float>(i[0] + 1); }
    Inputs  inputs;
    Outputs outputs;
    explicit Src(int n) : n_{n} {}
    void declare() override { declare_output(outputs.y, "y", n_); }
    std::unique_ptr<cs::LaunchDescription> define_forward() override { return create_launch_config(cs::range_desc(outputs.y.shape())); }

private:
    int n_;
};

// What a group stage writes, by kind (a compile-time parameter, so every body is straight-line code):
constexpr int kTag   = 0;  // segment * 1000 + groups(segment 0) * 100 + groups(segment 1) * 10 + segment count
constexpr int kRange = 1;  // the group range the stage sees
constexpr int kLeave = 2;  // shared local memory, first stage: t[l] = l + 1, y = x
constexpr int kTake  = 3;  // shared local memory, second stage: after a barrier y = t[wg - 1 - l]
constexpr int kWrite = 4;  // y = x + 1
template <int Mode>
constexpr int kReadNext = 10 + Mode;  // y(group g) = x(group g + 1): 0 nothing else, 1 wait_group, 2 wait_earlier_segments
template <int Base>
constexpr int kOwn = 1000 + Base;  // partitioned local memory: t[l] = Base + l, barrier, y = t[wg - 1 - l]

using TileLocal = cs::LocalMem<cs::LocalArray<DType::f32, 1>>;

template <int Kind>
class GStage : public cs::Component<GStage<Kind>> {
public:
    struct Inputs {
        cs::Input<DType::f32, 1> x;
    };
    struct Outputs {
        cs::Output<DType::f32, 1> y;
    };
    static void compute_forward(cs::GroupItem<1> it, const Inputs& in, const Outputs& out, TileLocal lm) {
        const std::size_t g = it.get_global_id(0);
        if constexpr (Kind == kTag) {
            if (g >= in.x.shape()[0]) return;
            const std::size_t segs = it.get_segment_count();
            const std::size_t gc0  = it.get_segment_group_count(0)[0];
            const std::size_t gc1  = segs > 1 ? it.get_segment_group_count(1)[0] : 0;
            out.y[g]               = static_cast<float>(it.get_segment() * 1000 + gc0 * 100 + gc1 * 10 + segs);
        } else if constexpr (Kind == kRange) {
            if (g < in.x.shape()[0]) out.y[g] = static_cast<float>(it.get_group_range(0));
        } else if constexpr (Kind == kLeave) {
            auto [t]                    = lm.get(it);
            t[it.get_local_linear_id()] = static_cast<float>(it.get_local_linear_id() + 1);
            if (g < in.x.shape()[0]) out.y[g] = in.x[g];
        } else if constexpr (Kind == kTake) {
            auto [t] = lm.get(it);
            cs::group_barrier(it.get_group());
            if (g < in.x.shape()[0]) out.y[g] = t[kWg - 1 - it.get_local_linear_id()];
        } else if constexpr (Kind >= 1000) {
            constexpr int Base          = Kind - 1000;
            auto [t]                    = lm.get(it);
            t[it.get_local_linear_id()] = static_cast<float>(Base + static_cast<int>(it.get_local_linear_id()));
            cs::group_barrier(it.get_group());
            if (g < in.x.shape()[0]) out.y[g] = t[kWg - 1 - it.get_local_linear_id()];
        } else if constexpr (Kind == kWrite) {
            if (g < in.x.shape()[0]) out.y[g] = in.x[g] + 1.0f;
        } else {
            constexpr int     Mode = Kind - 10;
            const std::size_t G = it.get_group_range(0), gi = it.get_group(0), wg = it.get_local_range(0), l = it.get_local_linear_id();
            if constexpr (Mode == 1) it.wait_group(0, cs::id<1>{(gi + 1) % G});
            if constexpr (Mode == 2) it.wait_earlier_segments();
            out.y[gi * wg + l] = in.x[((gi + 1) % G) * wg + l];
        }
    }

    Inputs  inputs;
    Outputs outputs;
    explicit GStage(std::size_t wg = kWg, bool local = false) : wg_{wg}, local_{local} {}
    void declare() override {
        this->declare_input(inputs.x, "x", cs::unspecified_shape<1>());
        this->declare_output(outputs.y, "y", cs::unspecified_shape<1>());
    }
    void prepare_forward() override { this->set_shape(outputs.y, this->shape_of(inputs.x)[0]); }
    std::unique_ptr<cs::LaunchDescription> define_forward() override {
        const std::size_t n = (inputs.x.shape()[0] + wg_ - 1) / wg_ * wg_;
        if (local_)
            return this->create_launch_config(cs::range_desc(cs::range<1>{n}), cs::wg_desc(cs::range<1>{wg_}),
                                              cs::local_desc(cs::range<1>{kWg}));
        return this->create_launch_config(cs::range_desc(cs::range<1>{n}), cs::wg_desc(cs::range<1>{wg_}));
    }

private:
    std::size_t wg_;
    bool        local_;
};

/// Every group-function family over one work-group of W0 x W1 work-items (x = local linear id + 1) and a joint range of M
/// local-memory floats (L[i] = i + 1), after the reference's tests/test_range_split.cpp:284-343. r[0..31]: group-wide
/// results (leader); r[32 + 6 t + k]: work-item t's scans; r[32 + 6 n + i], r[32 + 6 n + M + i]: the joint exclusive and
/// inclusive scans of the range.
template <std::size_t W0, std::size_t W1, std::size_t M>
class FamStage : public cs::Component<FamStage<W0, W1, M>> {
public:
    static constexpr std::size_t n = W0 * W1;
    struct Inputs {};
    struct Outputs {
        cs::Output<DType::f32, 1> r;
    };
    static void compute_forward(cs::GroupItem<2> it, const Inputs&, const Outputs& out, TileLocal lm) {
        auto              g   = it.get_group();
        auto [Lv]             = lm.get(it);
        float*            L   = &Lv[std::size_t{0}];
        const std::size_t t   = it.get_local_linear_id();
        const float       x   = static_cast<float>(t + 1);
        const auto        pls = cs::plus<float>();
        for (std::size_t i = t; i < M; i += n) L[i] = static_cast<float>(i + 1);
        cs::group_barrier(g);

        const float red      = cs::reduce_over_group(g, x, pls);
        const float red_init = cs::reduce_over_group(g, x, 10.0f, pls);
        const float red_max  = cs::reduce_over_group(g, x, cs::maximum<float>());
        const float red_min  = cs::reduce_over_group(g, x, cs::minimum<float>());
        const float bc0      = cs::group_broadcast(g, x);
        const float bc_lin   = cs::group_broadcast(g, x, std::size_t{5});
        const float bc_id    = cs::group_broadcast(g, x, cs::id<2>{1, 2});
        const bool  any_t    = cs::any_of_group(g, x == static_cast<float>(n));
        const bool  any_f    = cs::any_of_group(g, x > 100000.0f);
        const bool  all_t    = cs::all_of_group(g, x > 0.0f);
        const bool  none_t   = cs::none_of_group(g, x > static_cast<float>(n));
        const float ex       = cs::exclusive_scan_over_group(g, x, pls);
        const float in       = cs::inclusive_scan_over_group(g, x, pls);
        const float ex_i     = cs::exclusive_scan_over_group(g, x, 100.0f, pls);
        const float in_i     = cs::inclusive_scan_over_group(g, x, pls, 100.0f);
        const float jred     = cs::joint_reduce(g, L, L + M, pls);
        const float jred_i   = cs::joint_reduce(g, L, L + M, 1.0f, pls);
        const float jmin     = cs::joint_reduce(g, L, L + M, cs::minimum<float>());
        const bool  jany     = cs::joint_any_of(g, L, L + M, [](float v) { return v == static_cast<float>(M); });
        const bool  jall     = cs::joint_all_of(g, L, L + M, [](float v) { return v > 0.0f; });
        const bool  jnone    = cs::joint_none_of(g, L, L + M, [](float v) { return v > static_cast<float>(M); });
        float* const jout    = &out.r[std::size_t{32 + 6 * n}];
        cs::joint_exclusive_scan(g, L, L + M, L + M, pls);
        cs::group_barrier(g);
        const float jex = t < M ? L[M + t] : 0.0f;
        for (std::size_t i = t; i < M; i += n) jout[i] = L[M + i];
        cs::group_barrier(g);
        cs::joint_inclusive_scan(g, L, L + M, L + M, pls);
        cs::group_barrier(g);
        const float jin = t < M ? L[M + t] : 0.0f;
        for (std::size_t i = t; i < M; i += n) jout[M + i] = L[M + i];

        cs::sub_group sg  = it.get_sub_group();
        const float   px  = cs::permute_group_by_xor(sg, x, 0u);
        const float   sel = cs::select_from_group(sg, x, sg.get_local_id());
        const float   sl  = cs::shift_group_left(sg, x, 0);
        const float   sr  = cs::shift_group_right(sg, x, 0);
        const float   px1 = cs::permute_group_by_xor(sg, x, 1u);
        const float   sel3 = cs::select_from_group(sg, x, cs::id<1>{3});
        const float   sl1 = cs::shift_group_left(sg, x);
        const float   sr1 = cs::shift_group_right(sg, x);

        out.r[32 + t * 6 + 0] = ex;
        out.r[32 + t * 6 + 1] = in;
        out.r[32 + t * 6 + 2] = ex_i;
        out.r[32 + t * 6 + 3] = in_i;
        out.r[32 + t * 6 + 4] = jex;
        out.r[32 + t * 6 + 5] = jin;
        // the wave shuffles by 1 where the source lane is inside the item's own wave: r[24..27] from work-item 5
        if (t == 5) {
            out.r[24] = px1;   // lane 5 ^ 1 = 4 -> x 5
            out.r[25] = sel3;  // lane 3 -> x 4
            out.r[26] = sl1;   // lane 6 -> x 7
            out.r[27] = sr1;   // lane 4 -> x 5
        }
        if (g.leader()) {
            const float v[24] = {red, red_init, bc0, bc_lin, bc_id, any_t ? 1.f : 0.f, all_t ? 1.f : 0.f, none_t ? 1.f : 0.f,
                                 any_f ? 1.f : 0.f, jred, jred_i, jany ? 1.f : 0.f, jall ? 1.f : 0.f, jnone ? 1.f : 0.f, red_max, px,
                                 sel, sl, sr, red_min, jmin, static_cast<float>(sg.get_local_linear_range()),
                                 static_cast<float>(sg.get_group_range()[0]), static_cast<float>(sg.get_max_local_range()[0])};
            for (std::size_t k = 0; k < 24; ++k) out.r[k] = v[k];
        }
    }

    Inputs  inputs;
    Outputs outputs;

    void declare() override { this->declare_output(outputs.r, "r", 32 + 6 * n + 2 * M); }
    std::unique_ptr<cs::LaunchDescription> define_forward() override {
        return this->create_launch_config(cs::range_desc(cs::range<2>{W0, W1}), cs::wg_desc(cs::range<2>{W0, W1}),
                                          cs::local_desc(cs::range<1>{2 * M}));
    }
};

CS_COMPONENTS_END

/// The closed forms of FamStage's results.
template <std::size_t W0, std::size_t W1, std::size_t M>
void group_functions_case() {
    constexpr std::size_t n     = W0 * W1;
    cs::Model             model = make([](cs::System& s) { s.add_subsystem(cs::subsystem_desc<FamStage<W0, W1, M>>("a")); });
    model.setup();
    model.run();
    model.wait();
    const std::vector<float> r = out_of(model, "a.r");
    CHECK(r.size() == 32 + 6 * n + 2 * M, "size %zu", r.size());
    if (r.size() != 32 + 6 * n + 2 * M) return;
    const auto f = [](std::size_t v) { return static_cast<float>(v); };
    const std::size_t lanes0 = n < 64 ? n : 64;
    const float want[24] = {f(n * (n + 1) / 2), f(n * (n + 1) / 2 + 10), 1, 6, f(W1 + 3), 1, 1, 1, 0, f(M * (M + 1) / 2),
                            f(M * (M + 1) / 2 + 1), 1, 1, 1, f(n), 1, 1, 1, 1, 1, 1, f(lanes0), f((n + 63) / 64), 64};
    for (std::size_t k = 0; k < 24; ++k) CHECK(r[k] == want[k], "%zux%zu group-wide result %zu: %g, want %g", W0, W1, k, r[k], want[k]);
    const float shuf[4] = {5, 4, 7, 5};
    for (std::size_t k = 0; k < 4; ++k) CHECK(r[24 + k] == shuf[k], "%zux%zu shuffle %zu: %g, want %g", W0, W1, k, r[24 + k], shuf[k]);
    for (std::size_t t = 0; t < n; ++t) {
        const float ex = f(t * (t + 1) / 2), in = f((t + 1) * (t + 2) / 2);
        const float jex = t < M ? ex : 0.0f, jin = t < M ? in : 0.0f;
        const float got[6] = {r[32 + t * 6], r[32 + t * 6 + 1], r[32 + t * 6 + 2], r[32 + t * 6 + 3], r[32 + t * 6 + 4], r[32 + t * 6 + 5]};
        const float exp[6] = {ex, in, ex + 100, in + 100, jex, jin};
        for (int k = 0; k < 6; ++k) CHECK(got[k] == exp[k], "%zux%zu work-item %zu scan %d: %g, want %g", W0, W1, t, k, got[k], exp[k]);
    }
    for (std::size_t i = 0; i < M; ++i) {
        CHECK(r[32 + 6 * n + i] == f(i * (i + 1) / 2), "%zux%zu joint exclusive %zu: %g", W0, W1, i, r[32 + 6 * n + i]);
        CHECK(r[32 + 6 * n + M + i] == f((i + 1) * (i + 2) / 2), "%zux%zu joint inclusive %zu: %g", W0, W1, i, r[32 + 6 * n + M + i]);
    }
}

using TagStage   = GStage<kTag>;
using RangeStage = GStage<kRange>;

/// y = the next group's x + 1 over `forwards` forwards with new inputs each time. Kind 1: group_waits_enabled, no
/// barrier; 2: a global_barrier before the second segment.
template <int Mode, int Kind>
void neighbor_case(std::size_t wg, std::size_t G, int forwards) {
    cs::Model model = make([wg](cs::System& s) {
        auto a = cs::subsystem_desc<GStage<kWrite>>("a", wg);
        auto b = cs::subsystem_desc<GStage<kReadNext<Mode>>>("b", wg);
        if constexpr (Kind == 2)
            s.add_subsystem(cs::fuse_components("f", cs::connects_desc({{"a.y", "b.x"}}), cs::segment(a), cs::global_barrier(), cs::segment(b)));
        else
            s.add_subsystem(cs::fuse_components<cs::group_waits_enabled>("f", cs::connects_desc({{"a.y", "b.x"}}), cs::segment(a), cs::segment(b)));
    });
    model.setup();
    const std::size_t N = G * wg;
    for (int k = 0; k < forwards; ++k) {
        std::vector<float> x(N), want(N);
        for (std::size_t i = 0; i < N; ++i) x[i] = static_cast<float>(k * 1000000 + static_cast<int>(i));
        for (std::size_t g = 0; g < G; ++g)
            for (std::size_t l = 0; l < wg; ++l) want[g * wg + l] = x[((g + 1) % G) * wg + l] + 1.0f;
        model.update_input("f.a.x", cs::init::from_host(x));
        model.run();
        model.wait();
        CHECK(same(out_of(model, "f.b.y"), want), "neighbor case mode %d kind %d, %zu groups of %zu, forward %d", Mode, Kind, G, wg, k);
    }
}

}  // namespace tfg

using namespace tfg;

#ifndef CS_SOURCE_HASH
#define CS_SOURCE_HASH "unknown"
#endif

int main() {
    std::printf("build %s\n", CS_SOURCE_HASH);
    int cus = 0;
    cs_hip_device_info(0, nullptr, 0, &cus, nullptr, nullptr);
    const std::size_t big_wg = 1024, big_groups = 4 * static_cast<std::size_t>(std::max(cus, 1));

    std::printf("group functions, one group of 4 x 4 and of 4 x 25 (a partly populated second wave)\n");
    group_functions_case<4, 4, 16>();
    group_functions_case<4, 25, 307>();
    group_functions_case<8, 32, 1000>();

    std::printf("two segments, each stage sees its own segment's group count\n");
    {
        cs::Model model = make([](cs::System& s) {
            s.add_subsystem(cs::subsystem_desc<Src>("sa", 200));  // 4 groups of 64
            s.add_subsystem(cs::subsystem_desc<Src>("sb", 100));  // 2 groups
            s.add_subsystem(cs::fuse_components("f", cs::segment(cs::subsystem_desc<TagStage>("a")), cs::segment(cs::subsystem_desc<TagStage>("b"))));
            s.connect("sa.y", "f.a.x");
            s.connect("sb.y", "f.b.x");
        });
        model.setup();
        for (int fwd = 0; fwd < 3; ++fwd) {
            model.run();
            model.wait();
            CHECK(same(out_of(model, "f.a.y"), filled(200, 422)), "f.a.y forward %d", fwd);
            CHECK(same(out_of(model, "f.b.y"), filled(100, 1422)), "f.b.y forward %d", fwd);
        }
    }
    std::printf("range mode own vs whole\n");
    for (bool whole : {false, true}) {
        cs::Model model = make([whole](cs::System& s) {
            s.add_subsystem(cs::subsystem_desc<Src>("sa", 200));
            s.add_subsystem(cs::subsystem_desc<Src>("sb", 100));
            if (whole)
                s.add_subsystem(cs::fuse_components("f", cs::segment<cs::range_mode::whole>(cs::subsystem_desc<RangeStage>("a"), cs::subsystem_desc<RangeStage>("b"))));
            else
                s.add_subsystem(cs::fuse_components("f", cs::segment<cs::range_mode::own>(cs::subsystem_desc<RangeStage>("a"), cs::subsystem_desc<RangeStage>("b"))));
            s.connect("sa.y", "f.a.x");
            s.connect("sb.y", "f.b.x");
        });
        model.setup();
        model.run();
        model.wait();
        CHECK(same(out_of(model, "f.a.y"), filled(200, 4.0f)), "range a (whole %d)", whole);
        CHECK(same(out_of(model, "f.b.y"), filled(100, whole ? 4.0f : 2.0f)), "range b (whole %d)", whole);
    }
    std::printf("shared local memory carries a stage's values to the next\n");
    {
        cs::Model model = make([](cs::System& s) {
            s.add_subsystem(cs::subsystem_desc<Src>("src", 128));
            s.add_subsystem(cs::fuse_components("f", cs::connects_desc({{"a.y", "b.x"}}),
                                                cs::segment<cs::local_mem::shared>(cs::subsystem_desc<GStage<kLeave>>("a", kWg, true),
                                                                                   cs::subsystem_desc<GStage<kTake>>("b", kWg, true))));
            s.connect("src.y", "f.a.x");
        });
        model.setup();
        for (int fwd = 0; fwd < 2; ++fwd) {
            model.run();
            model.wait();
            std::vector<float> want(128);
            for (std::size_t g = 0; g < want.size(); ++g) want[g] = static_cast<float>(kWg - (g % kWg));
            CHECK(same(out_of(model, "f.b.y"), want), "shared local forward %d", fwd);
        }
    }
    std::printf("partitioned local memory gives each stage its own slice\n");
    {
        cs::Model model = make([](cs::System& s) {
            s.add_subsystem(cs::subsystem_desc<Src>("src", 128));
            s.add_subsystem(cs::fuse_components("f", cs::segment<cs::local_mem::partitioned>(cs::subsystem_desc<GStage<kOwn<1>>>("a", kWg, true),
                                                                                             cs::subsystem_desc<GStage<kOwn<1000>>>("b", kWg, true))));
            s.connect("src.y", "f.a.x");
            s.connect("src.y", "f.b.x");
        });
        model.setup();
        for (int fwd = 0; fwd < 2; ++fwd) {
            model.run();
            model.wait();
            std::vector<float> wa(128), wb(128);
            for (std::size_t g = 0; g < 128; ++g) {
                const std::size_t l = g % kWg;
                wa[g]               = static_cast<float>(1 + (kWg - 1 - l));
                wb[g]               = static_cast<float>(1000 + (kWg - 1 - l));
            }
            CHECK(same(out_of(model, "f.a.y"), wa), "partitioned a forward %d", fwd);
            CHECK(same(out_of(model, "f.b.y"), wb), "partitioned b forward %d", fwd);
        }
    }
    std::printf("stages with different work-group shapes throw at forward, naming both\n");
    {
        cs::Model model = make([](cs::System& s) {
            s.add_subsystem(cs::subsystem_desc<Src>("src", 256));
            s.add_subsystem(cs::fuse_components("f", cs::segment(cs::subsystem_desc<RangeStage>("a", std::size_t{32}),
                                                                 cs::subsystem_desc<RangeStage>("b", std::size_t{64}))));
            s.connect("src.y", "f.a.x");
            s.connect("src.y", "f.b.x");
        });
        model.setup();
        std::string what = "<no exception>";
        try {
            model.run();
        } catch (const std::exception& e) {
            what = e.what();
        }
        CHECK(what.find("f.b") != std::string::npos && what.find("different work-group shape") != std::string::npos, "message: %s",
              what.c_str());
        model.wait();
    }
    std::printf("a bare component launches as a group of one stage\n");
    {
        cs::Model model = make([](cs::System& s) {
            s.add_subsystem(cs::subsystem_desc<Src>("src", 256));
            s.add_subsystem(cs::subsystem_desc<TagStage>("t"));
            s.connect("src.y", "t.x");
        });
        model.setup();
        for (int fwd = 0; fwd < 3; ++fwd) {
            model.run();
            model.wait();
            CHECK(same(out_of(model, "t.y"), filled(256, 401)), "bare forward %d", fwd);
        }
    }
    std::printf("one component type in two groups\n");
    {
        cs::Model model = make([](cs::System& s) {
            s.add_subsystem(cs::subsystem_desc<Src>("s1", 200));
            s.add_subsystem(cs::subsystem_desc<Src>("s2", 100));
            s.add_subsystem(cs::subsystem_desc<Src>("s3", 64));
            s.add_subsystem(cs::subsystem_desc<Src>("s4", 192));
            s.add_subsystem(cs::fuse_components("f1", cs::subsystem_desc<TagStage>("a"), cs::subsystem_desc<TagStage>("b")));
            s.add_subsystem(cs::fuse_components("f2", cs::subsystem_desc<TagStage>("c"), cs::subsystem_desc<TagStage>("d")));
            s.connect("s1.y", "f1.a.x");
            s.connect("s2.y", "f1.b.x");
            s.connect("s3.y", "f2.c.x");
            s.connect("s4.y", "f2.d.x");
        });
        model.setup();
        for (int fwd = 0; fwd < 2; ++fwd) {
            model.run();
            model.wait();
            CHECK(same(out_of(model, "f1.a.y"), filled(200, 422)), "f1.a forward %d", fwd);
            CHECK(same(out_of(model, "f1.b.y"), filled(100, 1422)), "f1.b forward %d", fwd);
            CHECK(same(out_of(model, "f2.c.y"), filled(64, 132)), "f2.c forward %d", fwd);
            CHECK(same(out_of(model, "f2.d.y"), filled(192, 1132)), "f2.d forward %d", fwd);
        }
    }
    std::printf("group_waits_enabled keeps the segments' values\n");
    {
        cs::Model model = make([](cs::System& s) {
            s.add_subsystem(cs::subsystem_desc<Src>("sa", 200));
            s.add_subsystem(cs::subsystem_desc<Src>("sb", 100));
            s.add_subsystem(cs::fuse_components<cs::group_waits_enabled>("f", cs::segment(cs::subsystem_desc<TagStage>("a")),
                                                                         cs::segment(cs::subsystem_desc<TagStage>("b"))));
            s.connect("sa.y", "f.a.x");
            s.connect("sb.y", "f.b.x");
        });
        model.setup();
        for (int fwd = 0; fwd < 3; ++fwd) {
            model.run();
            model.wait();
            CHECK(same(out_of(model, "f.a.y"), filled(200, 422)), "waits a forward %d", fwd);
            CHECK(same(out_of(model, "f.b.y"), filled(100, 1422)), "waits b forward %d", fwd);
        }
    }
    std::printf("a global barrier keeps the segments' values (its launches see the whole segment table)\n");
    {
        cs::Model model = make([](cs::System& s) {
            s.add_subsystem(cs::subsystem_desc<Src>("sa", 200));
            s.add_subsystem(cs::subsystem_desc<Src>("sb", 100));
            s.add_subsystem(cs::fuse_components("f", cs::segment(cs::subsystem_desc<TagStage>("a")), cs::global_barrier(),
                                                cs::segment(cs::subsystem_desc<TagStage>("b"))));
            s.connect("sa.y", "f.a.x");
            s.connect("sb.y", "f.b.x");
        });
        model.setup();
        for (int fwd = 0; fwd < 3; ++fwd) {
            model.run();
            model.wait();
            CHECK(same(out_of(model, "f.a.y"), filled(200, 422)), "barrier a forward %d", fwd);
            CHECK(same(out_of(model, "f.b.y"), filled(100, 1422)), "barrier b forward %d", fwd);
        }
    }
    for (int big = 0; big < 2; ++big) {
        const std::size_t wg = big ? big_wg : 64, G = big ? big_groups : 5;
        std::printf("cross-group reads, %zu groups of %zu\n", G, wg);
        neighbor_case<0, 2>(wg, G, 5);  // global_barrier
        neighbor_case<1, 1>(wg, G, 5);  // group_waits_enabled + wait_group on the neighbour
        neighbor_case<1, 2>(wg, G, 5);  // wait_group in a group that has only a global barrier
        neighbor_case<2, 1>(wg, G, 5);  // wait_earlier_segments == a global barrier
    }
    if (g_fail) {
        std::printf("%d failure(s)\n", g_fail);
        return 1;
    }
    std::printf("OK\n");
    return 0;
}
