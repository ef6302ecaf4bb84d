// tests/framework/range_split.cpp — the range_split strategies and launch shapes through cs::Model on the GPU (a test program).
//
// The cases and their closed-form expectations follow the reference's tests/test_range_split.cpp:569-667: more than 65535
// groups per stage (a fused group under flatten_and_split, pad_and_split and padded; a bare group-range and a bare
// index-range component), rank-3 stages (flatten, pad, a bare rank-3 component), rank-1 stages (pad), the physical launch
// of a bare component (its declared range) and the physical group coordinates each strategy promises, and padded refusing
// rank-3 stages. One more case: a stage 70000 groups tall, past the 65535 a hardware grid allows in y, under pad and padded
// (the launcher then flattens the groups along x; the logical coordinates stay the same). The components are this file's
// own. Prints one line per case and "OK" at the end when every check holds; exits 1 otherwise.
#include <cstdio>
#include <functional>
#include <stdexcept>
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

namespace trs {

using Build = std::function<void(cs::System&)>;
class Top : public cs::System {
    Build build_;

public:
    explicit Top(Build b) : build_{std::move(b)} {}
    void setup() override { build_(*this); }
};
cs::Model make(Build b) { return cs::create_model<Top>(std::move(b)); }

template <int R>
std::vector<float> out_of(cs::Model& m, const char* path) {
    auto v = m.view_output<DType::f32, R>(path);
    v.sync();
    return std::vector<float>(v.data(), v.data() + v.size());
}

/// The work-group shape of the rank-R stages: 8, 4 x 4, 2 x 2 x 2.
template <int R>
constexpr cs::range<R> wg_of() {
    if constexpr (R == 1)
        return cs::range<1>{8};
    else if constexpr (R == 2)
        return cs::range<2>{4, 4};
    else
        return cs::range<3>{2, 2, 2};
}

/// What IdStage<R, Seg> leaves over `groups`: at every element, Seg * 1e6 + the linear id of the group holding it (its
/// rank-R group coordinate, row-major over `groups`).
template <int R>
std::vector<float> expected_ids(cs::range<R> groups, int seg) {
    const cs::range<R> wg = wg_of<R>();
    std::size_t        extent[3] = {1, 1, 1}, total = 1;
    for (int d = 0; d < R; ++d) total *= (extent[d] = groups[d] * wg[d]);
    std::vector<float> v(total);
    std::size_t        idx[3] = {0, 0, 0};
    for (std::size_t flat = 0; flat < total; ++flat) {
        std::size_t rest = flat;
        for (int d = R - 1; d >= 0; --d) {
            idx[d] = rest % extent[d];
            rest /= extent[d];
        }
        std::size_t lin = 0;
        for (int d = 0; d < R; ++d) lin = lin * groups[d] + idx[d] / wg[d];
        v[flat] = static_cast<float>(seg * 1000000 + static_cast<long long>(lin));
    }
    return v;
}

template <int R>
cs::range<R> scaled(cs::range<R> groups) {
    cs::range<R> r = groups;
    for (int d = 0; d < R; ++d) r[d] *= wg_of<R>()[d];
    return r;
}

CS_COMPONENTS_BEGIN

/// Every work-item writes Seg * 1e6 + its group's linear id at its global id (segment-relative group ids).
template <int R, int Seg>
class IdStage : public cs::Component<IdStage<R, Seg>> {
public:
    struct Inputs {};
    struct Outputs {
        cs::Output<DType::f32, R> y;
    };
    static void compute_forward(cs::GroupItem<R> it, const Inputs&, const Outputs& out) {
        out.y[it.get_global_id()] = static_cast<float>(Seg * 1000000 + static_cast<long long>(it.get_group_linear_id()));
    }
    Inputs  inputs;
    Outputs outputs;
    explicit IdStage(cs::range<R> groups) : groups_{groups} {}
    void declare() override {
        const cs::range<R> e = scaled<R>(groups_);
        if constexpr (R == 1)
            this->declare_output(outputs.y, "y", static_cast<int>(e[0]));
        else if constexpr (R == 2)
            this->declare_output(outputs.y, "y", static_cast<int>(e[0]), static_cast<int>(e[1]));
        else
            this->declare_output(outputs.y, "y", static_cast<int>(e[0]), static_cast<int>(e[1]), static_cast<int>(e[2]));
    }
    std::unique_ptr<cs::LaunchDescription> define_forward() override {
        return this->create_launch_config(cs::range_desc(scaled<R>(groups_)), cs::wg_desc(wg_of<R>()));
    }

private:
    cs::range<R> groups_;
};

/// An index-range rank-2 stage: y[i][j] = i * width + j.
class IndexId : public cs::Component<IndexId> {
public:
    struct Inputs {};
    struct Outputs {
        cs::Output<DType::f32, 2> y;
    };
    static void compute_forward(cs::id<2> i, const Inputs&, const Outputs& out) {
        out.y[i] = static_cast<float>(i[0] * out.y.shape()[1] + i[1]);
    }
    Inputs  inputs;
    Outputs outputs;
    explicit IndexId(cs::range<2> r) : r_{r} {}
    void declare() override { declare_output(outputs.y, "y", static_cast<int>(r_[0]), static_cast<int>(r_[1])); }
    std::unique_ptr<cs::LaunchDescription> define_forward() override { return create_launch_config(cs::range_desc(r_)); }

private:
    cs::range<2> r_;
};

/// 1 where the work-group's PHYSICAL coordinates (blockIdx: physical dimension 0 is z, 1 is y, 2 is x) are the ones the
/// launch kind K promises for its logical group (gy, gx); 0 elsewhere.
///   K = 0 flatten_and_split: (0, 0, the group's linear id) for segment 0 (the later segments' ids follow on along x)
///   K = 1 pad_and_split:     (0, gy, anything)
///   K = 2 padded:            (segment, gy, gx)
///   K = 3 a bare component:  (0, gy, gx), its declared range
template <int K>
class RawStage : public cs::Component<RawStage<K>> {
public:
    struct Inputs {};
    struct Outputs {
        cs::Output<DType::f32, 2> y;
    };
    static void compute_forward(cs::GroupItem<2> it, const Inputs&, const Outputs& out) {
        const std::size_t p0 = blockIdx.z, p1 = blockIdx.y, p2 = blockIdx.x;
        const cs::id<2>   g  = it.get_group_id();
        bool              ok = false;
        if constexpr (K == 0)
            ok = p0 == 0 && p1 == 0 && (it.get_segment() != 0 || p2 == it.get_group_linear_id());
        else if constexpr (K == 1)
            ok = p0 == 0 && p1 == g[0];
        else if constexpr (K == 2)
            ok = p0 == it.get_segment() && p1 == g[0] && p2 == g[1];
        else
            ok = p0 == 0 && p1 == g[0] && p2 == g[1];
        out.y[it.get_global_id()] = ok ? 1.0f : 0.0f;
    }
    Inputs  inputs;
    Outputs outputs;
    explicit RawStage(cs::range<2> groups) : groups_{groups} {}
    void declare() override { this->declare_output(outputs.y, "y", static_cast<int>(groups_[0] * 4), static_cast<int>(groups_[1] * 4)); }
    std::unique_ptr<cs::LaunchDescription> define_forward() override {
        return this->create_launch_config(cs::range_desc(scaled<2>(groups_)), cs::wg_desc(wg_of<2>()));
    }

private:
    cs::range<2> groups_;
};

CS_COMPONENTS_END

template <class Split>
const char* name_of() {
    if constexpr (std::is_same_v<Split, cs::range_split::flatten_and_split>)
        return "flatten_and_split";
    else if constexpr (std::is_same_v<Split, cs::range_split::pad_and_split>)
        return "pad_and_split";
    else
        return "padded";
}

/// Two rank-2 stages of `groups` groups each in one fused group under Split.
template <class Split>
void many_groups_case(cs::range<2> groups) {
    cs::Model m = make([groups](cs::System& s) {
        s.add_subsystem(cs::fuse_components<Split>("f", cs::subsystem_desc<IdStage<2, 0>>("a", groups), cs::subsystem_desc<IdStage<2, 1>>("b", groups)));
    });
    m.setup();
    for (int fwd = 0; fwd < 3; ++fwd) {  // eager, captured, replayed
        m.run();
        m.wait();
    }
    CHECK(out_of<2>(m, "f.a.y") == expected_ids<2>(groups, 0), "%s: %zu x %zu groups, stage a", name_of<Split>(), groups[0], groups[1]);
    CHECK(out_of<2>(m, "f.b.y") == expected_ids<2>(groups, 1), "%s: %zu x %zu groups, stage b", name_of<Split>(), groups[0], groups[1]);
}

/// Rank-3 stages (2, 3, 4) and (3, 1, 5) groups in one fused group under Split, and a bare rank-3 component of (2, 2, 3).
template <class Split>
void rank3_case() {
    cs::Model m = make([](cs::System& s) {
        s.add_subsystem(cs::fuse_components<Split>("f", cs::subsystem_desc<IdStage<3, 0>>("a", cs::range<3>{2, 3, 4}),
                                                   cs::subsystem_desc<IdStage<3, 1>>("b", cs::range<3>{3, 1, 5})));
        s.add_subsystem(cs::subsystem_desc<IdStage<3, 2>>("c", cs::range<3>{2, 2, 3}));
    });
    m.setup();
    m.run();
    m.wait();
    CHECK(out_of<3>(m, "f.a.y") == expected_ids<3>(cs::range<3>{2, 3, 4}, 0), "%s: rank-3 stage a", name_of<Split>());
    CHECK(out_of<3>(m, "f.b.y") == expected_ids<3>(cs::range<3>{3, 1, 5}, 1), "%s: rank-3 stage b", name_of<Split>());
    CHECK(out_of<3>(m, "c.y") == expected_ids<3>(cs::range<3>{2, 2, 3}, 2), "%s: bare rank-3 component", name_of<Split>());
}

/// Every element 1: the physical coordinates of every group are the ones kind K promises.
template <class Split, int K>
void physical_case() {
    cs::Model m = make([](cs::System& s) {
        s.add_subsystem(cs::fuse_components<Split>("f", cs::subsystem_desc<RawStage<K>>("a", cs::range<2>{6, 9}),
                                                   cs::subsystem_desc<RawStage<K>>("b", cs::range<2>{2, 3})));
    });
    m.setup();
    m.run();
    m.wait();
    CHECK(out_of<2>(m, "f.a.y") == std::vector<float>(24u * 36u, 1.0f), "%s: physical coordinates of stage a", name_of<Split>());
    CHECK(out_of<2>(m, "f.b.y") == std::vector<float>(8u * 12u, 1.0f), "%s: physical coordinates of stage b", name_of<Split>());
}

}  // namespace trs

using namespace trs;
using Flatten = cs::range_split::flatten_and_split;
using Pad     = cs::range_split::pad_and_split;
using Padded  = cs::range_split::padded;

#ifndef CS_SOURCE_HASH
#define CS_SOURCE_HASH "unknown"
#endif

int main() {
    std::printf("build %s\n", CS_SOURCE_HASH);

    std::printf("260 x 260 = 67600 groups per stage in a fused group, per strategy\n");
    many_groups_case<Flatten>(cs::range<2>{260, 260});
    many_groups_case<Pad>(cs::range<2>{260, 260});
    many_groups_case<Padded>(cs::range<2>{260, 260});

    std::printf("a stage 70000 groups tall (past a hardware grid's 65535 in y) under pad and padded: flattened along x\n");
    many_groups_case<Pad>(cs::range<2>{70000, 2});
    many_groups_case<Padded>(cs::range<2>{70000, 2});

    std::printf("67600 groups in a bare rank-2 group-range component and a bare index-range component\n");
    {
        cs::Model m = make([](cs::System& s) {
            s.add_subsystem(cs::subsystem_desc<IdStage<2, 0>>("a", cs::range<2>{260, 260}));
            s.add_subsystem(cs::subsystem_desc<IndexId>("p", cs::range<2>{2080, 2080}));
        });
        m.setup();
        m.run();
        m.wait();
        CHECK(out_of<2>(m, "a.y") == expected_ids<2>(cs::range<2>{260, 260}, 0), "bare rank-2 component");
        const std::vector<float> p  = out_of<2>(m, "p.y");
        bool                     ok = p.size() == 2080u * 2080u;
        for (std::size_t k = 0; ok && k < p.size(); ++k) ok = p[k] == static_cast<float>(k);
        CHECK(ok, "bare index-range component: y[i][j] = i * 2080 + j");
    }

    std::printf("rank-3 stages under flatten_and_split and pad_and_split, and a bare rank-3 component\n");
    rank3_case<Flatten>();
    rank3_case<Pad>();

    std::printf("rank-1 stages under pad_and_split\n");
    {
        cs::Model m = make([](cs::System& s) {
            s.add_subsystem(cs::fuse_components<Pad>("f", cs::subsystem_desc<IdStage<1, 0>>("a", cs::range<1>{9}),
                                                     cs::subsystem_desc<IdStage<1, 1>>("b", cs::range<1>{4})));
        });
        m.setup();
        m.run();
        m.wait();
        CHECK(out_of<1>(m, "f.a.y") == expected_ids<1>(cs::range<1>{9}, 0), "rank-1 stage a");
        CHECK(out_of<1>(m, "f.b.y") == expected_ids<1>(cs::range<1>{4}, 1), "rank-1 stage b");
    }

    std::printf("a bare component's physical launch is its declared range\n");
    {
        cs::Model m = make([](cs::System& s) { s.add_subsystem(cs::subsystem_desc<RawStage<3>>("a", cs::range<2>{6, 9})); });
        m.setup();
        m.run();
        m.wait();
        CHECK(out_of<2>(m, "a.y") == std::vector<float>(24u * 36u, 1.0f), "bare component: physical = logical group coordinates");
    }

    std::printf("the physical group coordinates of flatten_and_split, pad_and_split and padded\n");
    physical_case<Flatten, 0>();
    physical_case<Pad, 1>();
    physical_case<Padded, 2>();

    std::printf("padded refuses rank-3 stages\n");
    {
        bool        threw = false;
        std::string msg;
        try {
            cs::Model m = make([](cs::System& s) {
                s.add_subsystem(cs::fuse_components<Padded>("f", cs::subsystem_desc<IdStage<3, 0>>("a", cs::range<3>{2, 3, 4}),
                                                            cs::subsystem_desc<IdStage<3, 1>>("b", cs::range<3>{3, 1, 5})));
            });
            m.setup();
            m.run();
            m.wait();
        } catch (const std::invalid_argument& e) {
            threw = true;
            msg   = e.what();
        }
        CHECK(threw && msg.find("rank 2 or lower") != std::string::npos, "padded with rank-3 stages: threw %d (%s)", threw ? 1 : 0, msg.c_str());
    }

    if (g_fail) {
        std::printf("%d failure(s)\n", g_fail);
        return 1;
    }
    std::printf("OK\n");
    return 0;
}
