// tests/framework/policies.cpp — run policies, rearm and depends_on through cs::Model (a test program).
//
// Built twice: for the GPU (make -C tests/framework), where every check runs, and for the CPU against the mock boundary
// (make -C tests/host, -DCS_HOST_MOCK), where the side-channel kernels run as CPU emulations, the other kernels are no-ops
// and the streams drain in an adversarial order (so a missing depends_on edge shows as a stale read every time).
//
// The cases follow the reference's tests/test_r04.cpp:448-540 and :673-730 (run_n, do_once, transient on leaves, rearm,
// a subtree out of scope skipped whole with its policies, transient producers and readers releasing their own memory,
// views of released variables refused) and tests/test_system.cpp:3237-3275 (depends_on through a side channel with no
// variable between the two subsystems, declaration errors), with this file's own components; and the side channel over
// captured and replayed runs. Prints one line per case and "OK" at the end when every check holds; exits 1 otherwise.
#include <cstdint>
#include <cstdio>
#include <functional>
#include <string>
#include <vector>

#include <naga/cospan/cospan.hpp>

#ifdef CS_HOST_MOCK
#include "mock_hip.hpp"
constexpr bool kData = false;  // the components' kernels are not emulated: their outputs are not looked at
#else
constexpr bool kData = true;
#endif

namespace cs = naga::cospan;
using cs::DType;

static int g_fail = 0;
#define CHECK(cond, ...)                               \
    do {                                               \
        if (!(cond)) {                                 \
            ++g_fail;                                  \
            std::printf("  FAIL line %d: ", __LINE__); \
            std::printf(__VA_ARGS__);                  \
            std::printf("\n");                         \
        }                                              \
    } while (0)

namespace tpol {

constexpr std::size_t kN = 64;

void expect_throws(const std::function<void()>& f, const std::vector<std::string>& needles, const char* what) {
    std::string msg = "<no exception>";
    try {
        f();
    } catch (const std::exception& e) {
        msg = e.what();
    }
    bool all = true;
    for (const std::string& n : needles) all = all && msg.find(n) != std::string::npos;
    CHECK(all, "%s: message: %s", what, msg.c_str());
}
template <class E>
bool throws_as(const std::function<void()>& f) {
    try {
        f();
    } catch (const E&) {
        return true;
    } catch (...) {
    }
    return false;
}

/// The side channel of the depends_on case: one float in device memory that no variable names.
float* g_side = nullptr;

/// The one argument of the side-channel kernels.
struct SideArgs {
    float* side;
    float* got;
};
/// A writer that is slow on purpose (a bounded sleep of about a millisecond), then adds 1 to the side channel.
__global__ void slow_increment(SideArgs a) {
    for (int i = 0; i < 300; ++i) __builtin_amdgcn_s_sleep(127);
    if (threadIdx.x == 0) a.side[0] += 1.0f;
}
/// got[0] = the side channel.
__global__ void copy_side(SideArgs a) {
    if (threadIdx.x == 0) a.got[0] = a.side[0];
}

void launch(const void* k, void** args, cs_hip_stream s) {
    const std::uint32_t grid[3] = {1, 1, 1}, block[3] = {64, 1, 1};
    cs::detail::hip_check(cs_hip_launch(k, grid, block, 0, args, s), "side-channel launch");
}

CS_COMPONENTS_BEGIN

/// y = v (v a free input, initialized to `value`).
class Gen : public cs::Component<Gen> {
public:
    struct Inputs {
        cs::Input<DType::f32, 1> v;
    };
    struct Outputs {
        cs::Output<DType::f32, 1> y;
    };
    static void compute_forward(cs::id<1> i, const Inputs& in, const Outputs& out) { out.y[i] = in.v[i]; }
    Inputs  inputs;
    Outputs outputs;
    explicit Gen(float value) : value_{value} {}
    void declare() override {
        declare_input(inputs.v, "v", kN);
        init_input("v", cs::init::fill(value_));
        declare_output(outputs.y, "y", kN);
    }
    std::unique_ptr<cs::LaunchDescription> define_forward() override { return create_launch_config(cs::range_desc(outputs.y.shape())); }

private:
    float value_;
};

/// y = x + 1.
class Inc : public cs::Component<Inc> {
public:
    struct Inputs {
        cs::Input<DType::f32, 1> x;
    };
    struct Outputs {
        cs::Output<DType::f32, 1> y;
    };
    static void compute_forward(cs::id<1> i, const Inputs& in, const Outputs& out) { out.y[i] = in.x[i] + 1.0f; }
    Inputs  inputs;
    Outputs outputs;
    void    declare() override {
        declare_input(inputs.x, "x", kN);
        declare_output(outputs.y, "y", kN);
    }
    std::unique_ptr<cs::LaunchDescription> define_forward() override { return create_launch_config(cs::range_desc(outputs.y.shape())); }
};

/// A: the slow writer of the side channel (reads x, which only places it after its producer; writes a token nobody reads).
class SideWrite : public cs::CustomComponent<SideWrite> {
public:
    struct Inputs {
        cs::Input<DType::f32, 1> x;
    };
    struct Outputs {
        cs::Output<DType::f32, 1> token;
    };
    static void compute_forward(cs::CustomOp& op, const Inputs&, const Outputs&) {
        SideArgs a{g_side, nullptr};
        void*    args[] = {&a};
        launch(reinterpret_cast<const void*>(&slow_increment), args, op.stream);
    }
    Inputs  inputs;
    Outputs outputs;
    void    declare() override {
        declare_input(inputs.x, "x", kN);
        declare_output(outputs.token, "token", 1);
    }
};

/// B: reads the side channel into its output (reads x too: as the second reader of x it goes on another stream than A).
class SideRead : public cs::CustomComponent<SideRead> {
public:
    struct Inputs {
        cs::Input<DType::f32, 1> x;
    };
    struct Outputs {
        cs::Output<DType::f32, 1> got;
    };
    static void compute_forward(cs::CustomOp& op, const Inputs&, const Outputs& out) {
        SideArgs a{g_side, cs::raw::data(out.got)};
        void*    args[] = {&a};
        launch(reinterpret_cast<const void*>(&copy_side), args, op.stream);
    }
    Inputs  inputs;
    Outputs outputs;
    void    declare() override {
        declare_input(inputs.x, "x", kN);
        declare_output(outputs.got, "got", 1);
    }
};

CS_COMPONENTS_END

using Build = std::function<void(cs::System&)>;
class Top : public cs::System {
    Build build_;

public:
    explicit Top(Build b) : build_{std::move(b)} {}
    void setup() override { build_(*this); }
};
cs::Model make(Build b) { return cs::create_model<Top>(std::move(b)); }

/// A System of two leaves: g (Gen 1) -> c (Inc).
class Pair : public cs::System {
public:
    void setup() override {
        add_subsystem(cs::subsystem_desc<Gen>("g", 1.0f));
        add_subsystem(cs::subsystem_desc<Inc>("c"));
        connect("g.y", "c.x");
    }
};

void run(cs::Model& m) {
    m.run();
    m.wait();
}
bool all_equal(cs::Model& m, const char* path, float v) {
    auto h = m.view_output<DType::f32, 1>(path);
    h.sync();
    if (!kData) return h.size() == kN;
    for (std::size_t i = 0; i < h.size(); ++i)
        if (h[i] != v) return false;
    return h.size() == kN;
}
float side_got(cs::Model& m) {
    auto h = m.view_output<DType::f32, 1>("b.got");
    h.sync();
    return h[0];
}

}  // namespace tpol

using namespace tpol;

#ifndef CS_SOURCE_HASH
#define CS_SOURCE_HASH "unknown"
#endif

int main() {
    std::printf("build %s\n", CS_SOURCE_HASH);
#ifdef CS_HOST_MOCK
    mock::set_fallback_kernel([](const void*, const std::uint32_t*, const std::uint32_t*) {});
    mock::register_kernel(reinterpret_cast<const void*>(&slow_increment), sizeof(SideArgs), [](const void* a, const std::uint32_t*, const std::uint32_t*) {
        static_cast<const SideArgs*>(a)->side[0] += 1.0f;
    });
    mock::register_kernel(reinterpret_cast<const void*>(&copy_side), sizeof(SideArgs), [](const void* a, const std::uint32_t*, const std::uint32_t*) {
        const auto* c = static_cast<const SideArgs*>(a);
        c->got[0]     = c->side[0];
    });
#endif
    cs::detail::hip_check(cs_hip_init(0), "init");
    cs::detail::hip_check(cs_hip_malloc(reinterpret_cast<void**>(&g_side), 256), "side channel");
    cs_hip_stream s0 = nullptr;
    cs::detail::hip_check(cs_hip_stream_create(&s0), "stream");

    std::printf("run_n{2}: in scope for two turns, the reader keeps the last value\n");
    {
        cs::Model m = make([](cs::System& s) {
            s.add_subsystem(cs::property_list{cs::property::run_n{2}}, cs::subsystem_desc<Gen>("g", 1.0f));
            s.add_subsystem(cs::subsystem_desc<Inc>("c"));
            s.connect("g.y", "c.x");
        });
        m.setup();
        run(m);
        CHECK(m.in_scope("g"), "g after one turn");
        run(m);
        CHECK(!m.in_scope("g"), "g after two turns");
        run(m);
        run(m);  // c alone: stable, captured and replayed
        CHECK(all_equal(m, "c.y", 2.0f), "c.y after g left");
        CHECK(m.stats().replays >= 1, "the remaining leaf is replayed");
    }
    std::printf("do_once, then rearm_subsystem runs it once more with the new value\n");
    {
        cs::Model m = make([](cs::System& s) {
            s.add_subsystem(cs::property_list{cs::property::do_once{}}, cs::subsystem_desc<Gen>("g", 1.0f));
            s.add_subsystem(cs::subsystem_desc<Inc>("c"));
            s.connect("g.y", "c.x");
        });
        m.setup();
        run(m);
        CHECK(!m.in_scope("g"), "g after its one turn");
        m.update_input("g.v", cs::init::fill(7.0f, cs::Shape{kN}));
        run(m);
        CHECK(all_equal(m, "c.y", 2.0f), "g did not run: the update has no effect yet");
        m.rearm_subsystem("g");
        CHECK(m.in_scope("g"), "rearmed");
        run(m);
        CHECK(all_equal(m, "c.y", 8.0f), "c.y after the rearmed turn");
        CHECK(!m.in_scope("g"), "g after its rearmed turn");
        CHECK(throws_as<std::invalid_argument>([&] { m.rearm_subsystem("zz"); }), "rearm of an unknown subsystem");
    }
    std::printf("transient: out of scope after every run, back in for the next\n");
    {
        cs::Model m = make([](cs::System& s) {
            s.add_subsystem(cs::property_list{cs::property::transient{}}, cs::subsystem_desc<Gen>("g", 1.0f));
            s.add_subsystem(cs::subsystem_desc<Inc>("c"));
            s.connect("g.y", "c.x");
        });
        m.setup();
        for (int k = 0; k < 3; ++k) {
            run(m);
            CHECK(!m.in_scope("g"), "g after run %d", k);
            CHECK(all_equal(m, "c.y", 2.0f), "c.y after run %d", k);
            CHECK(m.device_data("g.v") == nullptr, "g's own input is released after run %d", k);
            CHECK(m.device_data("c.x") != nullptr, "what c holds stays after run %d", k);
        }
    }
    std::printf("a subtree out of scope is skipped whole, its policies included\n");
    {
        cs::Model m = make([](cs::System& s) {
            s.add_subsystem(cs::property_list{cs::property::run_n{1}}, cs::subsystem_desc<Pair>("s"));
            s.set_run_policy("s.c", cs::property_list{cs::property::run_n{3}});
        });
        m.setup();
        run(m);
        CHECK(!m.in_scope("s") && !m.in_scope("s.c"), "s and s.c after one turn");
        run(m);
        run(m);
        CHECK(throws_as<std::logic_error>([&] {
                  auto v = m.view_output<DType::f32, 1>("s.c.y");
                  v.sync();
              }),
              "a view of a released variable");
        m.clear_run_policy("s");
        m.rearm_subsystem("s");
        run(m);
        CHECK(m.in_scope("s") && m.in_scope("s.c"), "s and s.c rearmed (s.c's counter starts over)");
        CHECK(all_equal(m, "s.c.y", 2.0f), "s.c.y after rearming s");
    }
    std::printf("A transient, B eternal, C transient: A and C release their own memory, B's stays\n");
    {
        const cs::property_list tr{cs::property::transient{}};
        cs::Model               m = make([tr](cs::System& s) {
            s.add_subsystem(tr, cs::subsystem_desc<Gen>("a", 1.0f));
            s.add_subsystem(cs::subsystem_desc<Inc>("b"));
            s.add_subsystem(tr, cs::subsystem_desc<Inc>("c"));
            s.connect("a.y", "b.x");
            s.connect("b.y", "c.x");
        });
        m.setup();
        for (int k = 1; k <= 3; ++k) {
            run(m);
            CHECK(!m.in_scope("a") && m.in_scope("b") && !m.in_scope("c"), "scopes after run %d", k);
            CHECK(m.device_data("a.v") == nullptr && m.device_data("c.y") == nullptr, "released after run %d", k);
            CHECK(m.device_data("b.x") != nullptr && m.device_data("b.y") != nullptr, "kept after run %d", k);
            CHECK(all_equal(m, "b.y", 2.0f), "b.y after run %d", k);
            CHECK(throws_as<std::logic_error>([&] {
                      auto v = m.view_output<DType::f32, 1>("c.y");
                      v.sync();
                  }),
                  "a view of c.y after run %d", k);
        }
    }
    std::printf("policy errors\n");
    {
        expect_throws([] { cs::property_list p{cs::property::run_n{0}}; }, {"run_n needs n >= 1"}, "run_n{0}");
        expect_throws([] { cs::property_list p{cs::property::run_n{2}, cs::property::do_once{}}; }, {"two run policies"}, "run_n and do_once");
        cs::Model m = make([](cs::System& s) {
            s.add_subsystem(cs::property_list{cs::property::transient{}, cs::property::do_once{}}, cs::subsystem_desc<Gen>("g", 1.0f));
        });
        expect_throws([&] { m.setup(); }, {"transient cannot be combined"}, "transient with do_once");
        cs::Model m2 = make([](cs::System& s) {
            s.add_subsystem(cs::subsystem_desc<Gen>("g", 1.0f));
            s.set_run_policy("zz", cs::property_list{cs::property::do_once{}});
        });
        expect_throws([&] { m2.setup(); }, {"set_run_policy", "no subsystem 'zz'"}, "a policy on an unknown subsystem");
    }
    std::printf("depends_on: B reads what A wrote through a side channel, no variable between them\n");
    for (bool edge : {true, false}) {
        const float zero = 0.0f;
        cs::detail::hip_check(cs_hip_memcpy_async(g_side, &zero, 4, s0), "side reset");
        cs::detail::hip_check(cs_hip_stream_synchronize(s0), "side reset");
        cs::Model m = make([edge](cs::System& s) {
            s.add_subsystem(cs::subsystem_desc<Gen>("g", 1.0f));
            s.add_subsystem(cs::subsystem_desc<SideWrite>("a"));
            s.add_subsystem(cs::subsystem_desc<SideRead>("b"));
            s.connect("g.y", "a.x");
            s.connect("g.y", "b.x");
            if (edge) {
                s.depends_on("b", "a");
                s.depends_on("b", "a");  // a repeated edge is ignored
            }
        });
        m.setup();
        if (edge) {
            // five runs: eager, eager, captured, replayed, replayed; B sees A's increment of the same run every time
            for (int k = 1; k <= 5; ++k) {
                run(m);
                CHECK(side_got(m) == static_cast<float>(k), "run %d with the edge: b read %g", k, side_got(m));
            }
            CHECK(m.stats().captures >= 1 && m.stats().replays >= 2, "captured and replayed");
        } else {
#ifdef CS_HOST_MOCK
            // without the edge the mock drains b's stream first: b reads the side channel before a writes it
            run(m);
            CHECK(side_got(m) == 0.0f, "without the edge b ran first in the mock (read %g)", side_got(m));
#endif
        }
    }
    std::printf("depends_on: declaration errors\n");
    {
        cs::Model m = make([](cs::System& s) {
            s.add_subsystem(cs::subsystem_desc<Gen>("a", 1.0f));
            s.add_subsystem(cs::subsystem_desc<Gen>("b", 2.0f));
            s.depends_on("a", "b");
        });
        expect_throws([&] { m.setup(); }, {"depends_on", "can depend only on a subsystem added before it"}, "a later subsystem");
        cs::Model m2 = make([](cs::System& s) {
            s.add_subsystem(cs::subsystem_desc<Gen>("a", 1.0f));
            s.depends_on("a", "a");
        });
        expect_throws([&] { m2.setup(); }, {"can depend only on a subsystem added before it"}, "itself");
        cs::Model m3 = make([](cs::System& s) {
            s.add_subsystem(cs::subsystem_desc<Gen>("a", 1.0f));
            s.depends_on("a", "zz");
        });
        expect_throws([&] { m3.setup(); }, {"depends_on", "'zz' is not a subsystem"}, "an unknown subsystem");
    }
    cs::detail::hip_check(cs_hip_free(g_side), "side channel");
    cs::detail::hip_check(cs_hip_stream_destroy(s0), "stream");
#ifdef CS_HOST_MOCK
    CHECK(mock::failure().empty(), "mock: %s", mock::failure().c_str());
#endif
    if (g_fail) {
        std::printf("%d checks failed\n", g_fail);
        return 1;
    }
    std::printf("OK\n");
    return 0;
}
