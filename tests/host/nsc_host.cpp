// tests/host/nsc_host.cpp — the compressible NS residual's host logic on the CPU (TEST INFRASTRUCTURE ONLY).
//
// Builds workloads/nsc_systems.hpp's NscResidual in its three variants through the real Model / System / Executor code,
// linked against the CPU implementation of the C boundary (mock_hip.cpp) with a no-op stand-in for every kernel: the ties
// of the atomic accumulators (System::tie_outputs and cs::ties_desc inside a fused group), the connections into the
// reductions and the error stage, every shape, and the launches per run of each variant, eager and replayed. Prints "OK"
// when every check holds.
#include <cstdio>
#include <string>
#include <vector>

#include "mock_hip.hpp"
#include "naga/cospan/cospan.hpp"
#include "nsc_systems.hpp"

namespace cs = naga::cospan;

static int g_failures = 0;
#define CHECK(cond, ...)                                     \
    do {                                                     \
        if (!(cond)) {                                       \
            ++g_failures;                                    \
            std::printf("FAIL %s:%d: ", __FILE__, __LINE__); \
            std::printf(__VA_ARGS__);                        \
            std::printf("\n");                               \
        }                                                    \
    } while (0)

int main() {
    mock::set_fallback_kernel([](const void*, const std::uint32_t*, const std::uint32_t*) {});
    struct Case {
        nsc::Variant v;
        const char*  name;
        int          launches;
    } cases[] = {{nsc::Variant::unfused, "unfused", 6}, {nsc::Variant::fused, "fused", 1}, {nsc::Variant::launches, "launches", 4}};
    const std::size_t nx = 256, ny = 64, tiles = (nx / 64) * (ny / 8);
    for (const Case& c : cases) {
        try {
            cs::Model m = cs::create_model<nsc::NscResidual<nsc::kMinMod, true, 8>>(c.v);
            m.setup();
            m.update_input("U", cs::init::fill(1.0f, cs::Shape{4, ny, nx}));
            const std::uint64_t l0 = mock::stats().launches;
            m.run();
            m.wait();
            const std::uint64_t l1 = mock::stats().launches;
            CHECK(static_cast<int>(l1 - l0) == c.launches, "%s: %llu launches in an eager run, the plan has %d", c.name,
                  (unsigned long long)(l1 - l0), c.launches);
            for (int i = 0; i < 3; ++i) m.run();  // captured, then replayed
            m.wait();
            const cs::ModelStats st = m.stats();
            CHECK(st.captures >= 1 && st.replays >= 2, "%s: captures %llu replays %llu", c.name, (unsigned long long)st.captures,
                  (unsigned long long)st.replays);
            const std::vector<std::size_t> state{4, ny, nx};
            CHECK(m.shape_of("R") == state, "%s: R", c.name);
            CHECK(m.shape_of("Part") == (std::vector<std::size_t>{6, tiles}), "%s: Part", c.name);
            for (const char* v : {"Dt", "Ws", "DetDt", "DetWs"}) CHECK(m.shape_of(v) == std::vector<std::size_t>{1}, "%s: %s", c.name, v);
            for (const char* v : {"Sq", "DetSq"}) CHECK(m.shape_of(v) == std::vector<std::size_t>{4}, "%s: %s", c.name, v);
            CHECK(m.shape_of("Err") == std::vector<std::size_t>{6}, "%s: Err", c.name);
            // one variable per tie: the accumulators the tile adds into are the ones the init resets and err reads
            const std::string p = c.v == nsc::Variant::unfused ? "" : "ns.";
            for (const char* t : {"Dt", "Ws", "Sq"}) {
                void* a = m.device_data(t);
                for (const char* member : {"init.", "tile.", "err."})
                    CHECK(m.device_data(p + member + t) == a, "%s: %s%s is not the tie's variable", c.name, member, t);
            }
            CHECK(m.device_data(p + "detn.Part") == m.device_data("Part") && m.device_data(p + "err.DetSq") == m.device_data("DetSq"),
                  "%s: the reductions do not read the tile's partials", c.name);
            std::printf("%s: %d launches per evaluation, captured and replayed\n", c.name, c.launches);
        } catch (const std::exception& e) {
            CHECK(false, "%s: %s", c.name, e.what());
        }
    }
    // a plain tie of outputs needs the allow_plain policy
    try {
        struct Top : cs::System {
            void setup() override {
                add_subsystem(cs::subsystem_desc<nsc::NscDet<1>>("a"));
                add_subsystem(cs::subsystem_desc<nsc::NscDet<2>>("b"));
                tie_outputs<cs::DType::f32, 1>({"a.Det", "b.Det"}, "D");
            }
        };
        cs::Model   m = cs::create_model<Top>();
        std::string what;
        try {
            m.setup();
        } catch (const std::exception& e) {
            what = e.what();
        }
        CHECK(what.find("allow_plain") != std::string::npos, "plain tie without allow_plain: %s", what.c_str());
    } catch (const std::exception& e) {
        CHECK(false, "plain tie: %s", e.what());
    }
    const std::string f = mock::failure();
    CHECK(f.empty(), "mock: %s", f.c_str());
    if (g_failures) {
        std::printf("%d failure(s)\n", g_failures);
        return 1;
    }
    std::printf("OK\n");
    return 0;
}
