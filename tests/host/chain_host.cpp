// tests/host/chain_host.cpp — the GEMVER -> BiCGK -> GEMVER chain's host logic on the CPU (TEST INFRASTRUCTURE ONLY).
//
// Builds workloads/chain_systems.hpp's Chain in each of its six styles through the real Model / System / Executor code,
// linked against the CPU implementation of the C boundary (mock_hip.cpp) with a no-op stand-in for every kernel: it checks
// the wiring (every input has a source or an initial value, every shape resolves) and the number of launches per run
// against the styles' plans (examples/gemver-chain/systems.hpp:33-53), eager and replayed, and which replays ReplayMode::direct
// serves as plain launches (single-stream captures) or as graph launches (the parallel style). Prints "OK" when every check
// holds.
#include <cstdio>
#include <string>

#include "mock_hip.hpp"
#include "naga/cospan/cospan.hpp"
#include "chain_systems.hpp"

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
    // launches of one chain: G1 + BiCGK + G2 (a fused group with global barriers is one launch per barrier-free run)
    struct Case {
        chain::Style style;
        const char*  name;
        int          launches;
    } cases[] = {{chain::Style::unfused, "unfused", 4 + 3 + 4}, {chain::Style::parallel, "parallel", 4 + 3 + 4},
                 {chain::Style::fused, "fused", 3 + 2 + 3},       {chain::Style::onepass, "onepass", 3 + 3 + 3},
                 {chain::Style::naive, "naive", 3 + 2 + 3},       {chain::Style::crafted, "crafted", 3 + 2 + 3}};
    for (const Case& c : cases) {
        try {
            cs::Model m = cs::create_model<chain::Chain>(c.style, 512, 64, 0u);
            m.setup();
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
            for (const char* v : {"x1", "w1", "q", "s", "x2", "w2"}) {
                const auto shape = m.shape_of(v);
                CHECK(shape.size() == 1 && shape[0] == 512, "%s: '%s' has a wrong shape", c.name, v);
            }
            for (const char* v : {"B1", "B2"}) {
                const auto shape = m.shape_of(v);
                CHECK(shape.size() == 2 && shape[0] == 512 && shape[1] == 512, "%s: '%s' has a wrong shape", c.name, v);
            }
            std::printf("%s: %d launches per chain, captured and replayed\n", c.name, c.launches);

            // ReplayMode::direct: a capture that stayed on stream 0 is re-issued as its launches (no graph launch); one that
            // spans streams (the parallel style's independent components) still replays its graph
            cs::Model d = cs::create_model<chain::Chain>(c.style, 512, 64, 0u);
            d.setup();
            d.set_replay_mode(cs::ReplayMode::direct);
            for (int i = 0; i < 3; ++i) d.run();  // eager, captured, replayed
            d.wait();
            const mock::Stats    s0 = mock::stats();
            const std::uint64_t  r0 = d.stats().relaunches;
            for (int i = 0; i < 2; ++i) d.run();  // replays
            d.wait();
            const mock::Stats    s1     = mock::stats();
            const std::uint64_t  r1     = d.stats().relaunches;
            const std::uint64_t  plain  = s1.launches - s0.launches, graphs = s1.graph_launches - s0.graph_launches;
            const bool           multi  = graphs != 0;
            // every replay is one or the other, whole: the chain's launches on stream 0, or one graph launch
            CHECK((r1 - r0 == 2 && graphs == 0 && static_cast<int>(plain) == 2 * c.launches) || (r1 == r0 && graphs == 2 && plain == 0),
                  "%s direct: relaunches %llu, graph launches %llu, launches %llu", c.name, (unsigned long long)(r1 - r0),
                  (unsigned long long)graphs, (unsigned long long)plain);
            // the parallel style runs its independent components on side streams: its capture spans streams
            CHECK(c.style != chain::Style::parallel || multi, "%s direct: the parallel style's capture should span streams", c.name);
            std::printf("%s direct: %s\n", c.name, multi ? "multi-stream capture, graph replays" : "relaunched on stream 0");
        } catch (const std::exception& e) {
            CHECK(false, "%s: %s", c.name, e.what());
        }
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
