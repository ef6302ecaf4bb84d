// The shared L2 write-back of a work-group's stores (naga/cospan/kernel.hpp: publish_stores, write_back_step), simulated
// on the CPU. Work-groups on several XCDs publish their stores one job after another: a job's stores reach its XCD's L2,
// the group reads s0 = started, then polls with write_back_step; an `issue` claims the next write-back number by
// compare-and-swap, and the write-back itself is two more steps (it starts: it takes every store in the L2 at that moment;
// it completes: those stores are in memory), after which the issuer sets completed. A seeded generator picks which
// work-group steps next, so every interleaving of polls, claims, write-back starts and completions and store arrivals is
// tried. Checks: when a job is published, every one of its stores is in memory; the counters keep
// completed <= started <= completed + 1 (one write-back in flight per XCD); some work-group can always step until every
// job is done. The same simulation with a planted mistake in the decision (covered once completed >= s0) must fail the
// first check: the check is not vacuous.

#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <functional>
#include <set>
#include <vector>

#include "naga/cospan/cospan.hpp"

namespace cs = naga::cospan;
using cs::detail::write_back_step;
using cs::detail::WriteBackStep;

static int g_failures = 0;
#define CHECK(cond, ...)                                                                                    \
    do {                                                                                                    \
        if (!(cond)) {                                                                                      \
            ++g_failures;                                                                                   \
            std::printf("FAIL %s:%d: ", __FILE__, __LINE__);                                                \
            std::printf(__VA_ARGS__);                                                                       \
            std::printf("\n");                                                                              \
        }                                                                                                   \
    } while (0)

static std::uint64_t g_rng = 0x2545f4914f6cdd1dull;
static std::uint32_t next_rand() {
    g_rng ^= g_rng << 13;
    g_rng ^= g_rng >> 7;
    g_rng ^= g_rng << 17;
    return static_cast<std::uint32_t>(g_rng >> 16);
}

struct Xcd {
    std::uint32_t     started = 0, completed = 0;
    std::set<int>     l2;              ///< stores in the L2 not yet written back
    std::set<int>     in_flight;       ///< the stores the write-back in flight took when it started
};

enum class At { arrive, read_s0, poll_d, poll_s, wb_start, wb_done, set_completed, done };

struct Group {
    int           xcd = 0, jobs_left = 0;
    At            at  = At::arrive;
    std::uint32_t s0 = 0, d = 0, claimed = 0;
    std::uint64_t blocked = ~0ull;  ///< the counters' version this group saw when it decided to wait
    std::vector<int> job;  ///< the stores of the current job
};

using Decide = std::function<WriteBackStep(std::uint32_t, std::uint32_t, std::uint32_t)>;

/// One simulation: `nx` XCDs, `per` work-groups on each, `jobs` publications per work-group of 1..3 stores each. Returns
/// the number of published jobs whose stores were not all in memory (0 is correct).
static int simulate(int nx, int per, int jobs, const Decide& decide, std::uint64_t& steps_out) {
    std::vector<Xcd>   x(nx);
    std::vector<Group> g(nx * per);
    std::set<int>      memory;
    int                store_id = 0, stale = 0, done = 0;
    for (int i = 0; i < nx * per; ++i) {
        g[i].xcd       = i % nx;
        g[i].jobs_left = jobs;
    }
    std::uint64_t steps = 0, version = 0;  // version: bumped whenever a counter changes
    auto step = [&](Group& w) -> bool {
        Xcd& X = x[w.xcd];
        switch (w.at) {
            case At::arrive: {  // the job's stores reach the L2 (every wave waited for them before the barrier)
                w.job.clear();
                const int k = 1 + static_cast<int>(next_rand() % 3);
                for (int i = 0; i < k; ++i) {
                    w.job.push_back(store_id);
                    X.l2.insert(store_id++);
                }
                w.at = At::read_s0;
                return true;
            }
            case At::read_s0:
                w.s0 = X.started;
                w.at = At::poll_d;
                return true;
            case At::poll_d:  // the device's first load of a poll: completed (a covered poll reads nothing else)
                w.d = X.completed;
                if (decide(w.s0, w.d, w.d) == WriteBackStep::covered) {
                    for (int id : w.job) stale += memory.count(id) ? 0 : 1;
                    w.at = --w.jobs_left > 0 ? At::arrive : At::done;
                    done += w.at == At::done;
                    return true;
                }
                w.at = At::poll_s;
                return true;
            case At::poll_s: {  // the second load, started, read later than completed; then the decision
                const std::uint32_t s    = X.started;
                const WriteBackStep step = decide(w.s0, w.d, s);
                w.at                     = At::poll_d;
                if (step != WriteBackStep::issue || X.started != s) {  // wait, or the compare-and-swap loses: poll again,
                    if (w.d == X.completed) w.blocked = version;        // after a counter changes if what it read is current
                    return true;
                }
                w.claimed = ++X.started;
                ++version;
                w.at = At::wb_start;
                return true;
            }
            case At::wb_start:  // the write-back takes every dirty line of the L2
                X.in_flight = X.l2;
                X.l2.clear();
                w.at = At::wb_done;
                return true;
            case At::wb_done:  // ... and has written them to memory when the fence returns
                memory.insert(X.in_flight.begin(), X.in_flight.end());
                X.in_flight.clear();
                w.at = At::set_completed;
                return true;
            case At::set_completed:
                X.completed = w.claimed;
                ++version;
                // the issuer's own job: its write-back started after the job's stores arrived
                for (int id : w.job) stale += memory.count(id) ? 0 : 1;
                w.at = --w.jobs_left > 0 ? At::arrive : At::done;
                done += w.at == At::done;
                return true;
            default: return false;
        }
    };
    std::vector<int> can;
    while (done < static_cast<int>(g.size())) {
        for (const Xcd& X : x)
            CHECK(X.completed <= X.started && X.started <= X.completed + 1, "counters: started %u completed %u", X.started, X.completed);
        can.clear();
        for (int i = 0; i < static_cast<int>(g.size()); ++i)
            if (g[i].at != At::done && g[i].blocked != version) can.push_back(i);
        bool moved = false;
        while (!can.empty() && !moved) {
            const std::size_t pick = next_rand() % can.size();
            moved                  = step(g[can[pick]]);
            if (!moved) can.erase(can.begin() + static_cast<std::ptrdiff_t>(pick));
        }
        ++steps;
        if (!moved) {
            CHECK(false, "stuck: no work-group can step (%d of %zu done)", done, g.size());
            break;
        }
    }
    steps_out += steps;
    return stale;
}

int main() {
    std::uint64_t steps = 0;
    int           runs  = 0;
    for (int seed = 0; seed < 200; ++seed) {
        g_rng = 0x2545f4914f6cdd1dull ^ (0x9e3779b97f4a7c15ull * static_cast<std::uint64_t>(seed + 1));
        const int nx = 1 + seed % 4, per = 1 + static_cast<int>(next_rand() % 24), jobs = 1 + static_cast<int>(next_rand() % 6);
        const int stale = simulate(nx, per, jobs, write_back_step, steps);
        CHECK(stale == 0, "seed %d (%d XCDs x %d groups x %d jobs): %d published stores not in memory", seed, nx, per, jobs, stale);
        ++runs;
    }
    std::printf("shared write-back: %d simulations, %llu steps, every published store in memory\n", runs, (unsigned long long)steps);

    // a planted mistake: covered as soon as completed >= s0 (a write-back that may have started before the stores arrived)
    int caught = 0;
    for (int seed = 0; seed < 200; ++seed) {
        g_rng = 0x51ed27087c3b2a1full ^ (0x9e3779b97f4a7c15ull * static_cast<std::uint64_t>(seed + 1));
        const Decide wrong = [](std::uint32_t s0, std::uint32_t d, std::uint32_t s) {
            if (d >= s0 && d > 0) return WriteBackStep::covered;
            return s == d ? WriteBackStep::issue : WriteBackStep::wait;
        };
        std::uint64_t unused = 0;
        caught += simulate(2, 12, 4, wrong, unused) > 0 ? 1 : 0;
    }
    CHECK(caught > 0, "the planted mistake (covered once completed >= s0) went unnoticed in 200 simulations");
    std::printf("planted mistake caught in %d of 200 simulations\n", caught);
    if (g_failures) {
        std::printf("%d failures\n", g_failures);
        return 1;
    }
    std::printf("OK\n");
    return 0;
}
