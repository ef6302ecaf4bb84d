// tests/host/batch_host.cpp — the batched ragged GEMVER model's host logic on the CPU (TEST INFRASTRUCTURE ONLY).
//
// Builds workloads/batch_systems.hpp's BatchedGemver through the real Model / System / Executor code, linked against the
// CPU implementation of the C boundary in mock_hip.cpp, with CPU emulations of the model's kernels (BatchPlan, the
// persistent RaggedGemver, and the phased form's RaggedPass1 / RaggedReduce / RaggedPass2) registered under their host
// stubs. It checks what the kernels receive (shapes, pointers, parameters, grids against the unit counts, the scan's
// offsets, the work-queue state zeroed at every start) and what the model computes (x bit for bit against the oracle's
// strip partials + column reduction, w against the FP64 oracle), over eager runs, a capture, replays and parameter
// changes, in both forms. Prints "OK" and exits 0 when every check holds.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "mock_hip.hpp"
#include "naga/cospan/cospan.hpp"
#include "batch_systems.hpp"
#include "oracle.h"

namespace cs = naga::cospan;
using GAPlan = cs::detail::GroupArgs<batch::BatchPlan>;
using GAGem  = cs::detail::GroupArgs<batch::RaggedGemver>;
using GAMap  = cs::detail::GroupArgs<batch::UnitMap>;
using GAP1   = cs::detail::GroupArgs<batch::RaggedPass1>;
using GARed  = cs::detail::GroupArgs<batch::RaggedReduce>;
using GAP2   = cs::detail::GroupArgs<batch::RaggedPass2>;

static int g_failures = 0;
#define CHECK(cond, ...)                                             \
    do {                                                             \
        if (!(cond)) {                                               \
            ++g_failures;                                            \
            std::printf("FAIL %s:%d: ", __FILE__, __LINE__);         \
            std::printf(__VA_ARGS__);                                \
            std::printf("\n");                                       \
        }                                                            \
    } while (0)

static std::uint64_t g_gemver_launches = 0;

// ---- BatchPlan: the counts of batch_components.hpp, per instance ----------------------------------------------------
static void emulate_plan(const void* bytes, const std::uint32_t grid[3], const std::uint32_t block[3]) {
    const auto&       a = static_cast<const GAPlan*>(bytes)->args.head;
    const std::size_t K = a.units[2];
    CHECK(a.units[0] == 1 && a.units[1] == 1, "plan: units %zu %zu", (size_t)a.units[0], (size_t)a.units[1]);
    CHECK(static_cast<std::size_t>(grid[0]) * block[0] >= K, "plan: grid %u x %u does not cover %zu instances", grid[0], block[0], K);
    CHECK(a.in.sizes.size() == K && a.in.rank.size() == K && a.out.counts.size() == batch::kColumns * K, "plan: sizes %zu rank %zu counts %zu for K %zu",
          a.in.sizes.size(), a.in.rank.size(), a.out.counts.size(), K);
    CHECK(mock::is_device(cs::raw::data(a.in.sizes)) && mock::is_device(cs::raw::data(a.out.counts)), "plan: a port outside device memory");
    const std::uint32_t* sz = cs::raw::data(a.in.sizes);
    const std::uint32_t* pr = cs::raw::data(a.in.params);
    const std::uint32_t* rk = cs::raw::data(a.in.rank);
    std::uint32_t*       c  = cs::raw::data(a.out.counts);
    const std::uint32_t  R = pr[0], RT = pr[1];
    // the work order: the instances by size, largest first, ties in index order
    std::vector<bool> seen(K, false);
    for (std::size_t k = 0; k < K; ++k) {
        CHECK(rk[k] < K && !seen[rk[k]], "plan: rank is not a permutation (instance %zu at %u)", k, rk[k]);
        if (rk[k] < K) seen[rk[k]] = true;
        for (std::size_t m = 0; m < K; ++m)
            if (m != k && (sz[m] > sz[k] || (sz[m] == sz[k] && m < k))) CHECK(rk[m] < rk[k], "plan: instance %zu (n %u) ranks after %zu (n %u)", m, sz[m], k, sz[k]);
    }
    for (std::size_t k = 0; k < K; ++k) {
        const std::uint32_t n = sz[k], st = (n + R - 1) / R, bl = (n + batch::kCols - 1) / batch::kCols;
        const std::uint32_t u[3] = {st * ((n + batch::kTile1 - 1) / batch::kTile1), bl, (n + RT - 1) / RT};
        c[k]         = n;
        c[K + k]     = n * batch::ld_of(n) / 4;
        c[2 * K + k] = st * n;
        for (int p = 0; p < 3; ++p) {
            c[(3 + p) * K + k] = u[p];
            if (rk[k] < K) c[(6 + p) * K + rk[k]] = u[p];
        }
    }
}

// ---- the units, as the device runs them (same fma order) -------------------------------------------------------------
struct Emu {
    const std::uint32_t *s = nullptr, *sz = nullptr;
    std::vector<std::uint32_t> order;  ///< order[r] = the instance at place r of the work order
    std::uint32_t        K = 0, R = 0, RT = 0;
    const float *        A = nullptr, *U1 = nullptr, *V1 = nullptr, *U2 = nullptr, *V2 = nullptr, *Y = nullptr, *Z = nullptr;
    const float *        Pin = nullptr, *Xin = nullptr;
    float *              P = nullptr, *X = nullptr, *W = nullptr;
    float                alpha = 0.0f, beta = 0.0f;

    std::uint32_t off(int c, std::uint32_t k) const { return s[c * K + k] - s[c * K]; }
    std::uint32_t total(int c) const { return s[(c + 1) * K] - s[c * K]; }
    std::uint32_t find(int c, std::uint32_t u) const {
        std::uint32_t lo = 0, hi = K - 1;
        while (lo < hi) {
            const std::uint32_t mid = (lo + hi + 1) / 2;
            if (off(c, mid) <= u)
                lo = mid;
            else
                hi = mid - 1;
        }
        return lo;
    }
    /// The instance of unit u of phase p (its place in the phase's queue) and the unit's index t within the instance.
    std::uint32_t locate(int p, std::uint32_t u, std::uint32_t& t) const {
        const std::uint32_t r = find(6 + p, u);
        t                     = u - off(6 + p, r);
        return order[r];
    }
    /// pass-1 unit u (global over the instances); returns its instance. A tile is R rows x kTile1 columns.
    std::uint32_t pass1(std::uint32_t u) const {
        std::uint32_t       t;
        const std::uint32_t k = locate(0, u, t), n = sz[k], ld = batch::ld_of(n);
        const std::uint32_t blocks = (n + batch::kTile1 - 1) / batch::kTile1, c = t / blocks, b = t % blocks;
        const std::size_t   vo = off(0, k);
        const float*        Ak = A + 4ull * off(1, k);
        float*              Pk = P + off(2, k);
        const std::uint32_t i0 = c * R, i1 = std::min(i0 + R, n);
        for (std::uint32_t j = b * batch::kTile1; j < std::min<std::uint32_t>(n, (b + 1) * batch::kTile1); ++j) {
            float acc = 0.0f;
            for (std::uint32_t i = i0; i < i1; ++i)
                acc = std::fma(std::fma(U2[vo + i], V2[vo + j], std::fma(U1[vo + i], V1[vo + j], Ak[std::size_t(i) * ld + j])), Y[vo + i], acc);
            Pk[std::size_t(c) * n + j] = acc;
        }
        return k;
    }
    std::uint32_t reduce(std::uint32_t u) const {
        std::uint32_t       b;
        const std::uint32_t k = locate(1, u, b), n = sz[k];
        const std::uint32_t strips = (n + R - 1) / R;
        const std::size_t   vo     = off(0, k);
        const float*        Pk     = Pin + off(2, k);
        for (std::uint32_t j = b * batch::kCols; j < std::min<std::uint32_t>(n, (b + 1) * batch::kCols); ++j) {
            float acc = 0.0f;
            for (std::uint32_t c = 0; c < strips; ++c) acc += Pk[std::size_t(c) * n + j];
            X[vo + j] = std::fma(beta, acc, Z[vo + j]);
        }
        return k;
    }
    std::uint32_t pass2(std::uint32_t u) const {
        std::uint32_t       g;
        const std::uint32_t k = locate(2, u, g), n = sz[k], ld = batch::ld_of(n);
        const std::size_t   vo = off(0, k);
        const float*        Ak = A + 4ull * off(1, k);
        for (std::uint32_t r = g * RT; r < std::min(n, g * RT + RT); ++r) {
            double acc = 0.0;
            for (std::uint32_t j = 0; j < n; ++j)
                acc += (double(Ak[std::size_t(r) * ld + j]) + double(U1[vo + r]) * V1[vo + j] + double(U2[vo + r]) * V2[vo + j]) * Xin[vo + j];
            W[vo + r] = static_cast<float>(alpha * acc);
        }
        return k;
    }
};

/// The fields every ragged kernel reads: params, the scan (checked against its expected length), sizes (the instances'
/// sizes as the host knows them, for the kernels that read them from the unit map).
static const std::uint32_t* g_sizes = nullptr;
static std::uint32_t        g_K     = 0;

/// The work order the host gives the model: the instances by size, largest first, ties in index order.
static std::vector<std::uint32_t> work_order() {
    std::vector<std::uint32_t> o(g_K);
    for (std::uint32_t k = 0; k < g_K; ++k) o[k] = k;
    std::stable_sort(o.begin(), o.end(), [](std::uint32_t a, std::uint32_t b) { return g_sizes[a] > g_sizes[b]; });
    return o;
}

template <class In>
static Emu emu_common(const In& in, const char* who) {
    Emu e;
    e.K  = g_K;
    e.s  = cs::raw::data(in.offs);
    e.sz = g_sizes;
    e.R  = cs::raw::data(in.params)[0];
    e.RT = cs::raw::data(in.params)[1];
    e.order = work_order();
    CHECK(in.offs.size() == batch::kColumns * e.K + 1, "%s: offs holds %zu words, not %u", who, in.offs.size(), batch::kColumns * e.K + 1);
    CHECK(std::size_t(e.s[batch::kColumns * e.K]) <= (std::size_t(1) << 31), "%s: the scan total %u is implausible", who, e.s[batch::kColumns * e.K]);
    return e;
}

/// The unit map's entry for unit u of phase p must name the unit the work order puts there.
static void check_map(const Emu& e, const std::uint32_t* map, std::size_t map_elems, int p, std::uint32_t u, const char* who) {
    CHECK((std::size_t(u) + 1) * batch::kMapWords <= map_elems, "%s: unit %u is outside its map (%zu words)", who, u, map_elems);
    if ((std::size_t(u) + 1) * batch::kMapWords > map_elems) return;
    const std::uint32_t* m = map + std::size_t(u) * batch::kMapWords;
    std::uint32_t        t;
    const std::uint32_t  k = e.locate(p, u, t);
    const bool ok = m[0] == e.sz[k] && m[1] == t && m[2] == e.off(0, k) && m[3] == e.off(1, k) && m[4] == e.off(2, k) && m[5] == k;
    CHECK(ok, "%s: map entry of unit %u reads n %u t %u vo %u a4 %u poff %u k %u; the scan says instance %u", who, u, m[0], m[1], m[2], m[3],
          m[4], m[5], k);
}

// ---- RaggedGemver: the persistent work queue, its protocol run by simulated work-groups -------------------------------
// The kernel's queue decisions are batch::queue's host-and-device functions; here G simulated work-groups step through
// them (first fetch; per unit: publish, wait, fetch the next ticket, run, hand over; leave) in an order a seeded generator
// picks, with at most `resident` work-groups started at a time (the rest start only when one leaves: no co-residency).
// A unit runs at once when its step is taken and publishes (counts itself in its instance's ready word) when it ends,
// so the checks are that every reduction of an instance comes after all of its pass-1 units and every pass-2 group
// after all of its reductions, that the unit map names each unit's instance, that a work-group never fetches more than
// T + 1 tickets, that the queue always drains (some work-group can step until all have left), and that it leaves the
// state zeroed.
static std::uint64_t g_rng = 0x9e3779b97f4a7c15ull;
static std::uint32_t next_rand() {
    g_rng ^= g_rng << 13;
    g_rng ^= g_rng >> 7;
    g_rng ^= g_rng << 17;
    return static_cast<std::uint32_t>(g_rng >> 16);
}
static std::uint64_t g_queue_runs = 0, g_queue_waits = 0;

static void emulate_gemver(const void* bytes, const std::uint32_t grid[3], const std::uint32_t block[3]) {
    namespace q = batch::queue;
    ++g_gemver_launches;
    const auto&         a = static_cast<const GAGem*>(bytes)->args.head;
    const std::uint32_t G = static_cast<std::uint32_t>(a.units[2]);
    CHECK(grid[0] == G && grid[1] == 1 && grid[2] == 1 && block[0] == batch::kWg, "gemver: grid %u %u %u block %u for G %u", grid[0],
          grid[1], grid[2], block[0], G);
    const std::uint32_t K = static_cast<std::uint32_t>(a.in.sizes.size());
    const std::size_t words = q::words(K);
    CHECK(a.out.state.size() == words, "gemver: state holds %zu words, not %zu", a.out.state.size(), words);
    std::uint32_t* st   = cs::raw::data(a.out.state);
    bool           zero = true;
    for (std::size_t w = 0; w < words; ++w) zero = zero && st[w] == 0;
    CHECK(zero, "gemver launch %llu: the work-queue state is not zero at the start (ticket %u, finished %u)",
          (unsigned long long)g_gemver_launches, st[q::kTicket], st[q::kFinished]);
    if (!zero) return;

    // the launch carries the device's fault word and the current poll bound; a bound of 0 makes every wait give up
    const auto&    geo = static_cast<const GAGem*>(bytes)->geo;
    std::uint32_t* fw  = nullptr;
    cs_hip_fault_word(&fw);
    CHECK(geo.fault == fw, "gemver: the launch does not carry the fault word");
    CHECK(geo.spin_limit == cs_hip_spin_limit(), "gemver: the launch carries spin limit %u, the process %u", geo.spin_limit,
          cs_hip_spin_limit());

    Emu e = emu_common(a.in, "gemver");
    e.A = cs::raw::data(a.in.A), e.U1 = cs::raw::data(a.in.u1), e.V1 = cs::raw::data(a.in.v1), e.U2 = cs::raw::data(a.in.u2);
    e.V2 = cs::raw::data(a.in.v2), e.Y = cs::raw::data(a.in.y), e.Z = cs::raw::data(a.in.z);
    e.P = cs::raw::data(a.out.P), e.X = cs::raw::data(a.out.x), e.W = cs::raw::data(a.out.w);
    e.Pin = e.P, e.Xin = e.X;
    e.alpha = cs::raw::data(a.in.alpha)[0], e.beta = cs::raw::data(a.in.beta)[0];
    const q::Totals n{e.total(3), e.total(4), e.total(5)};
    const std::uint32_t T = n.all();
    CHECK(a.in.A.size() >= 4ull * e.total(1), "gemver: A holds %zu floats, the scan asks %llu", a.in.A.size(), 4ull * e.total(1));
    CHECK(a.out.P.size() >= e.total(2), "gemver: P holds %zu floats, the scan asks %u", a.out.P.size(), e.total(2));
    CHECK(a.out.x.size() >= e.total(0) && a.in.u1.size() >= e.total(0), "gemver: vectors shorter than the scan's total");
    const std::uint32_t* maps[3]     = {cs::raw::data(a.in.map1), cs::raw::data(a.in.mapR), cs::raw::data(a.in.map2)};
    const std::size_t    map_el[3]   = {a.in.map1.size(), a.in.mapR.size(), a.in.map2.size()};
    const bool           give_up     = geo.spin_limit == 0;

    enum Where { kNotStarted, kTop, kRun, kLeft };
    struct WG {
        Where         at      = kNotStarted;
        std::uint32_t u = 0, next = 0, fetches = 0, seen = 0;
    };
    const std::uint32_t R = e.R;
    // the instance of unit i of phase p, from the unit map (checked against the scan)
    auto instance = [&](int p, std::uint32_t i) {
        check_map(e, maps[p], map_el[p], p, i, "gemver");
        std::uint32_t       t;
        const std::uint32_t k = e.locate(p, i, t);
        CHECK(maps[p][batch::kMapWords * std::size_t{i} + 5] == k, "gemver: unit %u of phase %d names instance %u, the scan %u", i, p,
              maps[p][batch::kMapWords * std::size_t{i} + 5], k);
        return k;
    };
    std::vector<WG>            wg(G);
    std::vector<std::uint32_t> p1(K, 0), red(K, 0);
    const std::uint32_t        resident = 1 + next_rand() % std::max<std::uint32_t>(G, 1);
    std::uint32_t              started = 0, left = 0;
    auto fetch = [&](WG& w) {
        ++w.fetches;
        CHECK(w.fetches <= T + 1, "gemver: a work-group fetched %u tickets, the bound is T + 1 = %u", w.fetches, T + 1);
        return st[q::kTicket]++;
    };
    // one step of work-group w; false when it cannot step now (not started while the device is full, or waiting)
    auto step = [&](WG& w) -> bool {
        switch (w.at) {
            case kNotStarted:
                if (started - left >= resident) return false;
                ++started;
                w.u  = fetch(w);
                w.at = kTop;
                return true;
            case kTop: {
                const int p = q::phase_of(w.u, n);
                if (p == 3) {
                    w.at = kLeft;
                    if (++st[q::kFinished] == G) std::memset(st, 0, words * 4);  // the last to leave zeroes the state
                    ++left;
                    return true;
                }
                const int c = p - 1;
                if (p > 0 && ((w.seen >> c) & 1u) == 0u) {  // the instance's units of the phase before, all published
                    const std::uint32_t i = w.u - n.base(p), k = instance(p, i);
                    if (st[q::published(c)] >= n.of(c)) {
                        w.seen |= 1u << c;
                    } else if (give_up) {
                        *geo.fault |= CS_HIP_FAULT_UNIT_WAIT;  // a poll bound of 0 gives up at once and runs on
                    } else if (st[q::ready(c, k, K)] < q::awaits(p, e.sz[k], R)) {
                        ++g_queue_waits;
                        return false;
                    }
                }
                w.next = fetch(w);
                w.at   = kRun;
                return true;
            }
            case kRun: {
                const int           p = q::phase_of(w.u, n);
                const std::uint32_t i = w.u - n.base(p), k = instance(p, i);
                if (p == 0) {
                    CHECK(e.pass1(i) == k, "gemver: pass-1 unit %u ran for another instance", i);
                    ++p1[k];
                } else if (p == 1) {
                    CHECK(give_up || p1[k] == e.s[3 * K + k + 1] - e.s[3 * K + k], "gemver: reduction of instance %u before its pass-1 units", k);
                    ++red[e.reduce(i)];
                } else {
                    CHECK(give_up || red[k] == e.s[4 * K + k + 1] - e.s[4 * K + k], "gemver: pass 2 of instance %u before its reductions", k);
                    e.pass2(i);
                }
                if (p < 2) {  // published
                    ++st[q::ready(p, k, K)];
                    ++st[q::published(p)];
                }
                ++g_queue_runs;
                w.u  = w.next;
                w.at = kTop;
                return true;
            }
            default: return false;
        }
    };
    std::vector<std::uint32_t> can;
    while (left < G) {
        can.clear();
        for (std::uint32_t g = 0; g < G; ++g)
            if (wg[g].at != kLeft) can.push_back(g);
        // try the work-groups in a random order until one steps; none can: the queue is stuck
        bool moved = false;
        for (std::size_t tries = can.size(); tries > 0 && !moved; --tries) {
            const std::size_t pick = next_rand() % can.size();
            moved                  = step(wg[can[pick]]);
            if (!moved) can.erase(can.begin() + static_cast<std::ptrdiff_t>(pick));
        }
        if (!moved) {
            CHECK(false, "gemver: the work queue is stuck (G %u, resident %u, %u left, ticket %u, T %u %u %u)", G, resident, left, st[q::kTicket], n.T1,
                  n.TR, n.T2);
            for (std::uint32_t g = 0; g < G && g < 400; ++g)
                if (wg[g].at != kNotStarted)
                    std::printf("  wg %u at %d u %u next %u fetches %u phase %d\n", g, int(wg[g].at), wg[g].u, wg[g].next, wg[g].fetches, q::phase_of(wg[g].u, n));
            std::memset(st, 0, words * 4);
            return;
        }
    }
    for (std::size_t w = 0; w < words; ++w) CHECK(st[w] == 0, "gemver: state word %zu is %u after the launch", w, st[w]);
}

// ---- the phased form: one launch per phase, work-group u runs unit u; the grid must cover the device plan's count ----
static std::uint64_t g_phase_launches[3] = {0, 0, 0};

template <class GA>
static void check_phase_grid(const GA& a, const std::uint32_t grid[3], const std::uint32_t block[3], std::uint32_t want, const char* who) {
    CHECK(grid[0] == a.units[2] && grid[1] == 1 && grid[2] == 1 && block[0] == batch::kWg, "%s: grid %u %u %u block %u for %zu groups", who,
          grid[0], grid[1], grid[2], block[0], (size_t)a.units[2]);
    CHECK(a.units[2] == want, "%s: %zu work-groups for %u units", who, (size_t)a.units[2], want);
}

static std::uint64_t g_map_launches = 0;

static void emulate_map(const void* bytes, const std::uint32_t grid[3], const std::uint32_t block[3]) {
    ++g_map_launches;
    const auto& a = static_cast<const GAMap*>(bytes)->args.head;
    Emu         e;
    e.K  = static_cast<std::uint32_t>(a.in.sizes.size());
    e.s  = cs::raw::data(a.in.offs);
    e.sz = cs::raw::data(a.in.sizes);
    CHECK(e.K == g_K && a.in.offs.size() == batch::kColumns * e.K + 1 && a.in.rank.size() == e.K, "map: %u instances, offs %zu, rank %zu", e.K,
          a.in.offs.size(), a.in.rank.size());
    const std::uint32_t* rk = cs::raw::data(a.in.rank);
    CHECK(grid[0] == e.K && block[0] == batch::kWg, "map: grid %u block %u for %u instances", grid[0], block[0], e.K);
    std::uint32_t* maps[3] = {cs::raw::data(a.out.map1), cs::raw::data(a.out.mapR), cs::raw::data(a.out.map2)};
    const std::size_t caps[3] = {a.out.map1.size() / batch::kMapWords, a.out.mapR.size() / batch::kMapWords, a.out.map2.size() / batch::kMapWords};
    for (std::uint32_t k = 0; k < e.K; ++k)
        for (int p = 0; p < 3; ++p) {
            const std::uint32_t o = e.off(6 + p, rk[k]), cnt = e.s[(3 + p) * e.K + k + 1] - e.s[(3 + p) * e.K + k];
            CHECK(std::size_t(o) + cnt <= caps[p], "map: phase %d of instance %u ends at %u, the map holds %zu", p, k, o + cnt, caps[p]);
            if (std::size_t(o) + cnt > caps[p]) continue;
            for (std::uint32_t i = 0; i < cnt; ++i) {
                std::uint32_t* m = maps[p] + std::size_t(o + i) * batch::kMapWords;
                const std::uint32_t v[8] = {e.sz[k], i, e.off(0, k), e.off(1, k), e.off(2, k), k, 0, 0};
                std::memcpy(m, v, sizeof v);
            }
        }
}

static void emulate_pass1(const void* bytes, const std::uint32_t grid[3], const std::uint32_t block[3]) {
    ++g_phase_launches[0];
    const auto& a = static_cast<const GAP1*>(bytes)->args.head;
    Emu         e = emu_common(a.in, "pass 1");
    e.A = cs::raw::data(a.in.A), e.U1 = cs::raw::data(a.in.u1), e.V1 = cs::raw::data(a.in.v1), e.U2 = cs::raw::data(a.in.u2);
    e.V2 = cs::raw::data(a.in.v2), e.Y = cs::raw::data(a.in.y), e.P = cs::raw::data(a.out.P);
    check_phase_grid(a, grid, block, e.total(3), "pass 1");
    CHECK(a.out.P.size() >= e.total(2), "pass 1: P holds %zu floats, the scan asks %u", a.out.P.size(), e.total(2));
    CHECK(a.in.A.size() >= 4ull * e.total(1), "pass 1: A holds %zu floats, the scan asks %llu", a.in.A.size(), 4ull * e.total(1));
    for (std::uint32_t u = 0; u < std::min<std::uint32_t>(static_cast<std::uint32_t>(a.units[2]), e.total(3)); ++u) {
        check_map(e, cs::raw::data(a.in.map), a.in.map.size(), 0, u, "pass 1");
        e.pass1(u);
    }
}

static void emulate_reduce(const void* bytes, const std::uint32_t grid[3], const std::uint32_t block[3]) {
    ++g_phase_launches[1];
    const auto& a = static_cast<const GARed*>(bytes)->args.head;
    Emu         e = emu_common(a.in, "reduce");
    e.Pin = cs::raw::data(a.in.P), e.Z = cs::raw::data(a.in.z), e.X = cs::raw::data(a.out.x), e.beta = cs::raw::data(a.in.beta)[0];
    check_phase_grid(a, grid, block, e.total(4), "reduce");
    CHECK(a.out.x.size() >= e.total(0) && a.in.z.size() >= e.total(0), "reduce: vectors shorter than the scan's total");
    for (std::uint32_t u = 0; u < std::min<std::uint32_t>(static_cast<std::uint32_t>(a.units[2]), e.total(4)); ++u) {
        check_map(e, cs::raw::data(a.in.map), a.in.map.size(), 1, u, "reduce");
        e.reduce(u);
    }
}

static void emulate_pass2(const void* bytes, const std::uint32_t grid[3], const std::uint32_t block[3]) {
    ++g_phase_launches[2];
    const auto& a = static_cast<const GAP2*>(bytes)->args.head;
    Emu         e = emu_common(a.in, "pass 2");
    e.A = cs::raw::data(a.in.A), e.U1 = cs::raw::data(a.in.u1), e.V1 = cs::raw::data(a.in.v1), e.U2 = cs::raw::data(a.in.u2);
    e.V2 = cs::raw::data(a.in.v2), e.Xin = cs::raw::data(a.in.x), e.W = cs::raw::data(a.out.w), e.alpha = cs::raw::data(a.in.alpha)[0];
    check_phase_grid(a, grid, block, e.total(5), "pass 2");
    CHECK(a.out.w.size() >= e.total(0) && a.in.x.size() >= e.total(0), "pass 2: vectors shorter than the scan's total");
    for (std::uint32_t u = 0; u < std::min<std::uint32_t>(static_cast<std::uint32_t>(a.units[2]), e.total(5)); ++u) {
        check_map(e, cs::raw::data(a.in.map), a.in.map.size(), 2, u, "pass 2");
        e.pass2(u);
    }
}

// ---- inputs: instance k's A from one flat hash stream, written on the host into the mock's device memory ------------
static batch::AInit host_a_init(std::uint32_t seed) {
    return [seed](const batch::Layout& L) {
        auto layout = std::make_shared<batch::Layout>(L);
        return cs::init::custom<cs::DType::f32, 1>(
            [seed, layout](cs_hip_stream s, cs::Initializer<cs::DType::f32, 1> a) {
                cs::detail::hip_check(cs_hip_stream_synchronize(s), "host A init");
                float*        A   = cs::raw::data(a);
                std::uint64_t off = 0, base = 0;
                for (std::uint32_t n : layout->sizes) {
                    const std::uint32_t ld = batch::ld_of(n);
                    for (std::uint32_t i = 0; i < n; ++i)
                        for (std::uint32_t j = 0; j < ld; ++j)
                            A[4 * off + std::size_t(i) * ld + j] = j < n ? or_hash_unit(seed, base + std::uint64_t(i) * n + j) : 0.0f;
                    off += std::uint64_t(n) * ld / 4;
                    base += std::uint64_t(n) * n;
                }
            },
            cs::Shape{4 * layout->a4_total});
    };
}

// ---- the check of one run's outputs ------------------------------------------------------------------------------------
static void check_outputs(cs::Model& m, const std::vector<std::uint32_t>& sizes, std::uint32_t R, std::uint32_t seed, float alpha,
                          float beta, const char* what) {
    m.wait();
    const float* x = static_cast<const float*>(m.device_data("x"));
    const float* w = static_cast<const float*>(m.device_data("w"));
    std::uint64_t vo = 0, ab = 0;
    double        worst = 0.0;
    for (std::size_t k = 0; k < sizes.size(); ++k) {
        const std::uint32_t n = sizes[k], strips = (n + R - 1) / R;
        std::vector<float> A(std::size_t(n) * n), vec[6];
        or_fill_hash_f32_base(seed * 16 + 1, A.data(), n, n, n, ab);
        for (int i = 0; i < 6; ++i) {
            vec[i].resize(n);
            or_fill_hash_f32_base(seed * 16 + 2 + i, vec[i].data(), 1, n, n, vo);
        }
        std::vector<float> P(std::size_t(strips) * n), xf(n);
        or_strip_partials_f32(n, n, R, A.data(), vec[0].data(), vec[1].data(), vec[2].data(), vec[3].data(), vec[4].data(), P.data());
        or_col_reduce_f32(n, strips, P.data(), vec[5].data(), beta, xf.data());
        CHECK(std::memcmp(xf.data(), x + vo, n * 4) == 0, "%s: x of instance %zu (n %u) differs from the oracle's strip partials", what, k, n);
        std::vector<double> xr(n), wr(n);
        or_gemver_f64(n, 0, A.data(), n, nullptr, vec[0].data(), vec[1].data(), vec[2].data(), vec[3].data(), vec[4].data(), vec[5].data(), alpha,
                      beta, xr.data(), wr.data());
        const double ew = or_relative_l2_f32(w + vo, wr.data(), n);
        worst           = std::max(worst, ew);
        CHECK(ew < 1e-5, "%s: w of instance %zu (n %u): rel-L2 %.3g", what, k, n, ew);
        vo += n;
        ab += std::uint64_t(n) * n;
    }
    std::printf("%s: %zu instances, worst w rel-L2 %.2e\n", what, sizes.size(), worst);
}

/// One form of the model through eager runs, a capture, replays and a parameter change (replays as graph launches, or with
/// ReplayMode::direct as the captured launches re-issued on stream 0).
static void run_form(bool phased, cs::ReplayMode mode = cs::ReplayMode::graph) {
    const bool          direct = mode == cs::ReplayMode::direct;
    const char*         form   = phased ? (direct ? "phased, direct replays" : "phased") : "persistent";
    const std::uint32_t count = 24, seed = 1;
    std::vector<std::uint32_t> sizes(count);
    or_ragged_sizes(seed, count, 16, 700, sizes.data());
    sizes[5] = 4096 + 1000;  // an instance wider than one pass-1 tile (two column blocks, the second a partial one)
    sizes[9] = 4096;         // exactly one tile wide
    g_sizes = sizes.data();
    g_K     = count;
    const float alpha = 1.5f, beta = 1.2f;
    auto        layout = std::make_shared<batch::Layout>(sizes, batch::BatchedGemver::kMinR);
    auto        knobs  = std::make_shared<batch::Knobs>();
    std::vector<std::uint32_t> params{32, 4, 3};
    knobs->groups = 3;
    batch::count_units(sizes.data(), count, 32, 4, knobs->units);
    cs::Model m = cs::create_model<batch::BatchedGemver>(layout, knobs, params, alpha, beta, seed, host_a_init(seed * 16 + 1), phased);
    m.setup();
    m.set_replay_mode(mode);
    char what[160];

    m.run();
    std::snprintf(what, sizeof what, "%s: eager run (R 32, RT 4, G 3)", form);
    check_outputs(m, sizes, 32, seed, alpha, beta, what);
    for (int i = 0; i < 3; ++i) m.run();  // the second identical run is captured, then replays
    std::snprintf(what, sizeof what, "%s: capture + replays", form);
    check_outputs(m, sizes, 32, seed, alpha, beta, what);
    const cs::ModelStats st1 = m.stats();
    CHECK(st1.captures >= 1 && st1.replays >= 2, "%s: captures %llu replays %llu", form, (unsigned long long)st1.captures,
          (unsigned long long)st1.replays);
    // the batched model runs on one stream: direct replays are plain launches, graph replays one graph launch each
    CHECK(direct ? st1.relaunches == st1.replays - 1 : st1.relaunches == 0, "%s: relaunches %llu of %llu replays", form,
          (unsigned long long)st1.relaunches, (unsigned long long)st1.replays);

    params        = {64, 8, 17};  // new parameters: new counts, a new grid
    knobs->groups = 17;
    batch::count_units(sizes.data(), count, 64, 8, knobs->units);
    m.update_input("params", cs::init::from_host(params));
    for (int i = 0; i < 3; ++i) m.run();
    std::snprintf(what, sizeof what, "%s: after a parameter change (R 64, RT 8, G 17)", form);
    check_outputs(m, sizes, 64, seed, alpha, beta, what);
    // wave-level pass 2 (RT 16, 32) and long strips; the persistent form over grids from one work-group to more than
    // there are units, each run a new interleaving and residency cap
    const std::uint32_t more[][3] = {{256, 32, 5}, {32, 16, 1}, {128, 32, 2}, {64, 16, 400}, {32, 32, 7}};
    for (const auto& c : more) {
        params        = {c[0], c[1], c[2]};
        knobs->groups = static_cast<int>(c[2]);
        batch::count_units(sizes.data(), count, c[0], c[1], knobs->units);
        m.update_input("params", cs::init::from_host(params));
        for (int i = 0; i < 4; ++i) m.run();
        std::snprintf(what, sizeof what, "%s: R %u, RT %u, G %u", form, c[0], c[1], c[2]);
        check_outputs(m, sizes, c[0], seed, alpha, beta, what);
    }
}

/// Bounded waits on the host side: a launch carries the fault word and the current spin limit; with a limit of 0 the
/// work-queue kernel's waits give up, Model::wait throws std::runtime_error naming them and clears the word, and a
/// captured graph is not replayed with a stale limit (the limit is part of the launch's argument bytes).
static void run_fault_reporting() {
    const std::uint32_t        count = 6, seed = 2;
    std::vector<std::uint32_t> sizes(count);
    or_ragged_sizes(seed, count, 16, 300, sizes.data());
    g_sizes = sizes.data();
    g_K     = count;
    auto layout = std::make_shared<batch::Layout>(sizes, batch::BatchedGemver::kMinR);
    auto knobs  = std::make_shared<batch::Knobs>();
    std::vector<std::uint32_t> params{32, 4, 3};
    knobs->groups = 3;
    batch::count_units(sizes.data(), count, 32, 4, knobs->units);
    cs::Model m = cs::create_model<batch::BatchedGemver>(layout, knobs, params, 1.5f, 1.2f, seed, host_a_init(seed * 16 + 1), false);
    m.setup();
    for (int i = 0; i < 3; ++i) m.run();  // eager, captured, replayed
    m.wait();
    cs_hip_set_spin_limit(0);
    m.run();
    bool        threw = false;
    std::string msg;
    try {
        m.wait();
    } catch (const std::runtime_error& e) {
        threw = true;
        msg   = e.what();
    }
    CHECK(threw && msg.find("gave up") != std::string::npos && msg.find("work-queue") != std::string::npos,
          "faults: a wait that gave up was not reported (threw %d: %s)", threw ? 1 : 0, msg.c_str());
    bool again = false;
    try {
        m.wait();
    } catch (const std::exception&) {
        again = true;
    }
    CHECK(!again, "faults: the fault word was not cleared by the wait that reported it");
    cs_hip_set_spin_limit(CS_HIP_DEFAULT_SPIN_LIMIT);
    for (int i = 0; i < 3; ++i) m.run();
    check_outputs(m, sizes, 32, seed, 1.5f, 1.2f, "faults: runs after the limit is restored");
    std::printf("faults: a work-queue wait that gave up is reported by Model::wait\n");
}

int main() {
    mock::register_kernel(reinterpret_cast<const void*>(&cs::detail::index_kernel<batch::BatchPlan>), sizeof(GAPlan), emulate_plan);
    mock::register_kernel(reinterpret_cast<const void*>(&cs::detail::group_kernel<batch::kWg, 1, batch::RaggedGemver>), sizeof(GAGem),
                          emulate_gemver);
    mock::register_kernel(reinterpret_cast<const void*>(&cs::detail::group_kernel<batch::kWg, 1, batch::UnitMap>), sizeof(GAMap), emulate_map);
    mock::register_kernel(reinterpret_cast<const void*>(&cs::detail::group_kernel<batch::kWg, 1, batch::RaggedPass1>), sizeof(GAP1), emulate_pass1);
    mock::register_kernel(reinterpret_cast<const void*>(&cs::detail::group_kernel<batch::kWg, 1, batch::RaggedReduce>), sizeof(GARed),
                          emulate_reduce);
    mock::register_kernel(reinterpret_cast<const void*>(&cs::detail::group_kernel<batch::kWg, 1, batch::RaggedPass2>), sizeof(GAP2), emulate_pass2);

    run_form(false);
    run_form(true);
    run_form(true, cs::ReplayMode::direct);
    run_fault_reporting();

    const std::string f = mock::failure();
    CHECK(f.empty(), "mock: %s", f.c_str());
    CHECK(g_gemver_launches >= 4, "the persistent kernel ran %llu times", (unsigned long long)g_gemver_launches);
    CHECK(g_queue_waits > 0, "no simulated work-group ever had to wait (the interleavings never tested a wait)");
    std::printf("persistent queue: %llu units run, %llu blocked steps\n", (unsigned long long)g_queue_runs, (unsigned long long)g_queue_waits);
    for (int p = 0; p < 3; ++p) CHECK(g_phase_launches[p] >= 4, "phase %d ran %llu times", p, (unsigned long long)g_phase_launches[p]);
    CHECK(g_map_launches >= 4, "the unit map ran %llu times", (unsigned long long)g_map_launches);
    if (g_failures) {
        std::printf("%d failure(s)\n", g_failures);
        return 1;
    }
    std::printf("OK\n");
    return 0;
}
