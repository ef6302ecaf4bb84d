// tests/host/batch_host.cpp — the batched ragged GEMVER model's host logic on the CPU (TEST INFRASTRUCTURE ONLY).
//
// Builds workloads/batch_systems.hpp's BatchedGemver through the real Model / System / Executor code, linked against the
// CPU implementation of the C boundary in mock_hip.cpp, with CPU emulations of the model's two kernels (BatchPlan and the
// persistent RaggedGemver) registered under their host stubs. It checks what the kernels receive (shapes, pointers,
// parameters, the scan's offsets, the work-queue state zeroed at every start) and what the model computes (x bit for bit
// against the oracle's strip partials + column reduction, w against the FP64 oracle), over eager runs, a capture, replays
// and parameter changes. Prints "OK" and exits 0 when every check holds.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <memory>
#include <vector>

#include "mock_hip.hpp"
#include "naga/cospan/cospan.hpp"
#include "batch_systems.hpp"
#include "oracle.h"

namespace cs = naga::cospan;
using GAPlan = cs::detail::GroupArgs<batch::BatchPlan>;
using GAGem  = cs::detail::GroupArgs<batch::RaggedGemver>;

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
    CHECK(a.in.sizes.size() == K && a.out.counts.size() == 6 * K, "plan: sizes %zu counts %zu for K %zu", a.in.sizes.size(),
          a.out.counts.size(), K);
    CHECK(mock::is_device(cs::raw::data(a.in.sizes)) && mock::is_device(cs::raw::data(a.out.counts)), "plan: a port outside device memory");
    const std::uint32_t* sz = cs::raw::data(a.in.sizes);
    const std::uint32_t* pr = cs::raw::data(a.in.params);
    std::uint32_t*       c  = cs::raw::data(a.out.counts);
    const std::uint32_t  R = pr[0], RT = pr[1];
    for (std::size_t k = 0; k < K; ++k) {
        const std::uint32_t n = sz[k], st = (n + R - 1) / R, bl = (n + batch::kCols - 1) / batch::kCols;
        c[k]         = n;
        c[K + k]     = n * batch::ld_of(n) / 4;
        c[2 * K + k] = st * n;
        c[3 * K + k] = st * bl;
        c[4 * K + k] = bl;
        c[5 * K + k] = (n + RT - 1) / RT;
    }
}

// ---- RaggedGemver: the persistent work queue, its units in ticket order ----------------------------------------------
static void emulate_gemver(const void* bytes, const std::uint32_t grid[3], const std::uint32_t block[3]) {
    ++g_gemver_launches;
    const auto&         a = static_cast<const GAGem*>(bytes)->args.head;
    const std::uint32_t G = static_cast<std::uint32_t>(a.units[2]);
    CHECK(grid[0] == G && grid[1] == 1 && grid[2] == 1 && block[0] == batch::kWg, "gemver: grid %u %u %u block %u for G %u", grid[0],
          grid[1], grid[2], block[0], G);
    const std::uint32_t K = static_cast<std::uint32_t>(a.in.sizes.size());
    CHECK(a.out.state.size() == 2 + 2 * K, "gemver: state holds %zu words, not %u", a.out.state.size(), 2 + 2 * K);
    CHECK(a.in.offs.size() == 6 * K + 1, "gemver: offs holds %zu words, not %u", a.in.offs.size(), 6 * K + 1);
    std::uint32_t* st = cs::raw::data(a.out.state);
    bool           zero = true;
    for (std::uint32_t w = 0; w < 2 + 2 * K; ++w) zero = zero && st[w] == 0;
    CHECK(zero, "gemver launch %llu: the work-queue state is not zero at the start (ticket %u, finished %u)",
          (unsigned long long)g_gemver_launches, st[0], st[1]);
    if (!zero) return;

    const std::uint32_t* s  = cs::raw::data(a.in.offs);
    const std::uint32_t* sz = cs::raw::data(a.in.sizes);
    const std::uint32_t* pr = cs::raw::data(a.in.params);
    const std::uint32_t  R = pr[0], RT = pr[1];
    auto off   = [&](int c, std::uint32_t k) { return s[c * K + k] - s[c * K]; };
    auto total = [&](int c) { return s[(c + 1) * K] - s[c * K]; };
    auto find  = [&](int c, std::uint32_t u) {
        std::uint32_t lo = 0, hi = K - 1;
        while (lo < hi) {
            const std::uint32_t mid = (lo + hi + 1) / 2;
            if (off(c, mid) <= u)
                lo = mid;
            else
                hi = mid - 1;
        }
        return lo;
    };
    const std::uint32_t T1 = total(3), TR = total(4), T2 = total(5);
    CHECK(std::size_t(s[6 * K]) <= (std::size_t(1) << 31), "gemver: the scan total %u is implausible", s[6 * K]);
    const float *A = cs::raw::data(a.in.A), *U1 = cs::raw::data(a.in.u1), *V1 = cs::raw::data(a.in.v1), *U2 = cs::raw::data(a.in.u2);
    const float *V2 = cs::raw::data(a.in.v2), *Y = cs::raw::data(a.in.y), *Z = cs::raw::data(a.in.z);
    float *P = cs::raw::data(a.out.P), *X = cs::raw::data(a.out.x), *W = cs::raw::data(a.out.w);
    const float alpha = cs::raw::data(a.in.alpha)[0], beta = cs::raw::data(a.in.beta)[0];
    CHECK(a.in.A.size() >= 4ull * total(1), "gemver: A holds %zu floats, the scan asks %llu", a.in.A.size(), 4ull * total(1));
    CHECK(a.out.P.size() >= total(2), "gemver: P holds %zu floats, the scan asks %u", a.out.P.size(), total(2));
    CHECK(a.out.x.size() >= total(0) && a.in.u1.size() >= total(0), "gemver: vectors shorter than the scan's total");
    std::vector<std::uint32_t> p1(K, 0), red(K, 0);
    for (std::uint32_t u = 0; u < T1 + TR + T2; ++u) {
        if (u < T1) {
            const std::uint32_t k = find(3, u), t = u - off(3, k), n = sz[k], ld = batch::ld_of(n);
            const std::uint32_t blocks = (n + batch::kCols - 1) / batch::kCols, c = t / blocks, b = t % blocks;
            const std::size_t   vo = off(0, k);
            const float*        Ak = A + 4ull * off(1, k);
            float*              Pk = P + off(2, k);
            const std::uint32_t i0 = c * R, i1 = std::min(i0 + R, n);
            for (std::uint32_t j = b * batch::kCols; j < std::min<std::uint32_t>(n, (b + 1) * batch::kCols); ++j) {
                float acc = 0.0f;
                for (std::uint32_t i = i0; i < i1; ++i)
                    acc = std::fma(std::fma(U2[vo + i], V2[vo + j], std::fma(U1[vo + i], V1[vo + j], Ak[std::size_t(i) * ld + j])), Y[vo + i], acc);
                Pk[std::size_t(c) * n + j] = acc;
            }
            ++p1[k];
        } else if (u < T1 + TR) {
            const std::uint32_t k = find(4, u - T1), b = u - T1 - off(4, k), n = sz[k];
            CHECK(p1[k] == s[3 * K + k + 1] - s[3 * K + k], "gemver: reduction of instance %u before its pass-1 tiles", k);
            const std::uint32_t strips = (n + R - 1) / R;
            const std::size_t   vo     = off(0, k);
            const float*        Pk     = P + off(2, k);
            for (std::uint32_t j = b * batch::kCols; j < std::min<std::uint32_t>(n, (b + 1) * batch::kCols); ++j) {
                float acc = 0.0f;
                for (std::uint32_t c = 0; c < strips; ++c) acc += Pk[std::size_t(c) * n + j];
                X[vo + j] = std::fma(beta, acc, Z[vo + j]);
            }
            ++red[k];
        } else {
            const std::uint32_t k = find(5, u - T1 - TR), g = u - T1 - TR - off(5, k), n = sz[k], ld = batch::ld_of(n);
            CHECK(red[k] == s[4 * K + k + 1] - s[4 * K + k], "gemver: pass 2 of instance %u before its reductions", k);
            const std::size_t vo = off(0, k);
            const float*      Ak = A + 4ull * off(1, k);
            for (std::uint32_t r = g * RT; r < std::min(n, g * RT + RT); ++r) {
                double acc = 0.0;
                for (std::uint32_t j = 0; j < n; ++j)
                    acc += (double(Ak[std::size_t(r) * ld + j]) + double(U1[vo + r]) * V1[vo + j] + double(U2[vo + r]) * V2[vo + j]) * X[vo + j];
                W[vo + r] = static_cast<float>(alpha * acc);
            }
        }
    }
    // the last work-group to leave zeroes the state (it was zero, and the queue leaves it so)
    std::memset(st, 0, (2 + 2 * K) * 4);
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

int main() {
    mock::register_kernel(reinterpret_cast<const void*>(&cs::detail::index_kernel<batch::BatchPlan>), sizeof(GAPlan), emulate_plan);
    mock::register_kernel(reinterpret_cast<const void*>(&cs::detail::group_kernel<batch::kWg, batch::RaggedGemver>), sizeof(GAGem),
                          emulate_gemver);

    const std::uint32_t count = 24, seed = 1;
    std::vector<std::uint32_t> sizes(count);
    or_ragged_sizes(seed, count, 16, 700, sizes.data());
    const float alpha = 1.5f, beta = 1.2f;
    auto        layout = std::make_shared<batch::Layout>(sizes, batch::BatchedGemver::kMinR);
    auto        knobs  = std::make_shared<batch::Knobs>();
    std::vector<std::uint32_t> params{32, 4, 3};
    knobs->groups = 3;
    cs::Model m   = cs::create_model<batch::BatchedGemver>(layout, knobs, params, alpha, beta, seed, host_a_init(seed * 16 + 1));
    m.setup();

    m.run();
    check_outputs(m, sizes, 32, seed, alpha, beta, "eager run (R 32, RT 4, G 3)");
    for (int i = 0; i < 3; ++i) m.run();  // the second identical run is captured, then replays
    check_outputs(m, sizes, 32, seed, alpha, beta, "capture + replays");
    const cs::ModelStats st1 = m.stats();
    CHECK(st1.captures >= 1 && st1.replays >= 2, "captures %llu replays %llu", (unsigned long long)st1.captures,
          (unsigned long long)st1.replays);

    params        = {64, 8, 17};  // new parameters: new counts, a new grid
    knobs->groups = 17;
    m.update_input("params", cs::init::from_host(params));
    for (int i = 0; i < 3; ++i) m.run();
    check_outputs(m, sizes, 64, seed, alpha, beta, "after a parameter change (R 64, RT 8, G 17)");

    const std::string f = mock::failure();
    CHECK(f.empty(), "mock: %s", f.c_str());
    CHECK(g_gemver_launches >= 4, "the persistent kernel ran %llu times", (unsigned long long)g_gemver_launches);
    if (g_failures) {
        std::printf("%d failure(s)\n", g_failures);
        return 1;
    }
    std::printf("OK\n");
    return 0;
}
