// cospanmdao_amd/csrc/chain_lib.hip — the GEMVER -> BiCGK -> GEMVER chain behind the C boundary of include/cospan_workloads.h.
//
// Each handle owns a cs::Model whose root is workloads/chain_systems.hpp's Chain in one of the reference's six styles
// (examples/gemver-chain/systems.hpp). The entry points are the reference driver's actions (examples/gemver-chain/
// main.cpp:154-188): build, run, wait, read an output, time.

#include <hip/hip_runtime.h>

#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "cospan_hip.h"
#include "cospan_workloads.h"
#include "internal_workloads.hpp"
#include "naga/cospan/cospan.hpp"
#include "../workloads/chain_systems.hpp"

namespace cs = naga::cospan;

struct cs_chain_s {
    int                        style = 0;
    std::size_t                n = 0, R = 64;
    std::uint32_t              seed = 0;
    std::unique_ptr<cs::Model> model;
};

namespace {

cs::Model& model(cs_chain_s& g) {
    if (!g.model) {
        g.model = std::make_unique<cs::Model>(
            cs::create_model<chain::Chain>(static_cast<chain::Style>(g.style), static_cast<int>(g.n), static_cast<int>(g.R), g.seed));
        g.model->setup();
    }
    return *g.model;
}

}  // namespace

extern "C" {

int cs_chain_create(cs_chain* out, int style, size_t n, size_t R, uint32_t base_seed) {
    return wl::guard([&] {
        if (!out) throw std::invalid_argument("cs_chain_create: null handle");
        if (style < CS_CHAIN_UNFUSED || style > CS_CHAIN_CRAFTED) throw std::invalid_argument("gemver-chain: unknown style");
        // the reference's problem constraints (examples/gemver-chain/common/chain_data.hpp:76-82); the crafted style's strips
        // are kRt rows whatever R is
        if (n == 0 || n % 256 != 0)
            throw std::invalid_argument("gemver-chain: n must be a positive multiple of 256 (got " + std::to_string(n) + ")");
        if (R == 0 || n % R != 0)
            throw std::invalid_argument("gemver-chain: n must be a multiple of R (n = " + std::to_string(n) + ", R = " + std::to_string(R) + ")");
        auto g   = std::make_unique<cs_chain_s>();
        g->style = style;
        g->n     = n;
        g->R     = style == CS_CHAIN_CRAFTED ? gemver::kRt : R;
        g->seed  = base_seed;
        *out     = g.release();
    });
}

int cs_chain_run(cs_chain g) {
    return wl::guard([&] {
        if (!g) throw std::invalid_argument("cs_chain_run: null handle");
        model(*g).run();
    });
}

int cs_chain_wait(cs_chain g) {
    return wl::guard([&] {
        if (!g) throw std::invalid_argument("cs_chain_wait: null handle");
        model(*g).wait();
    });
}

int cs_chain_read(cs_chain g, const char* name, void* host, size_t bytes) {
    return wl::guard([&] {
        if (!g || !name || !host) throw std::invalid_argument("cs_chain_read: null argument");
        cs::Model&        m = model(*g);
        const std::string s = name;
        m.wait();
        const std::vector<std::size_t> shape = m.shape_of(s);
        std::size_t                    elems = 1;
        for (std::size_t e : shape) elems *= e;
        if (bytes != elems * 4)
            throw std::invalid_argument("gemver-chain: '" + s + "' holds " + std::to_string(elems * 4) + " bytes, not " + std::to_string(bytes));
        void* src = m.device_data(s);
        if (!src) throw std::logic_error("gemver-chain: '" + s + "' holds no data yet (run the model first)");
        cs::detail::hip_check(cs_hip_memcpy_async(host, src, bytes, m.stream()), "chain read");
        cs::detail::hip_check(cs_hip_stream_synchronize(m.stream()), "chain read");
    });
}

int cs_chain_time(cs_chain g, int iters, float* ms_per_run) {
    return wl::guard([&] {
        if (!g || !ms_per_run || iters <= 0) throw std::invalid_argument("cs_chain_time: bad argument");
        cs::Model& m = model(*g);
        m.wait();
        wl::EventPair ev;
        ev.record_start(m.stream());
        for (int i = 0; i < iters; ++i) m.run();
        m.wait();
        ev.record_stop(m.stream());
        *ms_per_run = ev.elapsed_ms() / static_cast<float>(iters);
    });
}

int cs_chain_time_launches(cs_chain g, int iters, float* ms, int cap, int* count) {
    return wl::guard([&] {
        if (!g || !ms || !count || iters <= 0) throw std::invalid_argument("cs_chain_time_launches: bad argument");
        cs::Model&          m = model(*g);
        std::vector<double> sum;
        for (int i = 0; i < iters; ++i) {
            const auto t = m.run_timed();
            if (sum.empty()) sum.assign(t.size(), 0.0);
            for (std::size_t k = 0; k < t.size() && k < sum.size(); ++k) sum[k] += t[k].second;
        }
        *count = static_cast<int>(sum.size());
        for (int k = 0; k < cap && k < *count; ++k) ms[k] = static_cast<float>(sum[k] / iters);
    });
}

int cs_chain_stats(cs_chain g, uint64_t* stats4) {
    return wl::guard([&] {
        if (!g || !stats4) throw std::invalid_argument("cs_chain_stats: null argument");
        const cs::ModelStats s = model(*g).stats();
        stats4[0]              = s.runs;
        stats4[1]              = s.launches;
        stats4[2]              = s.captures;
        stats4[3]              = s.replays;
    });
}

int cs_chain_destroy(cs_chain g) {
    return wl::guard([&] {
        if (!g) return;
        delete g;
    });
}

}  // extern "C"
