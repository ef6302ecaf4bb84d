// cospanmdao_amd/csrc/gemver_lib.hip — the GEMVER model behind the C boundary of include/cospan_workloads.h.
//
// Each handle owns a cs::Model whose root is the Gemver System of workloads/gemver_systems.hpp (the reference's
// examples/gemver-chain/systems.hpp:71-139 in its unfused / fused / naive styles, plus this build's two-pass styles). The
// C entry points are the reference driver's actions (examples/gemver-chain/main.cpp:154-234): build, run, wait, read an
// output, time.

#include <hip/hip_runtime.h>

#include <cstring>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "cospan_hip.h"
#include "cospan_workloads.h"
#include "internal_workloads.hpp"
#include "naga/cospan/cospan.hpp"
#include "../workloads/gemver_systems.hpp"

namespace cs = naga::cospan;
using cs::DType;

struct cs_gemver_s {
    int                        style = 0, store = 0;
    std::size_t                n = 0, R = 64;
    float                      alpha = 1.0f, beta = 1.0f;
    std::uint32_t              seed = 0;
    int                        rt2  = 4;
    std::unique_ptr<cs::Model> model;
};

namespace {

template <DType S, int RT>
std::unique_ptr<cs::Model> build_model(const cs_gemver_s& g) {
    auto m = std::make_unique<cs::Model>(cs::create_model<gemver::Gemver<S, RT>>(
        static_cast<gemver::Style>(g.style), static_cast<int>(g.n), static_cast<int>(g.R), g.alpha, g.beta, gemver::InputSeeds{g.seed}));
    m->setup();
    return m;
}

template <DType S>
std::unique_ptr<cs::Model> build_rt(const cs_gemver_s& g) {
    switch (g.rt2) {
        case 4: return build_model<S, 4>(g);
        case 16: return build_model<S, 16>(g);
        case 32: return build_model<S, 32>(g);
        default: return build_model<S, 8>(g);
    }
}

cs::Model& model(cs_gemver_s& g) {
    if (!g.model) {
        switch (g.store) {
            case 1: g.model = build_rt<DType::bf16>(g); break;
            case 2: g.model = build_rt<DType::e4m3>(g); break;
            default: g.model = g.style >= CS_GEMVER_TWOPASS ? build_rt<DType::f32>(g) : build_model<DType::f32, 8>(g); break;
        }
    }
    return *g.model;
}

}  // namespace

extern "C" {

int cs_gemver_create(cs_gemver* out, int style, int store, size_t n, size_t R, float alpha, float beta, uint32_t base_seed) {
    return wl::guard([&] {
        if (!out) throw std::invalid_argument("cs_gemver_create: null handle");
        if (style < CS_GEMVER_UNFUSED || style > CS_GEMVER_TWOPASS_1) throw std::invalid_argument("gemver: unknown style");
        if (store < 0 || store > 2) throw std::invalid_argument("gemver: unknown storage");
        if (store != 0 && style < CS_GEMVER_TWOPASS) throw std::invalid_argument("gemver: reduced-precision storage needs a two-pass style");
        if (n == 0) throw std::invalid_argument("gemver-chain: n must be positive");
        if (R == 0) throw std::invalid_argument("gemver-chain: R must be positive");
        // the reference's problem constraints for its own components (examples/gemver-chain/common/chain_data.hpp:76-82)
        if (style < CS_GEMVER_TWOPASS && (n % 256 != 0 || n % R != 0))
            throw std::invalid_argument("gemver-chain: n must be a positive multiple of 256 and of R (n = " + std::to_string(n) +
                                        ", R = " + std::to_string(R) + ")");
        // the two-pass form takes any n whose rows stay 16-byte aligned in every storage
        if (style >= CS_GEMVER_TWOPASS && n % 16 != 0)
            throw std::invalid_argument("gemver: the two-pass form keeps rows 16-byte aligned: n must be a multiple of 16");
        auto g   = std::make_unique<cs_gemver_s>();
        g->style = style;
        g->store = store;
        g->n     = n;
        g->R     = R;
        g->alpha = alpha;
        g->beta  = beta;
        g->seed  = base_seed;
        *out     = g.release();
    });
}

int cs_gemver_set_tile(cs_gemver g, int pass2_rows) {
    return wl::guard([&] {
        if (!g) throw std::invalid_argument("cs_gemver_set_tile: null handle");
        if (pass2_rows != 4 && pass2_rows != 8 && pass2_rows != 16 && pass2_rows != 32)
            throw std::invalid_argument("gemver: pass-2 rows are 4, 8, 16 or 32");
        if (g->model) throw std::logic_error("gemver: the tile is fixed once the model has run");
        g->rt2 = pass2_rows;
    });
}

int cs_gemver_run(cs_gemver g) {
    return wl::guard([&] {
        if (!g) throw std::invalid_argument("cs_gemver_run: null handle");
        model(*g).run();
    });
}

int cs_gemver_wait(cs_gemver g) {
    return wl::guard([&] {
        if (!g) throw std::invalid_argument("cs_gemver_wait: null handle");
        model(*g).wait();
    });
}

int cs_gemver_read(cs_gemver g, const char* name, void* host, size_t bytes) {
    return wl::guard([&] {
        if (!g || !name || !host) throw std::invalid_argument("cs_gemver_read: null argument");
        cs::Model&        m = model(*g);
        const std::string s = std::string{name} == "scale" ? "row_scale" : std::string{name};
        m.wait();
        const std::vector<std::size_t> shape = m.shape_of(s);
        std::size_t                    elems = 1;
        for (std::size_t e : shape) elems *= e;
        const std::size_t elem = s == "A" ? (g->store == 0 ? 4 : g->store == 1 ? 2 : 1) : 4;
        const std::size_t need = elems * elem;
        if (bytes != need) throw std::invalid_argument("gemver: '" + s + "' holds " + std::to_string(need) + " bytes, not " + std::to_string(bytes));
        void* src = m.device_data(s);
        if (!src) throw std::logic_error("gemver: '" + s + "' holds no data yet (run the model first)");
        cs::detail::hip_check(cs_hip_memcpy_async(host, src, need, m.stream()), "gemver read");
        cs::detail::hip_check(cs_hip_stream_synchronize(m.stream()), "gemver read");
    });
}

int cs_gemver_time(cs_gemver g, int iters, float* ms_per_run) {
    return wl::guard([&] {
        if (!g || !ms_per_run || iters <= 0) throw std::invalid_argument("cs_gemver_time: bad argument");
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

int cs_gemver_time_launches(cs_gemver g, int iters, float* ms, int cap, int* count) {
    return wl::guard([&] {
        if (!g || !ms || !count || iters <= 0) throw std::invalid_argument("cs_gemver_time_launches: bad argument");
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

int cs_gemver_time_stage(cs_gemver g, int leaf, int iters, float* ms) {
    return wl::guard([&] {
        if (!g || !ms || leaf < 0 || iters <= 0) throw std::invalid_argument("cs_gemver_time_stage: bad argument");
        *ms = model(*g).time_leaf(static_cast<std::size_t>(leaf), iters);
    });
}

int cs_gemver_stats(cs_gemver g, uint64_t* stats4) {
    return wl::guard([&] {
        if (!g || !stats4) throw std::invalid_argument("cs_gemver_stats: null argument");
        const cs::ModelStats s = model(*g).stats();
        stats4[0]              = s.runs;
        stats4[1]              = s.launches;
        stats4[2]              = s.captures;
        stats4[3]              = s.replays;
    });
}

int cs_gemver_set_replay(cs_gemver g, int mode) {
    return wl::guard([&] {
        if (!g) throw std::invalid_argument("cs_gemver_set_replay: null handle");
        if (mode != CS_REPLAY_GRAPH && mode != CS_REPLAY_DIRECT) throw std::invalid_argument("gemver: replay mode is CS_REPLAY_GRAPH or CS_REPLAY_DIRECT");
        model(*g).set_replay_mode(mode == CS_REPLAY_DIRECT ? cs::ReplayMode::direct : cs::ReplayMode::graph);
    });
}

int cs_gemver_destroy(cs_gemver g) {
    return wl::guard([&] {
        if (!g) return;
        delete g;
    });
}

}  // extern "C"
