// cospanmdao_amd/csrc/batch_lib.hip — the batched ragged GEMVER model behind the C boundary, with its autotuner.
//
// A handle owns a cs::Model whose root is workloads/batch_systems.hpp's BatchedGemver, in its phased form (plan, scan,
// unit map, pass 1, column reductions, pass 2: six launches) unless cs_batch_set_mode asks for the persistent one (plan,
// scan, unit map, then every unit of the three phases in one work-queue kernel). Runs after the first two are replays of one captured graph. The autotuner times every candidate
// (strip rows R, pass-2 rows RT and, persistent form only, work-groups G) on the model itself and keeps the fastest;
// results are cached per device, problem and form in a small text file so later processes skip the search.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <memory>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "cospan_hip.h"
#include "cospan_workloads.h"
#include "internal.hpp"
#include "internal_workloads.hpp"
#include "naga/cospan/cospan.hpp"
#include "../workloads/batch_systems.hpp"
#include "../workloads/inputs.hpp"

namespace cs = naga::cospan;
using cs::DType;

namespace {

/// A of every instance: element (i, j) of instance k = hash_unit(seed, base_k + i N_k + j) at A + 4 a4off_k + i ld_k + j.
/// grid: x = columns, y = rows (strided), z = instances. The padding columns [N_k, ld_k) are zero.
__global__ __launch_bounds__(256) void batch_a_kernel(std::uint32_t seed, const std::uint32_t* sizes, const std::uint64_t* a4off,
                                                      const std::uint64_t* base, float* A) {
    const std::uint32_t k = blockIdx.z, n = sizes[k], ld = batch::ld_of(n);
    const std::uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= ld) return;
    float* a = A + 4 * a4off[k];
    for (std::uint32_t i = blockIdx.y; i < n; i += gridDim.y)
        a[static_cast<std::size_t>(i) * ld + j] = j < n ? inputs::hash_unit(seed, base[k] + static_cast<std::uint64_t>(i) * n + j) : 0.0f;
}

batch::AInit a_init(std::uint32_t seed) {
    return [seed](const batch::Layout& L) {
        auto layout = std::make_shared<batch::Layout>(L);
        return cs::init::custom<DType::f32, 1>(
            [seed, layout](cs_hip_stream s, cs::Initializer<DType::f32, 1> a) {
                const std::size_t          K = layout->sizes.size();
                std::vector<std::uint64_t> tab(2 * K);
                std::uint64_t              off = 0, base = 0;
                std::uint32_t              maxn = 0;
                for (std::size_t k = 0; k < K; ++k) {
                    const std::uint32_t n = layout->sizes[k];
                    tab[k]                = off;
                    tab[K + k]            = base;
                    off += static_cast<std::uint64_t>(n) * batch::ld_of(n) / 4;
                    base += static_cast<std::uint64_t>(n) * n;
                    maxn = std::max(maxn, n);
                }
                void* d = nullptr;
                cs::detail::hip_check(cs_hip_malloc(&d, tab.size() * 8 + K * 4), "batch init tables");
                auto* dtab   = static_cast<std::uint64_t*>(d);
                auto* dsizes = reinterpret_cast<std::uint32_t*>(dtab + 2 * K);
                cs::detail::hip_check(cs_hip_memcpy_async(dtab, tab.data(), tab.size() * 8, s), "batch init tables");
                cs::detail::hip_check(cs_hip_memcpy_async(dsizes, layout->sizes.data(), K * 4, s), "batch init tables");
                const dim3 grid((batch::ld_of(maxn) + 255) / 256, 256, static_cast<unsigned>(K));
                hipLaunchKernelGGL(batch_a_kernel, grid, dim3(256), 0, cs_hip::raw(s), seed, dsizes, dtab, dtab + K, cs::raw::data(a));
                cs::detail::hip_check(hipGetLastError() == hipSuccess ? CS_HIP_OK : CS_HIP_ERR_RUNTIME, "batch A initializer");
                cs::detail::hip_check(cs_hip_stream_synchronize(s), "batch A initializer");
                cs::detail::hip_check(cs_hip_free(d), "batch init tables");
            },
            cs::Shape{4 * layout->a4_total});
    };
}

/// N_k = round(lo (hi / lo)^U_k), U_k in [0, 1) from the hash (seed, k): log-uniform sizes.
std::vector<std::uint32_t> ragged_sizes(std::uint32_t seed, std::uint32_t count, std::uint32_t lo, std::uint32_t hi) {
    std::vector<std::uint32_t> s(count);
    const double               ratio = static_cast<double>(hi) / static_cast<double>(lo);
    for (std::uint32_t k = 0; k < count; ++k) {
        const double u = (static_cast<double>(inputs::hash_unit(seed, k)) + 1.0) * 0.5;
        long long    r = std::llround(static_cast<double>(lo) * std::pow(ratio, u));
        r              = std::min<long long>(std::max<long long>(r, lo), hi);
        s[k]           = static_cast<std::uint32_t>(r);
    }
    return s;
}

}  // namespace

struct cs_batch_s {
    std::shared_ptr<batch::Layout> layout;
    std::shared_ptr<batch::Knobs>  knobs = std::make_shared<batch::Knobs>();
    std::vector<std::uint32_t>     params{64, 8, 1024};
    bool                           persistent = false;
    float                          alpha = 1.5f, beta = 1.2f;
    std::uint32_t                  seed = 0;
    std::unique_ptr<cs::Model>     model;
};

namespace {

/// Whether cs_batch_set_mode may select the persistent work-queue form (RaggedGemver).
constexpr bool kPersistentFormEnabled = true;

cs::Model& model(cs_batch_s& g) {
    if (!g.model) {
        g.knobs->groups = static_cast<int>(g.params[2]);
        batch::count_units(g.layout->sizes.data(), g.layout->sizes.size(), g.params[0], g.params[1], g.knobs->units);
        g.model = std::make_unique<cs::Model>(cs::create_model<batch::BatchedGemver>(g.layout, g.knobs, g.params, g.alpha, g.beta, g.seed,
                                                                                     a_init(g.seed * 16u + 1u), !g.persistent));
        g.model->setup();
    }
    return *g.model;
}

void set_params(cs_batch_s& g, std::uint32_t R, std::uint32_t RT, std::uint32_t G) {
    if (R < batch::BatchedGemver::kMinR) throw std::invalid_argument("batch: R must be at least 32 (the partials are sized for it)");
    if (RT != 4 && RT != 8 && RT != 16 && RT != 32) throw std::invalid_argument("batch: RT is 4, 8, 16 or 32");
    if (G < 1 || G > 65536) throw std::invalid_argument("batch: the persistent grid is 1 .. 65536 work-groups");
    g.params = {R, RT, G};
    g.knobs->groups = static_cast<int>(G);
    batch::count_units(g.layout->sizes.data(), g.layout->sizes.size(), R, RT, g.knobs->units);
    if (g.model) g.model->update_input("params", cs::init::from_host(g.params));
}

float time_runs(cs::Model& m, int iters) {
    m.wait();
    wl::EventPair ev;
    ev.record_start(m.stream());
    for (int i = 0; i < iters; ++i) m.run();
    ev.record_stop(m.stream());
    const float ms = ev.elapsed_ms();
    m.wait();
    return ms / static_cast<float>(iters);
}

std::string cache_key(const cs_batch_s& g) {
    char name[64] = {0};
    int  cus      = 0;
    cs_hip_device_info(0, name, sizeof name, &cus, nullptr, nullptr);
    std::uint64_t h = 1469598103934665603ull;
    for (std::uint32_t n : g.layout->sizes) h = (h ^ n) * 1099511628211ull;
    std::ostringstream o;
    o << name << ":cus" << cus << ":K" << g.layout->sizes.size() << ":sizes" << std::hex << h << (g.persistent ? ":persistent" : ":phased");
    return o.str();
}

bool cache_lookup(const std::string& path, const std::string& key, std::uint32_t (&p)[3]) {
    std::ifstream f(path);
    std::string   k;
    std::uint32_t a, b, c;
    while (f >> k >> a >> b >> c)
        if (k == key) {
            p[0] = a;
            p[1] = b;
            p[2] = c;
            return true;
        }
    return false;
}

void cache_store(const std::string& path, const std::string& key, const std::uint32_t (&p)[3]) {
    std::ofstream f(path, std::ios::app);
    if (f) f << key << ' ' << p[0] << ' ' << p[1] << ' ' << p[2] << '\n';
}

}  // namespace

extern "C" {

int cs_batch_create(cs_batch* out, uint32_t count, uint32_t lo, uint32_t hi, float alpha, float beta, uint32_t base_seed) {
    return wl::guard([&] {
        if (!out) throw std::invalid_argument("cs_batch_create: null handle");
        if (count == 0 || count > 65536) throw std::invalid_argument("batch: 1 .. 65536 instances");
        if (lo == 0 || hi < lo || hi > 65536) throw std::invalid_argument("batch: sizes must satisfy 1 <= lo <= hi <= 65536");
        auto g    = std::make_unique<cs_batch_s>();
        g->layout = std::make_shared<batch::Layout>(ragged_sizes(base_seed, count, lo, hi), batch::BatchedGemver::kMinR);
        if (g->layout->a4_total > 0xffffffffull) throw std::invalid_argument("batch: the instances' A exceeds 2^32 16-byte units");
        g->alpha = alpha;
        g->beta  = beta;
        g->seed  = base_seed;
        *out     = g.release();
    });
}

int cs_batch_sizes(cs_batch g, uint32_t* sizes, size_t count) {
    return wl::guard([&] {
        if (!g || !sizes || count != g->layout->sizes.size()) throw std::invalid_argument("cs_batch_sizes: bad argument");
        std::memcpy(sizes, g->layout->sizes.data(), count * 4);
    });
}

int cs_batch_set_params(cs_batch g, uint32_t R, uint32_t RT, uint32_t groups) {
    return wl::guard([&] {
        if (!g) throw std::invalid_argument("cs_batch_set_params: null handle");
        set_params(*g, R, RT, groups);
    });
}

int cs_batch_set_mode(cs_batch g, int persistent) {
    return wl::guard([&] {
        if (!g) throw std::invalid_argument("cs_batch_set_mode: null handle");
        if (persistent != 0 && persistent != 1) throw std::invalid_argument("cs_batch_set_mode: mode is 0 (phased) or 1 (persistent)");
        if (persistent == 1 && !kPersistentFormEnabled)
            throw std::invalid_argument("cs_batch_set_mode: the persistent form is switched off in this build");
        if ((persistent != 0) != g->persistent) {
            if (g->model) g->model->wait();
            g->model.reset();  // rebuilt in the other form at the next use
            g->persistent = persistent != 0;
            batch::count_units(g->layout->sizes.data(), g->layout->sizes.size(), g->params[0], g->params[1], g->knobs->units);
        }
    });
}

int cs_batch_get_params(cs_batch g, uint32_t* p3) {
    return wl::guard([&] {
        if (!g || !p3) throw std::invalid_argument("cs_batch_get_params: null argument");
        for (int i = 0; i < 3; ++i) p3[i] = g->params[i];
    });
}

int cs_batch_run(cs_batch g) {
    return wl::guard([&] {
        if (!g) throw std::invalid_argument("cs_batch_run: null handle");
        model(*g).run();
    });
}

int cs_batch_wait(cs_batch g) {
    return wl::guard([&] {
        if (!g) throw std::invalid_argument("cs_batch_wait: null handle");
        model(*g).wait();
    });
}

int cs_batch_read(cs_batch g, const char* name, void* host, size_t bytes) {
    return wl::guard([&] {
        if (!g || !name || !host) throw std::invalid_argument("cs_batch_read: null argument");
        cs::Model&        m = model(*g);
        const std::string s = name;
        m.wait();
        const std::vector<std::size_t> shape = m.shape_of(s);
        std::size_t                    elems = 1;
        for (std::size_t e : shape) elems *= e;
        const std::size_t need = elems * 4;
        if (bytes != need) throw std::invalid_argument("batch: '" + s + "' holds " + std::to_string(need) + " bytes, not " + std::to_string(bytes));
        void* src = m.device_data(s);
        if (!src) throw std::logic_error("batch: '" + s + "' holds no data yet (run the model first)");
        cs::detail::hip_check(cs_hip_memcpy_async(host, src, need, m.stream()), "batch read");
        cs::detail::hip_check(cs_hip_stream_synchronize(m.stream()), "batch read");
    });
}

int cs_batch_time(cs_batch g, int iters, float* ms_per_run) {
    return wl::guard([&] {
        if (!g || !ms_per_run || iters <= 0) throw std::invalid_argument("cs_batch_time: bad argument");
        *ms_per_run = time_runs(model(*g), iters);
    });
}

int cs_batch_time_each(cs_batch g, int iters, float* ms) {
    return wl::guard([&] {
        if (!g || !ms || iters <= 0) throw std::invalid_argument("cs_batch_time_each: bad argument");
        cs::Model& m = model(*g);
        m.wait();
        std::vector<cs_hip_event> ev(static_cast<std::size_t>(iters) + 1);
        for (auto& e : ev) cs::detail::hip_check(cs_hip_event_create(&e, 1), "event");
        cs::detail::hip_check(cs_hip_event_record(ev[0], m.stream()), "event");
        for (int i = 0; i < iters; ++i) {
            m.run();
            cs::detail::hip_check(cs_hip_event_record(ev[static_cast<std::size_t>(i) + 1], m.stream()), "event");
        }
        m.wait();
        for (int i = 0; i < iters; ++i)
            cs::detail::hip_check(cs_hip_event_elapsed_ms(&ms[i], ev[static_cast<std::size_t>(i)], ev[static_cast<std::size_t>(i) + 1]), "event");
        for (auto& e : ev) cs_hip_event_destroy(e);
    });
}

int cs_batch_autotune(cs_batch g, int reps, const char* cache_path, uint32_t* best3) {
    return wl::guard([&] {
        if (!g || reps <= 0) throw std::invalid_argument("cs_batch_autotune: bad argument");
        cs::Model&        m   = model(*g);
        const std::string key = cache_key(*g);
        std::uint32_t     p[3];
        if (cache_path && cache_lookup(cache_path, key, p)) {
            set_params(*g, p[0], p[1], p[2]);
        } else {
            int cus = 0;
            cs::detail::hip_check(cs_hip_device_info(0, nullptr, 0, &cus, nullptr, nullptr), "device info");
            float         best = 1e30f;
            std::uint32_t bp[3] = {64, 8, static_cast<std::uint32_t>(4 * cus)};
            const std::vector<std::uint32_t> gms = g->persistent ? std::vector<std::uint32_t>{2u, 3u, 4u} : std::vector<std::uint32_t>{4u};
            const std::vector<std::uint32_t> Rs{64u, 128u, 256u};
            const std::vector<std::uint32_t> RTs{8u, 16u, 32u};
            for (std::uint32_t R : Rs)
                for (std::uint32_t RT : RTs)
                    for (std::uint32_t gm : gms) {
                        set_params(*g, R, RT, gm * static_cast<std::uint32_t>(cus));
                        m.run();  // eager (new parameters), then a capture, then replays
                        m.run();
                        const float t = time_runs(m, reps);
                        if (t < best) {
                            best  = t;
                            bp[0] = R;
                            bp[1] = RT;
                            bp[2] = gm * static_cast<std::uint32_t>(cus);
                        }
                    }
            set_params(*g, bp[0], bp[1], bp[2]);
            if (cache_path) cache_store(cache_path, key, bp);
        }
        if (best3)
            for (int i = 0; i < 3; ++i) best3[i] = g->params[i];
    });
}

int cs_batch_time_stage(cs_batch g, int leaf, int iters, float* ms) {
    return wl::guard([&] {
        if (!g || !ms || leaf < 0 || iters <= 0) throw std::invalid_argument("cs_batch_time_stage: bad argument");
        *ms = model(*g).time_leaf(static_cast<std::size_t>(leaf), iters);
    });
}

int cs_batch_stats(cs_batch g, uint64_t* stats4) {
    return wl::guard([&] {
        if (!g || !stats4) throw std::invalid_argument("cs_batch_stats: null argument");
        const cs::ModelStats s = model(*g).stats();
        stats4[0]              = s.runs;
        stats4[1]              = s.launches;
        stats4[2]              = s.captures;
        stats4[3]              = s.replays;
    });
}

int cs_batch_destroy(cs_batch g) {
    return wl::guard([&] { delete g; });
}

}  // extern "C"
