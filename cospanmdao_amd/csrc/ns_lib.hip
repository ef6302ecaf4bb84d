// cospanmdao_amd/csrc/ns_lib.hip — the NS-residual model behind the C boundary of include/cospan_workloads.h.
//
// A handle owns a cs::Model whose root is workloads/ns_systems.hpp's NsResidual<S> (S = f32 or bf16 fields). The fields
// are generated on the device from host FP64 mode tables (workloads/inputs.hpp).

#include <hip/hip_runtime.h>

#include <cstring>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "cospan_hip.h"
#include "cospan_workloads.h"
#include "internal.hpp"
#include "internal_workloads.hpp"
#include "naga/cospan/cospan.hpp"
#include "../workloads/inputs.hpp"
#include "../workloads/ns_systems.hpp"

namespace cs = naga::cospan;
using cs::DType;

namespace {

/// out[i][j] = sum_m amp[m] (cx[m][j] cy[m][i] - sx[m][j] sy[m][i]) in FP64, mode order, no contraction; then rounded to the
/// field's storage (f32, or bf16 from the f32 value by round to nearest even).
template <DType S>
__global__ __launch_bounds__(256) void field_kernel(const double* __restrict__ t, int nx, int ny, cs::cpp_type_t<S>* __restrict__ out) {
    const int j = static_cast<int>(blockIdx.x * blockDim.x + threadIdx.x);
    if (j >= nx) return;
    const double* amp = t;
    const double* cx  = amp + 8;
    const double* sx  = cx + 8 * static_cast<std::size_t>(nx);
    const double* cy  = sx + 8 * static_cast<std::size_t>(nx);
    const double* sy  = cy + 8 * static_cast<std::size_t>(ny);
    for (int i = static_cast<int>(blockIdx.y); i < ny; i += static_cast<int>(gridDim.y)) {
        double s = 0.0;
#pragma unroll
        for (int m = 0; m < 8; ++m)
            s = __dadd_rn(s, __dmul_rn(amp[m], __dsub_rn(__dmul_rn(cx[m * nx + j], cy[m * ny + i]), __dmul_rn(sx[m * nx + j], sy[m * ny + i]))));
        const float f = __double2float_rn(s);
        if constexpr (S == DType::f32)
            out[static_cast<std::size_t>(i) * nx + j] = f;
        else
            out[static_cast<std::size_t>(i) * nx + j] = cs::bf16::from_float(f);
    }
}

template <DType S>
ns::FieldInit field_init(std::uint32_t seed, int nx, int ny) {
    return [seed, nx, ny](int field) {
        return cs::init::custom<S, 2>(
            [seed, nx, ny, field](cs_hip_stream s, cs::Initializer<S, 2> a) {
                const std::vector<double> t = inputs::ns_tables(seed, field, nx, ny);
                void*                     d = nullptr;
                cs::detail::hip_check(cs_hip_malloc(&d, t.size() * 8), "field tables");
                cs::detail::hip_check(cs_hip_memcpy_async(d, t.data(), t.size() * 8, s), "field tables");
                hipLaunchKernelGGL(field_kernel<S>, dim3((nx + 255) / 256, ny < 2048 ? ny : 2048), dim3(256), 0, cs_hip::raw(s),
                                   static_cast<const double*>(d), nx, ny, cs::raw::data(a));
                cs::detail::hip_check(cs_hip_stream_synchronize(s), "field initializer");
                cs::detail::hip_check(cs_hip_free(d), "field tables");
            },
            cs::Shape{static_cast<std::size_t>(ny), static_cast<std::size_t>(nx)});
    };
}

}  // namespace

struct cs_ns_s {
    int                        style = 0, store = 0;
    int                        nx = 0, ny = 0, ty = 16;
    double                     h = 0, nu = 0;
    std::uint32_t              seed = 0;
    std::unique_ptr<cs::Model> model;
};

namespace {

template <DType S>
std::unique_ptr<cs::Model> build(const cs_ns_s& g) {
    auto m = std::make_unique<cs::Model>(cs::create_model<ns::NsResidual<S>>(static_cast<ns::Style>(g.style), g.nx, g.ny, g.h, g.nu, g.ty,
                                                                            field_init<S>(g.seed, g.nx, g.ny)));
    m->setup();
    return m;
}

cs::Model& model(cs_ns_s& g) {
    if (!g.model) g.model = g.store == 1 ? build<DType::bf16>(g) : build<DType::f32>(g);
    return *g.model;
}

}  // namespace

extern "C" {

int cs_ns_create(cs_ns* out, int style, int store, size_t nx, size_t ny, double h, double nu, uint32_t seed, int tile_rows) {
    return wl::guard([&] {
        if (!out) throw std::invalid_argument("cs_ns_create: null handle");
        if (style != CS_NS_FUSED && style != CS_NS_UNFUSED) throw std::invalid_argument("ns: unknown style");
        if (store != 0 && store != 1) throw std::invalid_argument("ns: storage is 0 (f32) or 1 (bf16)");
        if (nx < 8 || ny < 2 || nx % 8 != 0) throw std::invalid_argument("ns: nx must be a multiple of 8 (>= 8) and ny >= 2");
        if (nx > (1u << 20) || ny > (1u << 20)) throw std::invalid_argument("ns: nx and ny are at most 2^20");
        if (!(h > 0.0)) throw std::invalid_argument("ns: h must be positive");
        if (tile_rows < 1 || tile_rows > 1024) throw std::invalid_argument("ns: tile rows must be in [1, 1024]");
        auto g   = std::make_unique<cs_ns_s>();
        g->style = style;
        g->store = store;
        g->nx    = static_cast<int>(nx);
        g->ny    = static_cast<int>(ny);
        g->h     = h;
        g->nu    = nu;
        g->seed  = seed;
        g->ty    = tile_rows;
        *out     = g.release();
    });
}

int cs_ns_run(cs_ns g) {
    return wl::guard([&] {
        if (!g) throw std::invalid_argument("cs_ns_run: null handle");
        model(*g).run();
    });
}

int cs_ns_wait(cs_ns g) {
    return wl::guard([&] {
        if (!g) throw std::invalid_argument("cs_ns_wait: null handle");
        model(*g).wait();
    });
}

int cs_ns_read(cs_ns g, const char* name, void* host, size_t bytes) {
    return wl::guard([&] {
        if (!g || !name || !host) throw std::invalid_argument("cs_ns_read: null argument");
        cs::Model&        m = model(*g);
        const std::string s = name;
        m.wait();
        const std::vector<std::size_t> shape = m.shape_of(s);
        std::size_t                    elems = 1;
        for (std::size_t e : shape) elems *= e;
        std::size_t elem = 4;
        if (s == "sq" || s == "norm") elem = 8;
        if ((s == "u" || s == "v" || s == "p") && g->store == 1) elem = 2;
        const std::size_t need = elems * elem;
        if (bytes != need) throw std::invalid_argument("ns: '" + s + "' holds " + std::to_string(need) + " bytes, not " + std::to_string(bytes));
        void* src = m.device_data(s);
        if (!src) throw std::logic_error("ns: '" + s + "' holds no data yet (run the model first)");
        cs::detail::hip_check(cs_hip_memcpy_async(host, src, need, m.stream()), "ns read");
        cs::detail::hip_check(cs_hip_stream_synchronize(m.stream()), "ns read");
    });
}

int cs_ns_write(cs_ns g, const char* name, const void* host, size_t bytes) {
    return wl::guard([&] {
        if (!g || !name || !host) throw std::invalid_argument("cs_ns_write: null argument");
        const std::string s = name;
        if (s != "u" && s != "v" && s != "p") throw std::invalid_argument("ns: only the fields u, v, p can be written");
        const std::size_t cells = static_cast<std::size_t>(g->nx) * static_cast<std::size_t>(g->ny);
        const std::size_t need  = cells * (g->store == 1 ? 2 : 4);
        if (bytes != need) throw std::invalid_argument("ns: '" + s + "' takes " + std::to_string(need) + " bytes, not " + std::to_string(bytes));
        auto             copy = std::make_shared<std::vector<std::byte>>(static_cast<const std::byte*>(host), static_cast<const std::byte*>(host) + bytes);
        cs::InitShape    shape;
        shape.rank       = 2;
        shape.extents[0] = static_cast<std::size_t>(g->ny);
        shape.extents[1] = static_cast<std::size_t>(g->nx);
        model(*g).update_input(s, cs::InitSpec{[copy](cs::detail::InitCtx& c) {
                                                   cs::detail::hip_check(cs_hip_memcpy_async(c.var.data, copy->data(), copy->size(), c.stream), "ns write");
                                                   cs::detail::hip_check(cs_hip_stream_synchronize(c.stream), "ns write");
                                               },
                                               shape});
    });
}

int cs_ns_time(cs_ns g, int iters, float* ms_per_run) {
    return wl::guard([&] {
        if (!g || !ms_per_run || iters <= 0) throw std::invalid_argument("cs_ns_time: bad argument");
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

int cs_ns_time_launches(cs_ns g, int iters, float* ms, int cap, int* count) {
    return wl::guard([&] {
        if (!g || !ms || !count || iters <= 0) throw std::invalid_argument("cs_ns_time_launches: bad argument");
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

int cs_ns_set_replay(cs_ns g, int mode) {
    return wl::guard([&] {
        if (!g) throw std::invalid_argument("cs_ns_set_replay: null handle");
        if (mode != CS_REPLAY_GRAPH && mode != CS_REPLAY_DIRECT) throw std::invalid_argument("ns: replay mode is CS_REPLAY_GRAPH or CS_REPLAY_DIRECT");
        model(*g).set_replay_mode(mode == CS_REPLAY_DIRECT ? cs::ReplayMode::direct : cs::ReplayMode::graph);
    });
}

int cs_ns_destroy(cs_ns g) {
    return wl::guard([&] { delete g; });
}

}  // extern "C"
