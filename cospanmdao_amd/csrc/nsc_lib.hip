// cospanmdao_amd/csrc/nsc_lib.hip — the compressible NS residual model behind the C boundary of include/cospan_workloads.h.
//
// A handle owns a cs::Model whose root is workloads/nsc_systems.hpp's NscResidual<Lim, Walls, TY> in one of its variants.
// The state comes from the device generator (the reference's init_cons, physics.hpp:199-214: uniform flow plus a smooth
// perturbation of fixed 32-cell wavelength) or from the host.

#include <hip/hip_runtime.h>

#include <cmath>
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
#include "../workloads/nsc_systems.hpp"

namespace cs = naga::cospan;
using cs::DType;

namespace {

/// U of cell (j, i): density 1 + 0.1 amp sx sy, velocity (0.3 + 0.05 amp cx sy, 0.2 + 0.05 amp sx cy), pressure
/// 1 + 0.05 amp cx cy, with s/c = sin/cos of 2 pi / 32 times the cell centre.
__global__ __launch_bounds__(256) void init_kernel(float* __restrict__ U, int nx, int ny, float amp) {
    const std::size_t nc = static_cast<std::size_t>(nx) * static_cast<std::size_t>(ny);
    for (std::size_t t = blockIdx.x * static_cast<std::size_t>(blockDim.x) + threadIdx.x; t < nc; t += static_cast<std::size_t>(gridDim.x) * blockDim.x) {
        const int   j = static_cast<int>(t / static_cast<std::size_t>(nx)), i = static_cast<int>(t % static_cast<std::size_t>(nx));
        const float k = static_cast<float>(2.0 * 3.14159265358979323846 / 32.0);
        const float x = static_cast<float>(i) + 0.5f, y = static_cast<float>(j) + 0.5f;
        const float sx = sinf(k * x), cx = cosf(k * x), sy = sinf(k * y), cy = cosf(k * y);
        const float r = 1.0f + 0.1f * amp * sx * sy, u = 0.3f + 0.05f * amp * cx * sy, v = 0.2f + 0.05f * amp * sx * cy;
        const float p = 1.0f + 0.05f * amp * cx * cy;
        U[t]          = r;
        U[nc + t]     = r * u;
        U[2 * nc + t] = r * v;
        U[3 * nc + t] = p / (1.4f - 1.0f) + 0.5f * r * (u * u + v * v);
    }
}

}  // namespace

struct cs_nsc_s {
    int                        variant = 0, limiter = 0, walls = 1, ty = 8;
    int                        nx = 0, ny = 0;
    std::unique_ptr<cs::Model> model;
};

namespace {

template <int Lim, bool Walls, int TY>
std::unique_ptr<cs::Model> build(const cs_nsc_s& g) {
    auto m = std::make_unique<cs::Model>(cs::create_model<nsc::NscResidual<Lim, Walls, TY>>(static_cast<nsc::Variant>(g.variant)));
    m->setup();
    return m;
}

template <int Lim, bool Walls>
std::unique_ptr<cs::Model> build_ty(const cs_nsc_s& g) {
    return g.ty == 16 ? build<Lim, Walls, 16>(g) : build<Lim, Walls, 8>(g);
}

cs::Model& model(cs_nsc_s& g) {
    if (!g.model) {
        if (g.limiter == CS_NSC_VANLEER)
            g.model = g.walls ? build_ty<nsc::kVanLeer, true>(g) : build_ty<nsc::kVanLeer, false>(g);
        else
            g.model = g.walls ? build_ty<nsc::kMinMod, true>(g) : build_ty<nsc::kMinMod, false>(g);
    }
    return *g.model;
}

cs::InitShape state_shape(const cs_nsc_s& g) {
    cs::InitShape s;
    s.rank       = 3;
    s.extents[0] = 4;
    s.extents[1] = static_cast<std::size_t>(g.ny);
    s.extents[2] = static_cast<std::size_t>(g.nx);
    return s;
}

}  // namespace

extern "C" {

int cs_nsc_create(cs_nsc* out, int variant, size_t nx, size_t ny, int limiter, int walls, int tile_rows) {
    return wl::guard([&] {
        if (!out) throw std::invalid_argument("cs_nsc_create: null handle");
        if (variant < CS_NSC_UNFUSED || variant > CS_NSC_LAUNCHES) throw std::invalid_argument("nsc: unknown variant");
        if (limiter != CS_NSC_MINMOD && limiter != CS_NSC_VANLEER) throw std::invalid_argument("nsc: the limiter is minmod (0) or van Leer (1)");
        if (walls != 0 && walls != 1) throw std::invalid_argument("nsc: walls is 0 (periodic in y) or 1 (no-slip walls in y)");
        if (tile_rows != 8 && tile_rows != 16) throw std::invalid_argument("nsc: tile rows are 8 or 16");
        if (nx == 0 || ny == 0 || nx % nsc::kTX != 0 || ny % static_cast<size_t>(tile_rows) != 0)
            throw std::invalid_argument("nsc: nx must be a positive multiple of 64 and ny of the tile rows");
        if (nx > (1u << 16) || ny > (1u << 16)) throw std::invalid_argument("nsc: nx and ny are at most 65536");
        auto g     = std::make_unique<cs_nsc_s>();
        g->variant = variant;
        g->limiter = limiter;
        g->walls   = walls;
        g->ty      = tile_rows;
        g->nx      = static_cast<int>(nx);
        g->ny      = static_cast<int>(ny);
        *out       = g.release();
    });
}

int cs_nsc_init(cs_nsc g, float amp) {
    return wl::guard([&] {
        if (!g) throw std::invalid_argument("cs_nsc_init: null handle");
        const int nx = g->nx, ny = g->ny;
        model(*g).update_input("U", cs::InitSpec{[nx, ny, amp](cs::detail::InitCtx& c) {
                                                     hipLaunchKernelGGL(init_kernel, dim3(2048), dim3(256), 0, cs_hip::raw(c.stream),
                                                                        static_cast<float*>(c.var.data), nx, ny, amp);
                                                     cs::detail::hip_check(cs_hip::check(hipGetLastError(), "nsc init"), "nsc init");
                                                 },
                                                 state_shape(*g)});
    });
}

int cs_nsc_write(cs_nsc g, const void* host, size_t bytes) {
    return wl::guard([&] {
        if (!g || !host) throw std::invalid_argument("cs_nsc_write: null argument");
        const std::size_t need = 16 * static_cast<std::size_t>(g->nx) * static_cast<std::size_t>(g->ny);
        if (bytes != need) throw std::invalid_argument("nsc: U takes " + std::to_string(need) + " bytes, not " + std::to_string(bytes));
        auto copy = std::make_shared<std::vector<std::byte>>(static_cast<const std::byte*>(host), static_cast<const std::byte*>(host) + bytes);
        model(*g).update_input("U", cs::InitSpec{[copy](cs::detail::InitCtx& c) {
                                                     cs::detail::hip_check(cs_hip_memcpy_async(c.var.data, copy->data(), copy->size(), c.stream), "nsc write");
                                                     cs::detail::hip_check(cs_hip_stream_synchronize(c.stream), "nsc write");
                                                 },
                                                 state_shape(*g)});
    });
}

int cs_nsc_run(cs_nsc g) {
    return wl::guard([&] {
        if (!g) throw std::invalid_argument("cs_nsc_run: null handle");
        model(*g).run();
    });
}

int cs_nsc_wait(cs_nsc g) {
    return wl::guard([&] {
        if (!g) throw std::invalid_argument("cs_nsc_wait: null handle");
        model(*g).wait();
    });
}

int cs_nsc_read(cs_nsc g, const char* name, void* host, size_t bytes) {
    return wl::guard([&] {
        if (!g || !name || !host) throw std::invalid_argument("cs_nsc_read: null argument");
        static const char* const names[] = {"U", "R", "Part", "Dt", "Ws", "Sq", "DetDt", "DetWs", "DetSq", "Err"};
        const std::string        s       = name;
        bool                     known   = false;
        for (const char* n : names) known = known || s == n;
        if (!known) throw std::invalid_argument("nsc: no output named '" + s + "'");
        cs::Model& m = model(*g);
        m.wait();
        const std::vector<std::size_t> shape = m.shape_of(s);
        std::size_t                    need  = 4;
        for (std::size_t e : shape) need *= e;
        if (bytes != need) throw std::invalid_argument("nsc: '" + s + "' holds " + std::to_string(need) + " bytes, not " + std::to_string(bytes));
        void* src = m.device_data(s);
        if (!src) throw std::logic_error("nsc: '" + s + "' holds no data yet (run the model first)");
        cs::detail::hip_check(cs_hip_memcpy_async(host, src, need, m.stream()), "nsc read");
        cs::detail::hip_check(cs_hip_stream_synchronize(m.stream()), "nsc read");
    });
}

int cs_nsc_time(cs_nsc g, int iters, float* ms_per_run) {
    return wl::guard([&] {
        if (!g || !ms_per_run || iters <= 0) throw std::invalid_argument("cs_nsc_time: bad argument");
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

int cs_nsc_time_launches(cs_nsc g, int iters, float* ms, int cap, int* count) {
    return wl::guard([&] {
        if (!g || !ms || !count || iters <= 0) throw std::invalid_argument("cs_nsc_time_launches: bad argument");
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

int cs_nsc_stats(cs_nsc g, uint64_t* stats4) {
    return wl::guard([&] {
        if (!g || !stats4) throw std::invalid_argument("cs_nsc_stats: null argument");
        const cs::ModelStats s = model(*g).stats();
        stats4[0]              = s.runs;
        stats4[1]              = s.launches;
        stats4[2]              = s.captures;
        stats4[3]              = s.replays;
    });
}

int cs_nsc_destroy(cs_nsc g) {
    return wl::guard([&] { delete g; });
}

}  // extern "C"
