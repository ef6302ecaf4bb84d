// cospanmdao_amd/csrc/gemver_lib.hip — the GEMVER model behind the C boundary of include/cospan_workloads.h.
//
// Styles follow examples/gemver-chain/systems.hpp:33-48 (unfused, fused, naive) plus this build's two-pass form; see the
// header for the list. Every style is a sequence of fused-group launches of the components in
// workloads/gemver_components.hpp.

#include <hip/hip_runtime.h>

#include <cstring>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "cospan_hip.h"
#include "cospan_workloads.h"
#include "internal_workloads.hpp"
#include "naga/cospan/group_launch.hpp"
#include "../workloads/gemver_components.hpp"

namespace {

namespace cs = naga::cospan;
namespace cd = naga::cospan::detail;
using cs::DType;
using namespace gemver;

template <class W>
void point(W& w, const void* p, std::size_t d0) {
    const std::size_t shape[1] = {d0}, stride[1] = {1};
    cd::WrapperAccess::point(w, const_cast<void*>(p), shape, stride);
}
template <class W>
void point(W& w, const void* p, std::size_t d0, std::size_t d1, std::size_t ld) {
    const std::size_t shape[2] = {d0, d1}, stride[2] = {ld, 1};
    cd::WrapperAccess::point(w, const_cast<void*>(p), shape, stride);
}

cd::StagePlan plan1(std::uint64_t groups, const char* who) {
    cd::StagePlan p;
    p.units[0] = 1;
    p.units[1] = 1;
    p.units[2] = groups;
    p.wg[0]    = 1;
    p.wg[1]    = 1;
    p.wg[2]    = kWg;
    p.who      = who;
    return p;
}
cd::StagePlan plan2(std::uint64_t g0, std::uint64_t g1, const char* who) {
    cd::StagePlan p = plan1(g1, who);
    p.units[1]      = g0;
    return p;
}

std::uint64_t ceil_div(std::uint64_t a, std::uint64_t b) { return (a + b - 1) / b; }

}  // namespace

struct cs_gemver_s {
    int           style = 0, store = 0;
    std::size_t   n = 0, R = 64;
    float         alpha = 1.0f, beta = 1.0f;
    std::uint32_t seed      = 0;
    int           rt2       = 8;
    bool          ready     = false;
    cs_hip_stream stream    = nullptr;
    wl::DeviceBuffers mem;

    void *A = nullptr, *B = nullptr;
    float *scale = nullptr, *u1 = nullptr, *v1 = nullptr, *u2 = nullptr, *v2 = nullptr, *y = nullptr, *z = nullptr;
    float *x = nullptr, *w = nullptr, *P = nullptr, *alpha_d = nullptr, *beta_d = nullptr;
    std::uint32_t* R_d = nullptr;

    cd::GroupLauncher<Rank2Update>                                        l_r2u;
    cd::GroupLauncher<GemvTPartial>                                       l_gtp;
    cd::GroupLauncher<Rank2Update, GemvTPartial>                          l_r2u_gtp;
    cd::GroupLauncher<ColReduce<true>>                                    l_cred;
    cd::GroupLauncher<Gemv>                                               l_gemv;
    cd::GroupLauncher<Rank2Update, GemvTPartial, ColReduce<true>, Gemv>   l_naive;

    std::size_t a_elem() const { return store == 0 ? 4 : store == 1 ? 2 : 1; }
    std::size_t strips() const { return (n + R - 1) / R; }
};

namespace {

void initialize(cs_gemver_s& g) {
    const std::size_t n = g.n;
    cd::hip_check(cs_hip_init(0), "gemver init");
    auto& m = g.mem;
    g.A     = m.alloc(n * n * g.a_elem());
    g.u1    = m.alloc_f32(n);
    g.v1    = m.alloc_f32(n);
    g.u2    = m.alloc_f32(n);
    g.v2    = m.alloc_f32(n);
    g.y     = m.alloc_f32(n);
    g.z     = m.alloc_f32(n);
    g.x     = m.alloc_f32(n);
    g.w     = m.alloc_f32(n);
    g.scale = m.alloc_f32(n);
    g.P     = m.alloc_f32(g.strips() * n);
    g.alpha_d = m.alloc_f32(1);
    g.beta_d  = m.alloc_f32(1);
    g.R_d     = static_cast<std::uint32_t*>(m.alloc(4));
    if (g.style <= CS_GEMVER_NAIVE) g.B = m.alloc(n * n * 4);
    const std::uint32_t sb = g.seed * 16u;
    // the reference's seeds: seed_A = 1, u1 = 2, v1 = 3, u2 = 4, v2 = 5, y = 6, z = 7 (chain_data.hpp:32)
    if (g.store == 0) {
        cd::hip_check(cs_hip_fill_hash_f32(sb + 1, static_cast<float*>(g.A), n, n, n, g.stream), "init A");
        cd::hip_check(cs_hip_fill_u32_async(g.scale, 0x3f800000u, n, g.stream), "init scale");
    } else {
        void* tmp = nullptr;
        cd::hip_check(cs_hip_malloc(&tmp, n * n * 4), "init A");
        cd::hip_check(cs_hip_fill_hash_f32(sb + 1, static_cast<float*>(tmp), n, n, n, g.stream), "init A");
        if (g.store == 1) {
            cd::hip_check(cs_hip_quantize_bf16(static_cast<float*>(tmp), static_cast<std::uint16_t*>(g.A), n, n, n, n, g.stream), "quantize");
            cd::hip_check(cs_hip_fill_u32_async(g.scale, 0x3f800000u, n, g.stream), "init scale");
        } else {
            cd::hip_check(cs_hip_quantize_e4m3_rows(static_cast<float*>(tmp), static_cast<std::uint8_t*>(g.A), g.scale, n, n, n, n, g.stream),
                          "quantize");
        }
        cd::hip_check(cs_hip_stream_synchronize(g.stream), "init A");
        cd::hip_check(cs_hip_free(tmp), "init A");
    }
    float* vecs[6] = {g.u1, g.v1, g.u2, g.v2, g.y, g.z};
    for (std::uint32_t k = 0; k < 6; ++k) cd::hip_check(cs_hip_fill_hash_f32(sb + 2 + k, vecs[k], 1, n, n, g.stream), "init vectors");
    std::uint32_t words[3];
    std::memcpy(&words[0], &g.alpha, 4);
    std::memcpy(&words[1], &g.beta, 4);
    words[2] = static_cast<std::uint32_t>(g.R);
    cd::hip_check(cs_hip_fill_u32_async(g.alpha_d, words[0], 1, g.stream), "init alpha");
    cd::hip_check(cs_hip_fill_u32_async(g.beta_d, words[1], 1, g.stream), "init beta");
    cd::hip_check(cs_hip_fill_u32_async(g.R_d, words[2], 1, g.stream), "init R");
    cd::hip_check(cs_hip_stream_synchronize(g.stream), "init");
    g.ready = true;
}

template <DType S, int RT>
void twopass_launch(cs_gemver_s& g, std::vector<wl::LaunchMark>* marks) {
    using P1 = AhatTPartial<S>;
    using P2 = AhatGemv<S, RT>;
    const std::size_t n = g.n, strips = g.strips();
    typename P1::Inputs  i1{};
    typename P1::Outputs o1{};
    point(i1.A, g.A, n, n, n);
    point(i1.row_scale, g.scale, n);
    point(i1.u1, g.u1, n);
    point(i1.v1, g.v1, n);
    point(i1.u2, g.u2, n);
    point(i1.v2, g.v2, n);
    point(i1.y, g.y, n);
    point(i1.R, g.R_d, 1);
    point(o1.P, g.P, strips, n, n);
    typename ColReduce<true>::Inputs ic{};
    typename ColReduce<true>::Outputs oc{};
    point(ic.P, g.P, strips, n, n);
    point(ic.z, g.z, n);
    point(ic.scale, g.beta_d, 1);
    point(oc.x, g.x, n);
    typename P2::Inputs  i2{};
    typename P2::Outputs o2{};
    point(i2.A, g.A, n, n, n);
    point(i2.row_scale, g.scale, n);
    point(i2.u1, g.u1, n);
    point(i2.v1, g.v1, n);
    point(i2.u2, g.u2, n);
    point(i2.v2, g.v2, n);
    point(i2.x, g.x, n);
    point(i2.scale, g.alpha_d, 1);
    point(o2.w, g.w, n);

    const cd::StagePlan sp1 = plan2(strips, ceil_div(n, P1::kCols), "'pass1'");
    const cd::StagePlan spc = plan1(ceil_div(n, kWg), "'cred'");
    const cd::StagePlan sp2 = plan1(ceil_div(n, RT), "'pass2'");
    if (g.style == CS_GEMVER_TWOPASS_1) {
        static auto& l = *new cd::GroupLauncher<P1, ColReduce<true>, P2>;  // process-wide, never destroyed
        const cd::StagePlan plans[3] = {sp1, spc, sp2};
        cd::ArgPack<P1, ColReduce<true>, P2> args{};
        args.head            = {{sp1.units[0], sp1.units[1], sp1.units[2]}, i1, o1, {}};
        args.tail.head       = {{spc.units[0], spc.units[1], spc.units[2]}, ic, oc, {}};
        args.tail.tail.head  = {{sp2.units[0], sp2.units[1], sp2.units[2]}, i2, o2, {}};
        std::vector<cd::SegmentPlan> segs(3);
        for (int k = 0; k < 3; ++k) {
            segs[k].first = k;
            segs[k].count = 1;
            segs[k].boundary = k == 0 ? cd::Boundary::none : cd::Boundary::global_barrier;
        }
        wl::mark(marks, g.stream);
        l.launch(cd::Split::flatten, true, segs, plans, args, g.stream, "gemver");
        wl::mark(marks, g.stream);
        return;
    }
    static auto& l1 = *new cd::GroupLauncher<P1>;
    static auto& l2 = *new cd::GroupLauncher<P2>;
    {
        const cd::StagePlan plans[1] = {sp1};
        cd::ArgPack<P1>     args{};
        args.head = {{sp1.units[0], sp1.units[1], sp1.units[2]}, i1, o1, {}};
        wl::mark(marks, g.stream);
        l1.launch(cd::Split::flatten, false, {cd::SegmentPlan{0, 1}}, plans, args, g.stream, "pass1");
    }
    {
        const cd::StagePlan          plans[1] = {spc};
        cd::ArgPack<ColReduce<true>> args{};
        args.head = {{spc.units[0], spc.units[1], spc.units[2]}, ic, oc, {}};
        wl::mark(marks, g.stream);
        g.l_cred.launch(cd::Split::flatten, false, {cd::SegmentPlan{0, 1}}, plans, args, g.stream, "cred");
    }
    {
        const cd::StagePlan plans[1] = {sp2};
        cd::ArgPack<P2>     args{};
        args.head = {{sp2.units[0], sp2.units[1], sp2.units[2]}, i2, o2, {}};
        wl::mark(marks, g.stream);
        l2.launch(cd::Split::flatten, false, {cd::SegmentPlan{0, 1}}, plans, args, g.stream, "pass2");
        wl::mark(marks, g.stream);
    }
}

template <DType S>
void twopass_dispatch(cs_gemver_s& g, std::vector<wl::LaunchMark>* marks) {
    switch (g.rt2) {
        case 4: return twopass_launch<S, 4>(g, marks);
        case 16: return twopass_launch<S, 16>(g, marks);
        default: return twopass_launch<S, 8>(g, marks);
    }
}

void reference_launch(cs_gemver_s& g, std::vector<wl::LaunchMark>* marks) {
    const std::size_t n = g.n, R = g.R, strips = g.strips();
    Rank2Update::Inputs  ir{};
    Rank2Update::Outputs orr{};
    point(ir.A, g.A, n, n, n);
    point(ir.u1, g.u1, n);
    point(ir.v1, g.v1, n);
    point(ir.u2, g.u2, n);
    point(ir.v2, g.v2, n);
    point(ir.R, g.R_d, 1);
    point(orr.B, g.B, n, n, n);
    GemvTPartial::Inputs  ig{};
    GemvTPartial::Outputs og{};
    point(ig.B, g.B, n, n, n);
    point(ig.y, g.y, n);
    point(og.P, g.P, strips, n, n);
    ColReduce<true>::Inputs  ic{};
    ColReduce<true>::Outputs oc{};
    point(ic.P, g.P, strips, n, n);
    point(ic.z, g.z, n);
    point(ic.scale, g.beta_d, 1);
    point(oc.x, g.x, n);
    Gemv::Inputs  iv{};
    Gemv::Outputs ov{};
    point(iv.B, g.B, n, n, n);
    point(iv.x, g.x, n);
    point(iv.scale, g.alpha_d, 1);
    point(ov.y, g.w, n);

    // launch ranges of the reference (components.hpp:121-125, 164-167, 216-219, 262-265), in work-groups of 256
    const std::uint64_t work = (n / R) * n;
    const cd::StagePlan spr  = plan1(ceil_div(work, kWg), "'r2u'");
    const cd::StagePlan spg  = plan1(ceil_div(strips * n, kWg), "'gtp'");
    const cd::StagePlan spc  = plan1(ceil_div(n, kWg), "'cred'");
    const cd::StagePlan spv  = plan1(n, "'gemv'");

    if (g.style == CS_GEMVER_NAIVE) {
        const cd::StagePlan plans[4] = {spr, spg, spc, spv};
        cd::ArgPack<Rank2Update, GemvTPartial, ColReduce<true>, Gemv> args{};
        args.head                     = {{1, 1, spr.units[2]}, ir, orr, {}};
        args.tail.head                = {{1, 1, spg.units[2]}, ig, og, {}};
        args.tail.tail.head           = {{1, 1, spc.units[2]}, ic, oc, {}};
        args.tail.tail.tail.head      = {{1, 1, spv.units[2]}, iv, ov, {}};
        std::vector<cd::SegmentPlan> segs{{0, 2, cd::RangeMode::whole, cd::LocalMode::partitioned, cd::Boundary::none},
                                          {2, 1, cd::RangeMode::own, cd::LocalMode::partitioned, cd::Boundary::global_barrier},
                                          {3, 1, cd::RangeMode::own, cd::LocalMode::partitioned, cd::Boundary::global_barrier}};
        wl::mark(marks, g.stream);
        g.l_naive.launch(cd::Split::flatten, true, segs, plans, args, g.stream, "fw");
        wl::mark(marks, g.stream);
        return;
    }
    if (g.style == CS_GEMVER_FUSED) {
        const cd::StagePlan                           plans[2] = {spr, spg};
        cd::ArgPack<Rank2Update, GemvTPartial>        args{};
        args.head      = {{1, 1, spr.units[2]}, ir, orr, {}};
        args.tail.head = {{1, 1, spg.units[2]}, ig, og, {}};
        wl::mark(marks, g.stream);
        g.l_r2u_gtp.launch(cd::Split::flatten, false, {cd::SegmentPlan{0, 2, cd::RangeMode::whole}}, plans, args, g.stream, "fw");
    } else {
        {
            const cd::StagePlan      plans[1] = {spr};
            cd::ArgPack<Rank2Update> args{};
            args.head = {{1, 1, spr.units[2]}, ir, orr, {}};
            wl::mark(marks, g.stream);
            g.l_r2u.launch(cd::Split::flatten, false, {cd::SegmentPlan{0, 1}}, plans, args, g.stream, "r2u");
        }
        {
            const cd::StagePlan       plans[1] = {spg};
            cd::ArgPack<GemvTPartial> args{};
            args.head = {{1, 1, spg.units[2]}, ig, og, {}};
            wl::mark(marks, g.stream);
            g.l_gtp.launch(cd::Split::flatten, false, {cd::SegmentPlan{0, 1}}, plans, args, g.stream, "gtp");
        }
    }
    {
        const cd::StagePlan          plans[1] = {spc};
        cd::ArgPack<ColReduce<true>> args{};
        args.head = {{1, 1, spc.units[2]}, ic, oc, {}};
        wl::mark(marks, g.stream);
        g.l_cred.launch(cd::Split::flatten, false, {cd::SegmentPlan{0, 1}}, plans, args, g.stream, "cred");
    }
    {
        const cd::StagePlan plans[1] = {spv};
        cd::ArgPack<Gemv>   args{};
        args.head = {{1, 1, spv.units[2]}, iv, ov, {}};
        wl::mark(marks, g.stream);
        g.l_gemv.launch(cd::Split::flatten, false, {cd::SegmentPlan{0, 1}}, plans, args, g.stream, "gemv");
        wl::mark(marks, g.stream);
    }
}

void forward(cs_gemver_s& g, std::vector<wl::LaunchMark>* marks) {
    if (!g.ready) initialize(g);
    if (g.style >= CS_GEMVER_TWOPASS) {
        switch (g.store) {
            case 1: return twopass_dispatch<DType::bf16>(g, marks);
            case 2: return twopass_dispatch<DType::e4m3>(g, marks);
            default: return twopass_dispatch<DType::f32>(g, marks);
        }
    }
    reference_launch(g, marks);
}

}  // namespace

extern "C" {

int cs_gemver_create(cs_gemver* out, int style, int store, size_t n, size_t R, float alpha, float beta, uint32_t base_seed) {
    return wl::guard([&] {
        if (!out) throw std::invalid_argument("cs_gemver_create: null handle");
        if (style < CS_GEMVER_UNFUSED || style > CS_GEMVER_TWOPASS_1) throw std::invalid_argument("gemver: unknown style");
        if (store < 0 || store > 2) throw std::invalid_argument("gemver: unknown storage");
        if (store != 0 && style < CS_GEMVER_TWOPASS) throw std::invalid_argument("gemver: reduced-precision storage needs a two-pass style");
        // the reference's problem constraints (chain_data.hpp:76-82) for its own components; the two-pass form takes any n
        if (n == 0) throw std::invalid_argument("gemver-chain: n must be positive");
        if (R == 0) throw std::invalid_argument("gemver-chain: R must be positive");
        if (style < CS_GEMVER_TWOPASS && (n % 256 != 0 || n % R != 0))
            throw std::invalid_argument("gemver-chain: n must be a positive multiple of 256 and of R (n = " + std::to_string(n) +
                                        ", R = " + std::to_string(R) + ")");
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
        cd::hip_check(cs_hip_init(0), "gemver");
        cd::hip_check(cs_hip_stream_create(&g->stream), "gemver stream");
        *out = g.release();
    });
}

int cs_gemver_set_tile(cs_gemver g, int pass2_rows) {
    return wl::guard([&] {
        if (!g) throw std::invalid_argument("cs_gemver_set_tile: null handle");
        if (pass2_rows != 4 && pass2_rows != 8 && pass2_rows != 16) throw std::invalid_argument("gemver: pass-2 rows are 4, 8 or 16");
        g->rt2 = pass2_rows;
    });
}

int cs_gemver_run(cs_gemver g) {
    return wl::guard([&] {
        if (!g) throw std::invalid_argument("cs_gemver_run: null handle");
        forward(*g, nullptr);
    });
}

int cs_gemver_wait(cs_gemver g) {
    return wl::guard([&] {
        if (!g) throw std::invalid_argument("cs_gemver_wait: null handle");
        cd::hip_check(cs_hip_stream_synchronize(g->stream), "gemver wait");
    });
}

int cs_gemver_read(cs_gemver g, const char* name, void* host, size_t bytes) {
    return wl::guard([&] {
        if (!g || !name || !host) throw std::invalid_argument("cs_gemver_read: null argument");
        if (!g->ready) initialize(*g);
        const std::string s = name;
        const std::size_t n = g->n;
        const void*       src = nullptr;
        std::size_t       need = n * 4;
        if (s == "x") src = g->x;
        else if (s == "w") src = g->w;
        else if (s == "u1") src = g->u1;
        else if (s == "v1") src = g->v1;
        else if (s == "u2") src = g->u2;
        else if (s == "v2") src = g->v2;
        else if (s == "y") src = g->y;
        else if (s == "z") src = g->z;
        else if (s == "scale") src = g->scale;
        else if (s == "A") { src = g->A; need = n * n * g->a_elem(); }
        else if (s == "B") { src = g->B; need = n * n * 4; }
        else if (s == "P") { src = g->P; need = g->strips() * n * 4; }
        else throw std::invalid_argument("gemver: no output or input named '" + s + "'");
        if (!src) throw std::invalid_argument("gemver: '" + s + "' does not exist in this style");
        if (bytes != need) throw std::invalid_argument("gemver: '" + s + "' holds " + std::to_string(need) + " bytes, not " + std::to_string(bytes));
        cd::hip_check(cs_hip_stream_synchronize(g->stream), "gemver read");
        cd::hip_check(cs_hip_memcpy_async(host, src, need, g->stream), "gemver read");
        cd::hip_check(cs_hip_stream_synchronize(g->stream), "gemver read");
    });
}

int cs_gemver_time(cs_gemver g, int iters, float* ms_per_run) {
    return wl::guard([&] {
        if (!g || !ms_per_run || iters <= 0) throw std::invalid_argument("cs_gemver_time: bad argument");
        if (!g->ready) initialize(*g);
        wl::EventPair ev;
        ev.record_start(g->stream);
        for (int i = 0; i < iters; ++i) forward(*g, nullptr);
        ev.record_stop(g->stream);
        *ms_per_run = ev.elapsed_ms() / static_cast<float>(iters);
    });
}

int cs_gemver_time_launches(cs_gemver g, int iters, float* ms, int cap, int* count) {
    return wl::guard([&] {
        if (!g || !ms || !count || iters <= 0) throw std::invalid_argument("cs_gemver_time_launches: bad argument");
        if (!g->ready) initialize(*g);
        std::vector<double> sum;
        for (int i = 0; i < iters; ++i) {
            std::vector<wl::LaunchMark> marks;
            forward(*g, &marks);
            const std::vector<float> d = wl::durations(marks);
            if (sum.empty()) sum.assign(d.size(), 0.0);
            for (std::size_t k = 0; k < d.size() && k < sum.size(); ++k) sum[k] += d[k];
        }
        *count = static_cast<int>(sum.size());
        for (int k = 0; k < cap && k < *count; ++k) ms[k] = static_cast<float>(sum[k] / iters);
    });
}

int cs_gemver_destroy(cs_gemver g) {
    return wl::guard([&] {
        if (!g) return;
        cs_hip_stream_synchronize(g->stream);
        cs_hip_stream_destroy(g->stream);
        delete g;
    });
}

}  // extern "C"
