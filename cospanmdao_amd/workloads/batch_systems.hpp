#pragma once

// =============================================================================
// cospanmdao_amd/workloads/batch_systems.hpp — the batched ragged GEMVER as a System
// =============================================================================
//
// Leaves: plan (BatchPlan) -> scan (ColumnScan) -> map (UnitMap), then gemver (RaggedGemver, the persistent form: one
// work-queue launch) or, in the phased form, p1 (RaggedPass1) -> red (RaggedReduce) -> p2 (RaggedPass2). Ports (both forms): inputs sizes, params
// {R, RT, G}, A (all instances back to back, instance k at 16-byte unit offs[K + k] - offs[K]), u1 v1 u2 v2 y z (back to
// back, instance k at offs[k]), rank (each instance's place in the work order), alpha, beta; outputs counts (9K), offs
// (the device scan of the counts, 9K + 1 entries), result ([total, flag] of the scan), x, w.
//
// Inputs: instance k's A(i, j) = hash_unit(seed_A, a_k + i N_k + j) with a_k = sum_{m<k} N_m^2 and its vectors
// u1_k[i] = hash_unit(seed_u1, v_k + i) with v_k = sum_{m<k} N_m: one flat hash stream per array (seeds as in
// examples/gemver-chain/common/chain_data.hpp:32), generated on the device.

#include <cstdint>
#include <memory>
#include <algorithm>
#include <numeric>
#include <vector>

#include "naga/cospan/cospan.hpp"
#include "batch_components.hpp"

namespace batch {

/// Host-side layout of the instances (what the allocations are sized by; the kernels use the device scan's offsets) and
/// the work order: rank[k] = instance k's place when the instances are sorted by size, largest first (ties in index
/// order), so that each phase's queue ends on the smallest units.
struct Layout {
    std::vector<std::uint32_t> sizes, rank;
    std::uint64_t              vec_total = 0, a4_total = 0, p_max = 0;
    std::uint64_t              units_max[3] = {0, 0, 0};  ///< the phases' unit counts at the smallest R and RT (4)
    explicit Layout(std::vector<std::uint32_t> s, std::uint32_t min_R) : sizes{std::move(s)}, rank(sizes.size()) {
        for (std::uint32_t n : sizes) {
            vec_total += n;
            a4_total += static_cast<std::uint64_t>(n) * ld_of(n) / 4;
            p_max += static_cast<std::uint64_t>(ceil_div(n, min_R)) * n;
        }
        count_units(sizes.data(), sizes.size(), min_R, 4, units_max);
        std::vector<std::uint32_t> order(sizes.size());
        std::iota(order.begin(), order.end(), 0u);
        std::stable_sort(order.begin(), order.end(), [this](std::uint32_t a, std::uint32_t b) { return sizes[a] > sizes[b]; });
        for (std::size_t r = 0; r < order.size(); ++r) rank[order[r]] = static_cast<std::uint32_t>(r);
    }
};

/// The A initializer is a kernel of the library that builds the model (it needs the per-instance layout).
using AInit = std::function<cs::InitSpec(const Layout&)>;

inline cs::InitSpec hash_stream(std::uint32_t seed, std::size_t n) {
    return cs::init::custom<DType::f32, 1>(
        [seed, n](cs_hip_stream s, cs::Initializer<DType::f32, 1> v) {
            cs::detail::hip_check(cs_hip_fill_hash_f32(seed, cs::raw::data(v), 1, n, n, s), "hash initializer");
        },
        cs::Shape{n});
}

class BatchedGemver : public cs::System {
public:
    static constexpr std::uint32_t kMinR = 32;  // the smallest strip the autotuner tries: P is sized for it

    BatchedGemver(std::shared_ptr<Layout> layout, std::shared_ptr<Knobs> knobs, std::vector<std::uint32_t> params, float alpha, float beta,
                  std::uint32_t base_seed, AInit a_init, bool phased = false)
        : layout_{std::move(layout)}, knobs_{std::move(knobs)}, params_{std::move(params)}, alpha_{alpha}, beta_{beta}, seed_{base_seed},
          a_init_{std::move(a_init)}, phased_{phased} {}

    void setup() override {
        add_subsystem(cs::subsystem_desc<BatchPlan>("plan"));
        add_subsystem(cs::subsystem_desc<ColumnScan>("scan"));
        connect("plan.counts", "scan.counts");
        add_subsystem(cs::subsystem_desc<UnitMap>("map"));
        connect("scan.offs", "map.offs");
        tie_inputs<DType::u32, 1>({"plan.rank", "map.rank"}, "rank");
        shape_hook("map.map1", cs::dim::fixed_at_run(kMapWords * layout_->units_max[0]));
        shape_hook("map.mapR", cs::dim::fixed_at_run(kMapWords * layout_->units_max[1]));
        shape_hook("map.map2", cs::dim::fixed_at_run(kMapWords * layout_->units_max[2]));
        if (phased_) {
            add_subsystem(cs::subsystem_desc<RaggedPass1>("p1", knobs_));
            add_subsystem(cs::subsystem_desc<RaggedReduce>("red", knobs_));
            add_subsystem(cs::subsystem_desc<RaggedPass2>("p2", knobs_));
            for (const char* leaf : {"p1.offs", "red.offs", "p2.offs"}) connect("scan.offs", leaf);
            connect("map.map1", "p1.map");
            connect("map.mapR", "red.map");
            connect("map.map2", "p2.map");
            connect("p1.P", "red.P");
            connect("red.x", "p2.x");
            tie_inputs<DType::u32, 1>({"plan.sizes", "map.sizes"}, "sizes");
            tie_inputs<DType::u32, 1>({"plan.params", "p1.params", "red.params", "p2.params"}, "params", cs::shape(3));
            for (const char* v : {"A", "u1", "v1", "u2", "v2"})
                tie_inputs<DType::f32, 1>({std::string{"p1."} + v, std::string{"p2."} + v}, v);
            promote<DType::f32, 1>("p1.y", "y");
            promote<DType::f32, 1>("red.z", "z");
            promote<DType::f32, 1>("p2.alpha", "alpha");
            promote<DType::f32, 1>("red.beta", "beta");
            promote<DType::f32, 1>("red.x", "x");
            promote<DType::f32, 1>("p2.w", "w");
            shape_hook("p1.P", cs::dim::fixed_at_run(layout_->p_max));
        } else {
            add_subsystem(cs::subsystem_desc<RaggedGemver>("gemver", knobs_));
            connect("scan.offs", "gemver.offs");
            connect("map.map1", "gemver.map1");
            connect("map.mapR", "gemver.mapR");
            connect("map.map2", "gemver.map2");
            tie_inputs<DType::u32, 1>({"plan.sizes", "map.sizes", "gemver.sizes"}, "sizes");
            tie_inputs<DType::u32, 1>({"plan.params", "gemver.params"}, "params", cs::shape(3));
            for (const char* v : {"A", "u1", "v1", "u2", "v2", "y", "z", "alpha", "beta"})
                promote<DType::f32, 1>(std::string{"gemver."} + v, v);
            promote<DType::f32, 1>("gemver.x", "x");
            promote<DType::f32, 1>("gemver.w", "w");
            shape_hook("gemver.P", cs::dim::fixed_at_run(layout_->p_max));
        }
        promote<DType::u32, 1>("scan.offs", "offs");
        promote<DType::u32, 1>("scan.result", "result");
        promote<DType::u32, 1>("plan.counts", "counts");

        const std::uint32_t sb = seed_ * 16u;
        init_input("sizes", cs::init::from_host(layout_->sizes));
        init_input("params", cs::init::from_host(params_));
        init_input("rank", cs::init::from_host(layout_->rank));
        init_input("A", a_init_(*layout_));
        const char* vecs[6] = {"u1", "v1", "u2", "v2", "y", "z"};
        for (std::uint32_t i = 0; i < 6; ++i) init_input(vecs[i], hash_stream(sb + 2 + i, layout_->vec_total));
        init_input("alpha", cs::init::fill(alpha_, {1}));
        init_input("beta", cs::init::fill(beta_, {1}));
    }

private:
    std::shared_ptr<Layout>    layout_;
    std::shared_ptr<Knobs>     knobs_;
    std::vector<std::uint32_t> params_;
    float                      alpha_, beta_;
    std::uint32_t              seed_;
    AInit                      a_init_;
    bool                       phased_;
};

}  // namespace batch
