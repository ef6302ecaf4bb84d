// tests/framework/ragged_gemver.cpp — a batch of GEMVERs of different sizes written only with the public ragged API
// (a test program for the GPU).
//
// Instance k has size N_k: Â = A + u1 v1^T + u2 v2^T, x = beta Â^T y + z, w = alpha Â x. Every per-instance array is a
// ragged port: A (N_k^2 elements), the six vectors as one block of 6 N_k ([u1 | v1 | u2 | v2 | y | z]), x and w (N_k). Their
// counts come from shape hooks: A from counts_from(N_k^2) and the vectors from counts_from(6 N_k), both computed on the
// device from the sizes by a component, and x, w from counts_from(sizes) directly; so nothing about the sizes is known on
// the host until the hooks have run. The two passes run as bare components and as one fused group with a global barrier.
// Checked per instance against a host FP64 GEMVER over forwards whose sizes (and number of instances) change. Prints one
// line per case and "OK" at the end when every check holds; exits 1 otherwise.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <functional>
#include <string>
#include <vector>

#include <naga/cospan/cospan.hpp>

namespace cs = naga::cospan;
using cs::DType;

static int g_fail = 0;
#define CHECK(cond, ...)                               \
    do {                                               \
        if (!(cond)) {                                 \
            ++g_fail;                                  \
            std::printf("  FAIL line %d: ", __LINE__); \
            std::printf(__VA_ARGS__);                  \
            std::printf("\n");                         \
        }                                              \
    } while (0)

namespace trg {

constexpr std::size_t kWg = 256;
constexpr float       kAlpha = 1.5f, kBeta = 1.2f;

/// A value in [-1, 1) from (stream, index): an integer hash (so host and device agree exactly).
__host__ __device__ inline float val(std::uint32_t stream, std::uint64_t i) {
    std::uint32_t h = static_cast<std::uint32_t>(i * 2654435761ull) ^ (stream * 0x9E3779B9u);
    h ^= h >> 16;
    h *= 0x7feb352du;
    h ^= h >> 15;
    h *= 0x846ca68bu;
    h ^= h >> 16;
    return static_cast<float>(h >> 8) * (2.0f / 16777216.0f) - 1.0f;
}

CS_COMPONENTS_BEGIN

/// sq[k] = N_k^2, six[k] = 6 N_k.
class Counts : public cs::Component<Counts> {
public:
    struct Inputs {
        cs::Input<DType::u32, 1> sizes;
    };
    struct Outputs {
        cs::Output<DType::u32, 1> sq, six;
    };
    static void compute_forward(cs::id<1> k, const Inputs& in, const Outputs& out) {
        const std::uint32_t n = in.sizes[k];
        out.sq[k]             = n * n;
        out.six[k]            = 6 * n;
    }
    Inputs  inputs;
    Outputs outputs;
    void    declare() override {
        declare_input(inputs.sizes, "sizes", cs::unspecified);
        declare_output(outputs.sq, "sq", cs::unspecified);
        declare_output(outputs.six, "six", cs::unspecified);
    }
    void prepare_forward() override {
        set_shape(outputs.sq, shape_of(inputs.sizes)[0]);
        set_shape(outputs.six, shape_of(inputs.sizes)[0]);
    }
    std::unique_ptr<cs::LaunchDescription> define_forward() override { return create_launch_config(cs::range_desc(outputs.sq.shape())); }
};

/// The instances' inputs: A block k from stream 1, the vector block k from stream 2, indexed within the block and by k.
class Generate : public cs::Component<Generate> {
public:
    struct Inputs {};
    struct Outputs {
        cs::RaggedOutput<DType::f32, 2> A, V;
    };
    static void compute_forward(cs::GroupItem<1> it, const Inputs&, const Outputs& out) {
        const std::size_t k = it.get_group(0), t = it.get_local_id(0);
        const auto        a = out.A.block(k);
        const auto        v = out.V.block(k);
        for (std::size_t e = t; e < a.size(); e += kWg) a[e] = val(1u + 16u * static_cast<std::uint32_t>(k), e);
        for (std::size_t e = t; e < v.size(); e += kWg) v[e] = val(2u + 16u * static_cast<std::uint32_t>(k), e);
    }
    Inputs  inputs;
    Outputs outputs;
    static constexpr int max_work_group_size = kWg;
    void                 declare() override {
        declare_output(outputs.A, "A");
        declare_output(outputs.V, "V");
    }
    std::unique_ptr<cs::LaunchDescription> define_forward() override {
        return create_launch_config(cs::range_desc(cs::range<1>{outputs.A.size() * kWg}), cs::wg_desc(cs::range<1>{kWg}));
    }
};

/// One work-group per instance: x_j = beta sum_i Â_ij y_i + z_j.
class PassX : public cs::Component<PassX> {
public:
    struct Inputs {
        cs::RaggedInput<DType::f32, 2> A, V;
    };
    struct Outputs {
        cs::RaggedOutput<DType::f32, 2> x;
    };
    static void compute_forward(cs::GroupItem<1> it, const Inputs& in, const Outputs& out) {
        const std::size_t k = it.get_group(0), t = it.get_local_id(0);
        const auto        a = in.A.block(k);
        const auto        v = in.V.block(k);
        const auto        x = out.x.block(k);
        const std::size_t n = x.size();
        const float *     u1 = &v[0], *v1 = u1 + n, *u2 = v1 + n, *v2 = u2 + n, *y = v2 + n, *z = y + n;
        for (std::size_t j = t; j < n; j += kWg) {
            float s = 0.0f;
            for (std::size_t i = 0; i < n; ++i) s += (a[i * n + j] + u1[i] * v1[j] + u2[i] * v2[j]) * y[i];
            x[j] = kBeta * s + z[j];
        }
    }
    Inputs  inputs;
    Outputs outputs;
    static constexpr int max_work_group_size = kWg;
    void                 declare() override {
        declare_input(inputs.A, "A");
        declare_input(inputs.V, "V");
        declare_output(outputs.x, "x");
    }
    std::unique_ptr<cs::LaunchDescription> define_forward() override {
        return create_launch_config(cs::range_desc(cs::range<1>{outputs.x.size() * kWg}), cs::wg_desc(cs::range<1>{kWg}));
    }
};

/// One work-group per instance: w_i = alpha sum_j Â_ij x_j.
class PassW : public cs::Component<PassW> {
public:
    struct Inputs {
        cs::RaggedInput<DType::f32, 2> A, V, x;
    };
    struct Outputs {
        cs::RaggedOutput<DType::f32, 2> w;
    };
    static void compute_forward(cs::GroupItem<1> it, const Inputs& in, const Outputs& out) {
        const std::size_t k = it.get_group(0), t = it.get_local_id(0);
        const auto        a = in.A.block(k);
        const auto        v = in.V.block(k);
        const auto        x = in.x.block(k);
        const auto        w = out.w.block(k);
        const std::size_t n = w.size();
        const float *     u1 = &v[0], *v1 = u1 + n, *u2 = v1 + n, *v2 = u2 + n;
        for (std::size_t i = t; i < n; i += kWg) {
            float s = 0.0f;
            for (std::size_t j = 0; j < n; ++j) s += (a[i * n + j] + u1[i] * v1[j] + u2[i] * v2[j]) * x[j];
            w[i] = kAlpha * s;
        }
    }
    Inputs  inputs;
    Outputs outputs;
    static constexpr int max_work_group_size = kWg;
    void                 declare() override {
        declare_input(inputs.A, "A");
        declare_input(inputs.V, "V");
        declare_input(inputs.x, "x");
        declare_output(outputs.w, "w");
    }
    std::unique_ptr<cs::LaunchDescription> define_forward() override {
        return create_launch_config(cs::range_desc(cs::range<1>{outputs.w.size() * kWg}), cs::wg_desc(cs::range<1>{kWg}));
    }
};

CS_COMPONENTS_END

/// counts -> gen -> (x pass, w pass) as bare components or one fused group "g" with a global barrier.
class Batch : public cs::System {
public:
    explicit Batch(bool fused) : fused_{fused} {}
    void setup() override {
        add_subsystem(cs::subsystem_desc<Counts>("counts"));
        add_subsystem(cs::subsystem_desc<Generate>("gen"));
        const std::string p = fused_ ? "g." : "";
        if (fused_) {
            add_subsystem(cs::fuse_components("g", cs::connects_desc({{"px.x", "pw.x"}}), cs::segment(cs::subsystem_desc<PassX>("px")), cs::global_barrier(),
                                              cs::segment(cs::subsystem_desc<PassW>("pw"))));
        } else {
            add_subsystem(cs::subsystem_desc<PassX>("px"));
            add_subsystem(cs::subsystem_desc<PassW>("pw"));
            connect("px.x", "pw.x");
        }
        connect("gen.A", p + "px.A");
        connect("gen.V", p + "px.V");
        connect("gen.A", p + "pw.A");
        connect("gen.V", p + "pw.V");
        promote<DType::u32, 1>("counts.sizes", "sizes");
        promote<DType::f32, 2>(p + "px.x", "x");
        promote<DType::f32, 2>(p + "pw.w", "w");
        shape_hook("gen.A", cs::ragged::counts_from("counts.sq"));
        shape_hook("gen.V", cs::ragged::counts_from("counts.six"));
        shape_hook("x", cs::ragged::counts_from("sizes"));
        shape_hook("w", cs::ragged::counts_from("sizes"));
    }

private:
    bool fused_;
};

/// x and w of one instance in FP64 from the same inputs.
void reference(std::size_t k, std::size_t n, std::vector<double>& x, std::vector<double>& w) {
    auto A = [&](std::size_t i, std::size_t j) { return static_cast<double>(val(1u + 16u * static_cast<std::uint32_t>(k), i * n + j)); };
    auto V = [&](int which, std::size_t i) { return static_cast<double>(val(2u + 16u * static_cast<std::uint32_t>(k), which * n + i)); };
    auto Ah = [&](std::size_t i, std::size_t j) { return A(i, j) + V(0, i) * V(1, j) + V(2, i) * V(3, j); };
    x.assign(n, 0.0);
    w.assign(n, 0.0);
    for (std::size_t j = 0; j < n; ++j) {
        double s = 0.0;
        for (std::size_t i = 0; i < n; ++i) s += Ah(i, j) * V(4, i);
        x[j] = static_cast<double>(kBeta) * s + V(5, j);
    }
    for (std::size_t i = 0; i < n; ++i) {
        double s = 0.0;
        for (std::size_t j = 0; j < n; ++j) s += Ah(i, j) * x[j];
        w[i] = static_cast<double>(kAlpha) * s;
    }
}

double rel_l2(std::span<const float> got, const std::vector<double>& want) {
    double num = 0.0, den = 0.0;
    for (std::size_t i = 0; i < want.size(); ++i) {
        const double d = static_cast<double>(got[i]) - want[i];
        num += d * d;
        den += want[i] * want[i];
    }
    return den > 0.0 ? std::sqrt(num / den) : std::sqrt(num);
}

}  // namespace trg

using namespace trg;

#ifndef CS_SOURCE_HASH
#define CS_SOURCE_HASH "unknown"
#endif

int main() {
    std::printf("build %s\n", CS_SOURCE_HASH);
    const std::vector<std::vector<std::uint32_t>> batches{{3, 64, 200, 17, 1, 256, 90}, {255, 8, 0, 129}, {33}, {120, 120, 5, 300, 64, 2, 2, 77}};
    for (bool fused : {false, true}) {
        std::printf("a batch of GEMVERs through ragged ports and counts_from hooks%s\n", fused ? ", the two passes fused" : "");
        cs::Model m = cs::create_model<Batch>(fused);
        m.setup();
        for (const auto& sizes : batches) {
            m.update_input("sizes", cs::init::from_host(sizes));
            m.run();
            m.wait();
            auto x = m.view_ragged_output<DType::f32, 2>("x");
            auto w = m.view_ragged_output<DType::f32, 2>("w");
            x.sync();
            w.sync();
            CHECK(x.n() == sizes.size() && w.n() == sizes.size(), "%zu instances", sizes.size());
            double worst = 0.0;
            for (std::size_t k = 0; k < sizes.size() && k < x.n(); ++k) {
                const std::size_t n = sizes[k];
                CHECK(x.block(k).size() == n && w.block(k).size() == n, "instance %zu: blocks of %zu", k, n);
                if (x.block(k).size() != n || w.block(k).size() != n) continue;
                std::vector<double> xr, wr;
                reference(k, n, xr, wr);
                const double ex = rel_l2(x.block(k), xr), ew = rel_l2(w.block(k), wr);
                worst           = std::max(worst, std::max(ex, ew));
                CHECK(ex <= 1e-5 && ew <= 1e-5, "instance %zu (n %zu): x %.2e, w %.2e", k, n, ex, ew);
            }
            std::printf("  %zu instances, worst rel-L2 %.2e\n", sizes.size(), worst);
        }
    }
    if (g_fail) {
        std::printf("%d checks failed\n", g_fail);
        return 1;
    }
    std::printf("OK\n");
    return 0;
}
