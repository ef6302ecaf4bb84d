// tests/framework/ragged.cpp — ragged ports through cs::Model (a test program).
//
// Built twice: for the GPU (make -C tests/framework), where every check runs, and for the CPU against the mock boundary
// (make -C tests/host, -DCS_HOST_MOCK), where the components' kernels do not run and only the host logic is checked:
// shapes, block offsets, the column builds, hook placement, wiring and error messages.
//
// The cases follow the reference's tests/test_ragged.cpp:516-710 (fill / scale / sum over forwards with changing counts,
// N = 0, counts_from column 0 of a rank-2 source, a hooked ragged OUTPUT whose counts grow and shrink, declared counts
// merged at connect and their mismatch errors, a RaggedInput reading a dense port, dense views refused, negative counts
// and totals past 32 bits) with this file's own components, and go further: offsets_from and its errors, a ragged port
// inside a fused group (the reference refuses those), a hook whose source is written after its target's first holder,
// and graph replay of a model with declared counts. Prints one line per case and "OK" at the end when every check
// holds; exits 1 otherwise.
#include <cstdint>
#include <cstdio>
#include <functional>
#include <string>
#include <vector>

#include <naga/cospan/cospan.hpp>

#ifdef CS_HOST_MOCK
#include "mock_hip.hpp"
constexpr bool kData = false;  // the kernels are not emulated: their outputs are not looked at
#else
constexpr bool kData = true;
#endif

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

using Sizes = std::vector<std::size_t>;

/// Runs f; it must throw a message holding every needle.
void expect_throws(const std::function<void()>& f, const std::vector<std::string>& needles, const char* what) {
    std::string msg = "<no exception>";
    try {
        f();
    } catch (const std::exception& e) {
        msg = e.what();
    }
    bool all = true;
    for (const std::string& n : needles) all = all && msg.find(n) != std::string::npos;
    CHECK(all, "%s: message: %s", what, msg.c_str());
}

CS_COMPONENTS_BEGIN

/// counts[i] = raw[i] (N fixed by the declaration of raw).
class Counts : public cs::Component<Counts> {
public:
    struct Inputs {
        cs::Input<DType::u32, 1> raw;
    };
    struct Outputs {
        cs::Output<DType::u32, 1> counts;
    };
    static void compute_forward(cs::id<1> i, const Inputs& in, const Outputs& out) { out.counts[i] = in.raw[i]; }
    Inputs  inputs;
    Outputs outputs;
    explicit Counts(std::size_t n) : n_{n} {}
    void declare() override {
        declare_input(inputs.raw, "raw", n_);
        declare_output(outputs.counts, "counts", cs::unspecified);
    }
    void prepare_forward() override { set_shape(outputs.counts, shape_of(inputs.raw)[0]); }
    std::unique_ptr<cs::LaunchDescription> define_forward() override { return create_launch_config(cs::range_desc(outputs.counts.shape())); }

private:
    std::size_t n_;
};

/// pts block i, element j = i * 1000 + j; the shape comes from the (hooked) ragged input `layout`.
class Fill : public cs::Component<Fill> {
public:
    struct Inputs {
        cs::RaggedInput<DType::f32, 2> layout;
    };
    struct Outputs {
        cs::RaggedOutput<DType::f32, 2> pts;
    };
    static void compute_forward(cs::id<1> id, const Inputs&, const Outputs& out) {
        const std::size_t i = id[0];
        const auto        b = out.pts.block(i);
        for (std::size_t j = 0; j < b.size(); ++j) b[j] = static_cast<float>(i * 1000 + j);
    }
    Inputs  inputs;
    Outputs outputs;
    void    declare() override {
        declare_input(inputs.layout, "layout");
        declare_output(outputs.pts, "pts");
    }
    void prepare_forward() override { set_shape(outputs.pts, shape_of(inputs.layout)); }
    std::unique_ptr<cs::LaunchDescription> define_forward() override {
        return create_launch_config(cs::range_desc(cs::range<1>{inputs.layout.size()}));
    }
};

/// out = 2 * in, block by block; the output takes the input's ragged shape.
class Scale : public cs::Component<Scale> {
public:
    struct Inputs {
        cs::RaggedInput<DType::f32, 2> in;
    };
    struct Outputs {
        cs::RaggedOutput<DType::f32, 2> out;
    };
    static void compute_forward(cs::id<1> id, const Inputs& in, const Outputs& out) {
        const auto bi = in.in.block(id[0]);
        const auto bo = out.out.block(id[0]);
        for (std::size_t q = 0; q < bi.size(); ++q) bo[q] = 2.0f * bi[q];
    }
    Inputs  inputs;
    Outputs outputs;
    void    declare() override {
        declare_input(inputs.in, "in");
        declare_output(outputs.out, "out");
    }
    void prepare_forward() override { set_shape(outputs.out, shape_of(inputs.in)); }
    std::unique_ptr<cs::LaunchDescription> define_forward() override {
        return create_launch_config(cs::range_desc(cs::range<1>{inputs.in.size()}));
    }
};

/// sums[i] = the sum of block i (in block order). Declared: a ragged input that DECLARES its counts (merged with its
/// source's at connect).
template <bool Declared>
class SumOf : public cs::Component<SumOf<Declared>> {
public:
    struct Inputs {
        cs::RaggedInput<DType::f32, 2> pts;
    };
    struct Outputs {
        cs::Output<DType::f32, 1> sums;
    };
    static void compute_forward(cs::id<1> id, const Inputs& in, const Outputs& out) {
        const auto b = in.pts.block(id[0]);
        float      s = 0.0f;
        for (std::size_t q = 0; q < b.size(); ++q) s += b[q];
        out.sums[id] = s;
    }
    Inputs  inputs;
    Outputs outputs;
    SumOf() = default;
    explicit SumOf(Sizes s) : sizes_{std::move(s)} {}
    void declare() override {
        if constexpr (Declared)
            this->declare_input(inputs.pts, "pts", cs::counts(sizes_));
        else
            this->declare_input(inputs.pts, "pts");
        this->declare_output(outputs.sums, "sums", cs::unspecified);
    }
    void prepare_forward() override { this->set_shape(outputs.sums, this->shape_of(inputs.pts).n); }
    std::unique_ptr<cs::LaunchDescription> define_forward() override {
        return this->create_launch_config(cs::range_desc(outputs.sums.shape()));
    }

private:
    Sizes sizes_;
};
using Sum     = SumOf<false>;
using TakeSum = SumOf<true>;

/// m block i, element j = i * 100 + j, with counts declared by a host array.
class Make : public cs::Component<Make> {
public:
    struct Inputs {};
    struct Outputs {
        cs::RaggedOutput<DType::f32, 2> m;
    };
    static void compute_forward(cs::id<1> id, const Inputs&, const Outputs& out) {
        const std::size_t i = id[0];
        const auto        b = out.m.block(i);
        for (std::size_t j = 0; j < b.size(); ++j) b[j] = static_cast<float>(i * 100 + j);
    }
    Inputs  inputs;
    Outputs outputs;
    explicit Make(Sizes s) : sizes_{std::move(s)} {}
    void declare() override { declare_output(outputs.m, "m", cs::counts(sizes_)); }
    std::unique_ptr<cs::LaunchDescription> define_forward() override { return create_launch_config(cs::range_desc(cs::range<1>{sizes_.size()})); }

private:
    Sizes sizes_;
};

/// A ragged OUTPUT whose N and counts a hook owns (from the values of `cnt`, DT_CNT of rank R_CNT); v block i, element j =
/// i * 100 + j. With Layout, a hooked ragged INPUT `layout` gives the shape and the output copies it.
template <DType DT_CNT, int R_CNT, bool Layout>
class Hooked : public cs::Component<Hooked<DT_CNT, R_CNT, Layout>> {
public:
    struct Inputs {
        cs::Input<DT_CNT, R_CNT>        cnt;
        cs::RaggedInput<DType::f32, 2> layout;
    };
    struct Outputs {
        cs::RaggedOutput<DType::f32, 2> v;
    };
    static void compute_forward(cs::id<1> id, const Inputs&, const Outputs& out) {
        const std::size_t i = id[0];
        const auto        b = out.v.block(i);
        for (std::size_t j = 0; j < b.size(); ++j) b[j] = static_cast<float>(i * 100 + j);
    }
    Inputs  inputs;
    Outputs outputs;
    void    declare() override {
        this->declare_input(inputs.cnt, "cnt", cs::unspecified);
        if constexpr (Layout) this->declare_input(inputs.layout, "layout");
        this->declare_output(outputs.v, "v");
    }
    void prepare_forward() override {
        if constexpr (Layout) this->set_shape(outputs.v, this->shape_of(inputs.layout));
    }
    std::unique_ptr<cs::LaunchDescription> define_forward() override {
        return this->create_launch_config(cs::range_desc(cs::range<1>{outputs.v.size()}));
    }
};

/// d[i, j] = 10 i + j, dense 3 x 4.
class DenseGen : public cs::Component<DenseGen> {
public:
    struct Inputs {};
    struct Outputs {
        cs::Output<DType::f32, 2> d;
    };
    static void compute_forward(cs::id<2> i, const Inputs&, const Outputs& out) { out.d[i] = static_cast<float>(i[0] * 10 + i[1]); }
    Inputs  inputs;
    Outputs outputs;
    void    declare() override { declare_output(outputs.d, "d", cs::shape(3, 4)); }
    std::unique_ptr<cs::LaunchDescription> define_forward() override { return create_launch_config(cs::range_desc(outputs.d.shape())); }
};

/// A dense rank-2 input with open dimensions.
class DenseIn : public cs::Component<DenseIn> {
public:
    struct Inputs {
        cs::Input<DType::f32, 2> in;
    };
    struct Outputs {};
    static void compute_forward(cs::id<1>, const Inputs&, const Outputs&) {}
    Inputs  inputs;
    Outputs outputs;
    void    declare() override { declare_input(inputs.in, "in", cs::unspecified); }
    std::unique_ptr<cs::LaunchDescription> define_forward() override { return create_launch_config(cs::range_desc(cs::range<1>{4})); }
};

CS_COMPONENTS_END

using Build = std::function<void(cs::System&)>;
class Top : public cs::System {
    Build build_;

public:
    explicit Top(Build b) : build_{std::move(b)} {}
    void setup() override { build_(*this); }
};
cs::Model make(Build b) { return cs::create_model<Top>(std::move(b)); }

// On the GPU the hook reads the Counts component's output; against the mock (no kernels run) its free input.
const char* const kCountsSource = kData ? "counts.counts" : "counts.raw";

/// counts -> fill -> scale -> sum; fill's layout hooked from the counts. fused: fill and scale in one fused group with a
/// global barrier between them (`f.fill`, `f.scale`).
cs::Model ragged_top(std::size_t n, bool fused) {
    return make([n, fused](cs::System& s) {
        s.add_subsystem(cs::subsystem_desc<Counts>("counts", n));
        const std::string p = fused ? "f." : "";
        if (fused) {
            s.add_subsystem(cs::fuse_components("f", cs::connects_desc({{"fill.pts", "scale.in"}}), cs::segment(cs::subsystem_desc<Fill>("fill")),
                                                cs::global_barrier(), cs::segment(cs::subsystem_desc<Scale>("scale"))));
        } else {
            s.add_subsystem(cs::subsystem_desc<Fill>("fill"));
            s.add_subsystem(cs::subsystem_desc<Scale>("scale"));
            s.connect("fill.pts", "scale.in");
        }
        s.add_subsystem(cs::subsystem_desc<Sum>("sum"));
        s.shape_hook(p + "fill.layout", cs::ragged::counts_from(kCountsSource));
        s.connect(p + "scale.out", "sum.pts");
    });
}

struct Ref {
    std::vector<float>         data;
    std::vector<std::uint64_t> offsets{0};
};
Ref ref_of(const Sizes& counts, std::size_t mul) {
    Ref r;
    for (std::size_t i = 0; i < counts.size(); ++i) {
        for (std::size_t j = 0; j < counts[i]; ++j) r.data.push_back(static_cast<float>(i * mul + j));
        r.offsets.push_back(r.data.size());
    }
    return r;
}
template <class V>
std::vector<float> flat(const V& v) {
    return std::vector<float>(v.span().begin(), v.span().end());
}

/// One forward of a ragged_top model with `counts`, every view checked.
void run_counts(cs::Model& model, const Sizes& counts, bool fused) {
    const std::string p = fused ? "f." : "";
    std::vector<std::uint32_t> raw(counts.begin(), counts.end());
    model.update_input("counts.raw", cs::init::from_host(raw));
    model.run();
    model.wait();
    const Ref ref = ref_of(counts, 1000);
    auto      sums = model.view_output<DType::f32, 1>("sum.sums");
    sums.sync();
    CHECK(sums.size() == counts.size(), "sums: %zu of %zu", sums.size(), counts.size());
    if (kData && sums.size() == counts.size())
        for (std::size_t i = 0; i < counts.size(); ++i) {
            float want = 0.0f;
            for (std::uint64_t q = ref.offsets[i]; q < ref.offsets[i + 1]; ++q) want += 2.0f * ref.data[q];
            CHECK(sums[i] == want, "sum of block %zu (count %zu): %g, want %g", i, counts[i], sums[i], want);
        }
    auto pts = model.view_ragged_output<DType::f32, 2>(p + "fill.pts");
    pts.sync();
    CHECK(pts.n() == counts.size(), "pts.n() %zu", pts.n());
    CHECK(pts.block_offsets() == ref.offsets, "pts offsets");
    CHECK(pts.size() == ref.data.size(), "pts size %zu, want %zu", pts.size(), ref.data.size());
    if (kData) CHECK(flat(pts) == ref.data, "pts data");
    for (std::size_t i = 0; i < counts.size() && i + 1 < pts.block_offsets().size(); ++i)
        CHECK(pts.block(i).size() == counts[i], "block %zu size %zu", i, pts.block(i).size());
    auto scaled = model.view_ragged_output<DType::f32, 2>(p + "scale.out");
    scaled.sync();
    CHECK(scaled.block_offsets() == ref.offsets, "scale.out offsets");
    CHECK(scaled.size() == ref.data.size(), "scale.out size");
    if (kData && scaled.size() == ref.data.size())
        for (std::size_t q = 0; q < ref.data.size(); ++q) CHECK(scaled[q] == 2.0f * ref.data[q], "scale.out[%zu]", q);
    auto peer = model.view_ragged_input<DType::f32, 2>(p + "scale.in");
    peer.sync();
    CHECK(peer.block_offsets() == ref.offsets, "scale.in offsets");
    if (kData) CHECK(flat(peer) == ref.data, "scale.in data");
}

}  // namespace trg

using namespace trg;

#ifndef CS_SOURCE_HASH
#define CS_SOURCE_HASH "unknown"
#endif

int main() {
    std::printf("build %s\n", CS_SOURCE_HASH);
#ifdef CS_HOST_MOCK
    mock::set_fallback_kernel([](const void*, const std::uint32_t*, const std::uint32_t*) {});
#endif
    for (bool fused : {false, true}) {
        std::printf("fill / scale / sum over forwards with changing counts%s\n", fused ? ", fill and scale in a fused group" : "");
        cs::Model model = ragged_top(6, fused);
        model.setup();
        run_counts(model, {3, 0, 2, 4, 0, 1}, fused);  // empty blocks in the middle and at the end
        run_counts(model, {0, 5, 1, 0, 2, 2}, fused);  // an empty first block
        run_counts(model, {1, 0, 0, 0, 0, 0}, fused);  // one block
        run_counts(model, {0, 0, 0, 0, 0, 0}, fused);  // every block empty
        run_counts(model, {2, 2, 2, 2, 2, 7}, fused);  // back up
        run_counts(model, {2, 2, 2, 2, 2, 7}, fused);  // the same again (no graph: a hooked model runs eagerly)
        CHECK(model.stats().captures == 0, "a model with a ragged hook captured a graph");
        expect_throws([&] { (void)model.view_output<DType::f32, 2>(fused ? "f.fill.pts" : "fill.pts"); }, {"fill.pts", "view_ragged_output"},
                      "dense view of a ragged output");
        expect_throws([&] { (void)model.view_input<DType::f32, 2>("sum.pts"); }, {"sum.pts", "view_ragged_input"}, "dense view of a ragged input");
        expect_throws([&] { (void)model.view_ragged_output<DType::u32, 2>(fused ? "f.fill.pts" : "fill.pts"); }, {"wrong dtype or rank"},
                      "ragged view of the wrong dtype");
    }
    std::printf("N = 0 stays ragged-capable\n");
    {
        cs::Model model = ragged_top(0, false);
        model.setup();
        run_counts(model, {}, false);
        auto pts = model.view_ragged_output<DType::f32, 2>("fill.pts");
        pts.sync();
        CHECK(pts.n() == 0 && pts.size() == 0 && pts.block_offsets() == std::vector<std::uint64_t>{0}, "N = 0 view");
    }
    std::printf("a hooked ragged output over growing and shrinking counts\n");
    {
        cs::Model model = make([](cs::System& s) {
            s.add_subsystem(cs::subsystem_desc<Hooked<DType::u32, 1, false>>("c"));
            s.shape_hook("c.v", cs::ragged::counts_from("c.cnt"));
        });
        model.setup();
        for (const Sizes& counts : {Sizes{1, 2}, Sizes{3, 1, 0, 2}, Sizes{1, 1, 1, 1, 1, 1, 1, 1, 1}, Sizes{2, 0}, Sizes(12, 5), Sizes(300, 7)}) {
            model.update_input("c.cnt", cs::init::from_host(std::vector<std::uint32_t>(counts.begin(), counts.end())));
            model.run();
            model.wait();
            const Ref ref = ref_of(counts, 100);
            auto      v   = model.view_ragged_output<DType::f32, 2>("c.v");
            v.sync();
            CHECK(v.n() == counts.size() && v.block_offsets() == ref.offsets, "hooked output, n = %zu", counts.size());
            if (kData) CHECK(flat(v) == ref.data, "hooked output data, n = %zu", counts.size());
        }
    }
    std::printf("counts_from a rank-2 [N, k] source reads column 0\n");
    {
        cs::Model model = make([](cs::System& s) {
            s.add_subsystem(cs::subsystem_desc<Hooked<DType::u32, 2, true>>("c"));
            s.shape_hook("c.layout", cs::ragged::counts_from("c.cnt"));
        });
        model.setup();
        model.update_input("c.cnt", cs::init::from_host(std::vector<std::uint32_t>{2, 9, 0, 9, 3, 9, 1, 9}, cs::Shape{4, 2}));
        model.run();
        model.wait();
        auto v = model.view_ragged_output<DType::f32, 2>("c.v");
        v.sync();
        CHECK(v.n() == 4 && v.block_offsets() == (std::vector<std::uint64_t>{0, 2, 2, 5, 6}), "column 0 offsets");
        if (kData) CHECK(flat(v) == (std::vector<float>{0, 1, 200, 201, 202, 300}), "column 0 data");
    }
    std::printf("offsets_from: N + 1 offsets of any integer dtype\n");
    {
        cs::Model model = make([](cs::System& s) {
            s.add_subsystem(cs::subsystem_desc<Hooked<DType::i64, 1, false>>("c"));
            s.shape_hook("c.v", cs::ragged::offsets_from("c.cnt"));
        });
        model.setup();
        model.update_input("c.cnt", cs::init::from_host(std::vector<std::int64_t>{0, 2, 2, 5, 6}));
        model.run();
        model.wait();
        auto v = model.view_ragged_output<DType::f32, 2>("c.v");
        v.sync();
        CHECK(v.n() == 4 && v.block_offsets() == (std::vector<std::uint64_t>{0, 2, 2, 5, 6}), "offsets_from offsets");
        if (kData) CHECK(flat(v) == (std::vector<float>{0, 1, 200, 201, 202, 300}), "offsets_from data");
        model.update_input("c.cnt", cs::init::from_host(std::vector<std::int64_t>{0}));  // N = 0
        model.run();
        model.wait();
        auto z = model.view_ragged_output<DType::f32, 2>("c.v");
        z.sync();
        CHECK(z.n() == 0 && z.size() == 0, "offsets_from N = 0");
        for (const auto& bad : {std::vector<std::int64_t>{0, 3, 2}, std::vector<std::int64_t>{1, 2}, std::vector<std::int64_t>{0, -1}}) {
            model.update_input("c.cnt", cs::init::from_host(bad));
            expect_throws([&] { model.run(); }, {"shape hook on 'c.v'", "offsets that are negative, do not start at 0, or decrease"},
                          "bad offsets");
            model.wait();
        }
        model.update_input("c.cnt", cs::init::from_host(std::vector<std::int64_t>{}));
        expect_throws([&] { model.run(); }, {"shape hook on 'c.v'", "holds no offsets"}, "no offsets");
        model.wait();
    }
    std::printf("declared counts: a ragged output feeding a ragged input, graph replays\n");
    {
        const Sizes sizes{1, 0, 3, 2};
        cs::Model   model = make([&](cs::System& s) {
            s.add_subsystem(cs::subsystem_desc<Make>("make", sizes));
            s.add_subsystem(cs::subsystem_desc<TakeSum>("take", sizes));
            s.connect("make.m", "take.pts");
        });
        model.setup();
        for (int fwd = 0; fwd < 4; ++fwd) {
            model.run();
            model.wait();
            auto sums = model.view_output<DType::f32, 1>("take.sums");
            sums.sync();
            CHECK(sums.size() == 4, "take.sums size");
            if (kData && sums.size() == 4)
                for (std::size_t i = 0; i < 4; ++i) {
                    float want = 0.0f;
                    for (std::size_t j = 0; j < sizes[i]; ++j) want += static_cast<float>(i * 100 + j);
                    CHECK(sums[i] == want, "take.sums[%zu] forward %d", i, fwd);
                }
            auto m = model.view_ragged_output<DType::f32, 2>("make.m");
            m.sync();
            CHECK(m.block_offsets() == (std::vector<std::uint64_t>{0, 1, 1, 4, 6}), "make.m offsets");
            if (kData) CHECK(flat(m) == (std::vector<float>{0, 200, 201, 202, 300, 301}), "make.m data forward %d", fwd);
        }
        CHECK(model.stats().captures == 1 && model.stats().replays >= 2, "declared counts: captures %llu replays %llu",
              static_cast<unsigned long long>(model.stats().captures), static_cast<unsigned long long>(model.stats().replays));
        cs::Model zero = make([](cs::System& s) {
            s.add_subsystem(cs::subsystem_desc<Make>("make", Sizes{}));
            s.add_subsystem(cs::subsystem_desc<TakeSum>("take", Sizes{}));
            s.connect("make.m", "take.pts");
        });
        zero.setup();
        zero.run();
        zero.wait();
        auto z = zero.view_ragged_output<DType::f32, 2>("make.m");
        z.sync();
        CHECK(z.n() == 0 && z.block_offsets() == std::vector<std::uint64_t>{0}, "declared N = 0");
    }
    std::printf("mismatching declared counts are errors at connect\n");
    {
        auto per_index = [](Sizes a, Sizes b) {
            return make([a, b](cs::System& s) {
                s.add_subsystem(cs::subsystem_desc<Make>("make", a));
                s.add_subsystem(cs::subsystem_desc<TakeSum>("take", b));
                s.connect("make.m", "take.pts");
            });
        };
        cs::Model m1 = per_index({1, 0, 3, 2}, {2, 1, 2, 1});
        expect_throws([&] { m1.setup(); }, {"connect('make.m', 'take.pts')", "dimension 1", "the counts must match", "[1,0,3,2]", "[2,1,2,1]"},
                      "same total, other counts");
        cs::Model m2 = per_index({1, 0, 3, 2}, {1, 1, 1, 1});
        expect_throws([&] { m2.setup(); }, {"connect('make.m', 'take.pts')", "the counts must match", "[1,1,1,1]", "[1,0,3,2]"}, "other total");
        cs::Model m3 = per_index({1, 2, 3, 4}, {1, 2, 3});
        expect_throws([&] { m3.setup(); }, {"connect('make.m', 'take.pts')", "dimension 0", "declares 4", "declares 3", "the extents must match"},
                      "other N");
    }
    std::printf("a RaggedInput reading a dense port: its blocks are the rows\n");
    {
        cs::Model model = make([](cs::System& s) {
            s.add_subsystem(cs::subsystem_desc<DenseGen>("gen"));
            s.add_subsystem(cs::subsystem_desc<Sum>("sum"));
            s.connect("gen.d", "sum.pts");
        });
        model.setup();
        model.run();
        model.wait();
        auto sums = model.view_output<DType::f32, 1>("sum.sums");
        sums.sync();
        CHECK(sums.size() == 3, "dense sums size");
        if (kData && sums.size() == 3)
            for (std::size_t i = 0; i < 3; ++i) CHECK(sums[i] == static_cast<float>(40 * i + 6), "dense sums[%zu] = %g", i, sums[i]);
        auto d = model.view_ragged_output<DType::f32, 2>("gen.d");
        d.sync();
        CHECK(d.n() == 3 && d.block_offsets() == (std::vector<std::uint64_t>{0, 4, 8, 12}), "dense blocks");
        if (kData) CHECK(d.block(2)[3] == 23.0f, "dense block 2");
        auto in = model.view_ragged_input<DType::f32, 2>("sum.pts");
        in.sync();
        CHECK(in.block_offsets() == d.block_offsets(), "dense input blocks");
        auto dense = model.view_output<DType::f32, 2>("gen.d");  // the dense view of a dense variable is unchanged
        dense.sync();
        CHECK(dense.size() == 12, "dense view");
        if (kData) CHECK(dense[5] == 11.0f, "dense view value");
    }
    std::printf("wiring and value errors\n");
    {
        cs::Model model = make([](cs::System& s) {
            s.add_subsystem(cs::subsystem_desc<Make>("make", Sizes{1, 0, 3, 2}));
            s.add_subsystem(cs::subsystem_desc<DenseIn>("c"));
            s.connect("make.m", "c.in");
        });
        expect_throws([&] { model.setup(); }, {"connect('make.m', 'c.in')", "'make.m' is ragged", "a dense Input cannot see per-index extents", "RaggedInput"},
                      "dense input on a ragged source");
    }
    {
        cs::Model model = make([](cs::System& s) {
            s.add_subsystem(cs::subsystem_desc<Hooked<DType::i32, 1, true>>("c"));
            s.shape_hook("c.layout", cs::ragged::counts_from("c.cnt"));
        });
        model.setup();
        model.update_input("c.cnt", cs::init::from_host(std::vector<std::int32_t>{1, -1, 2}));
        expect_throws([&] { model.run(); }, {"shape hook on 'c.layout'", "holds a negative value"}, "negative counts");
        model.wait();
        model.update_input("c.cnt", cs::init::from_host(std::vector<std::int32_t>{1, 4, 2}));  // the model recovers
        model.run();
        model.wait();
        auto v = model.view_ragged_output<DType::f32, 2>("c.v");
        v.sync();
        CHECK(v.block_offsets() == (std::vector<std::uint64_t>{0, 1, 5, 7}), "after the error");
    }
    {
        cs::Model model = make([](cs::System& s) {
            s.add_subsystem(cs::subsystem_desc<Hooked<DType::u32, 1, true>>("c"));
            s.shape_hook("c.layout", cs::ragged::counts_from("c.cnt"));
        });
        model.setup();
        model.update_input("c.cnt", cs::init::from_host(std::vector<std::uint32_t>{4294967295u, 2u}));
        expect_throws([&] { model.run(); }, {"shape hook on 'c.layout'", "limited to 2^32 - 1 elements"}, "total past 32 bits");
        model.wait();
    }
    {
        cs::Model model = make([](cs::System& s) {
            s.add_subsystem(cs::subsystem_desc<Fill>("fill"));  // the hook's source is written after its target's holder
            s.add_subsystem(cs::subsystem_desc<Counts>("counts", 3));
            s.shape_hook("fill.layout", cs::ragged::counts_from("counts.counts"));
        });
        model.setup();
        model.update_input("counts.raw", cs::init::from_host(std::vector<std::uint32_t>{1, 2, 3}));
        expect_throws([&] { model.run(); }, {"shape hook on 'fill.layout'", "written by 'counts'", "'fill'"}, "source after the target's holder");
        model.wait();
    }
    {
        cs::Model model = make([](cs::System& s) {
            s.add_subsystem(cs::subsystem_desc<Make>("make", Sizes{1, 2}));
            s.add_subsystem(cs::subsystem_desc<Sum>("sum"));
            s.connect("make.m", "sum.pts");
            s.shape_hook("sum.pts", cs::ragged::counts_from("make.m"));
        });
        expect_throws([&] { model.setup(); }, {"shape_hook", "'make.m' is f32", "integer"}, "a float source");
        cs::Model m2 = make([](cs::System& s) {
            s.add_subsystem(cs::subsystem_desc<Make>("make", Sizes{1, 2}));
            s.add_subsystem(cs::subsystem_desc<Sum>("sum"));
            s.add_subsystem(cs::subsystem_desc<Counts>("counts", 2));
            s.connect("make.m", "sum.pts");
            s.shape_hook("sum.pts", cs::ragged::counts_from("counts.counts"));
        });
        expect_throws([&] { m2.setup(); }, {"shape_hook", "'sum.pts' are declared"}, "a hook on declared counts");
        cs::Model m3 = make([](cs::System& s) {
            s.add_subsystem(cs::subsystem_desc<Counts>("counts", 2));
            s.shape_hook("counts.counts", cs::ragged::counts_from("counts.raw"));
        });
        expect_throws([&] { m3.setup(); }, {"shape_hook", "is not a ragged port"}, "a hook on a dense port");
    }
#ifdef CS_HOST_MOCK
    CHECK(mock::failure().empty(), "mock: %s", mock::failure().c_str());
#endif
    if (g_fail) {
        std::printf("%d checks failed\n", g_fail);
        return 1;
    }
    std::printf("OK\n");
    return 0;
}
