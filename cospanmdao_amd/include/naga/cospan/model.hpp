#pragma once

// =============================================================================
// naga/cospan/model.hpp — Model: setup, update_input, run, wait, view_output
// =============================================================================
//
// The reference's Model API (cospanmdao/include/naga/cospan/model.hpp:447-532): create_model<S>(args...), add_subsystem,
// setup, update_input(path, init), run, wait, view_output<DT, R>(path) / view_input.
//
// run() (reference model.hpp:474-480, walk :904-941): the shape hooks set their dimensions, then every leaf is prepared in
// declaration order on the host (inputs materialized, user hooks, outputs allocated, its ONE launch described). The
// launches then go to the executor (model_meta.hpp). Graph replay: when a run's launches are byte-identical to the
// previous run's (same kernels, shapes, pointers and arguments), the run is captured into a HIP graph; every further
// identical run replays it with one hipGraphLaunch. Any change (a resize, a new input shape, a different argument)
// falls back to eager launches, and the next stable run captures again. update_input with the same shape reuses the
// allocation, so value updates keep the graph.

#include <memory>
#include <stdexcept>
#include <string>
#include <string_view>
#include <type_traits>
#include <utility>
#include <vector>

#include "naga/cospan/component.hpp"
#include "naga/cospan/fused_component.hpp"
#include "naga/cospan/model_meta.hpp"
#include "naga/cospan/system.hpp"
#include "naga/cospan/variable.hpp"

namespace naga::cospan {

/// A host copy of a variable (reference host_view.hpp): sync() waits for the last writer and copies.
template <DType DT, int Rank>
class HostView {
public:
    using value_type = cpp_type_t<DT>;
    HostView(std::shared_ptr<detail::VarMeta> v, detail::Executor* ex) : var_{std::move(v)}, ex_{ex} {}

    void sync() {
        detail::VarMeta& v = *var_;
        if (v.write) detail::hip_check(cs_hip_event_synchronize(v.write->get()), "view sync");
        for (int d = 0; d < Rank; ++d) shape_[d] = v.dims[d];
        data_.resize(v.elements());
        if (!data_.empty()) {
            detail::hip_check(cs_hip_memcpy_async(data_.data(), v.data, data_.size() * sizeof(value_type), ex_->stream(0)), "view sync");
            detail::hip_check(cs_hip_stream_synchronize(ex_->stream(0)), "view sync");
        }
    }
    range<Rank>       shape() const { return shape_; }
    std::size_t       size() const { return data_.size(); }
    const value_type& operator[](std::size_t flat) const { return data_[flat]; }
    const value_type& operator[](id<Rank> idx) const {
        std::size_t o = 0;
        for (int d = 0; d < Rank; ++d) o = o * shape_[d] + idx[d];
        return data_[o];
    }
    const value_type* data() const { return data_.data(); }

private:
    std::shared_ptr<detail::VarMeta> var_;
    detail::Executor*                ex_;
    range<Rank>                      shape_;
    std::vector<value_type>          data_;
};

/// Counters of a model's executor.
struct ModelStats {
    std::uint64_t runs = 0, launches = 0, captures = 0, replays = 0;
};

class Model {
public:
    /// Building a model touches no device: the executor's streams are created by the first run (so structure and wiring
    /// errors surface, and can be tested, on a host without a GPU).
    explicit Model(std::unique_ptr<System> root) : meta_{std::make_unique<detail::ModelMeta>()}, root_{std::move(root)} {
        detail::SystemAccess::bind(*root_, "", meta_.get());
    }
    Model(Model&&) noexcept            = default;
    Model& operator=(Model&&) noexcept = default;
    ~Model() {
        if (meta_) meta_->exec.synchronize();
    }

    template <class T, class... A>
    void add_subsystem(SubsystemDesc<T, A...> desc) {
        root_->add_subsystem(std::move(desc));
    }

    /// Runs the root System's setup() (structure and wiring).
    void setup() {
        if (set_up_) throw std::logic_error("naga::cospan: Model::setup: the model is already set up");
        root_->setup();
        set_up_ = true;
    }

    /// The next run re-initializes `path` (an input) from `init`.
    void update_input(std::string_view path, InitSpec init) {
        auto inst = meta_->at(path, "update_input");
        if (inst->kind != detail::IOKind::input) throw std::invalid_argument("naga::cospan: update_input: '" + std::string{path} + "' is an output");
        if (inst->var->has_producer)
            throw std::invalid_argument("naga::cospan: update_input: '" + std::string{path} + "' is fed by a connection");
        inst->var->init    = std::move(init);
        inst->var->pending = true;
    }

    /// One forward of every leaf (asynchronous; wait() or a view's sync() waits).
    void run() {
        detail::ModelMeta& m = *meta_;
        m.exec.init();
        for (auto& h : m.hooks) h.apply(m, *h.inst->var);
        std::vector<detail::LaunchRecord> recs;
        recs.reserve(m.leaves.size());
        std::vector<std::byte> sig;
        for (detail::Leaf* leaf : m.leaves) {
            recs.push_back(leaf->prepare(m));
            recs.back().append_signature(sig);
        }
        ++stats_.runs;
        detail::Executor& ex = m.exec;
        if (replay_enabled_ && ex.has_graph() && sig == graph_sig_) {
            ex.replay();
            after_graph(m);
            return;
        }
        if (replay_enabled_ && !sig.empty() && sig == last_sig_) {
            // stable for two runs: capture this one, then launch it
            ex.begin_capture();
            for (auto& [path, inst] : m.instances) {
                inst->var->write.reset();
                inst->var->reads.clear();
            }
            for (std::size_t k = 0; k < recs.size(); ++k) ex.submit(*m.leaves[k], recs[k]);
            ex.end_capture();
            graph_sig_ = sig;
            ex.replay();
            after_graph(m);
            return;
        }
        if (ex.has_graph()) {
            ex.drop_graph();
            graph_sig_.clear();
        }
        for (std::size_t k = 0; k < recs.size(); ++k) ex.submit(*m.leaves[k], recs[k]);
        last_sig_ = std::move(sig);
    }

    /// Waits for every launch of the model.
    void wait() { meta_->exec.synchronize(); }

    /// One eager forward with device timing events around every launch: (launch name, milliseconds) in launch order.
    /// No graph is captured or replayed by it.
    std::vector<std::pair<std::string, float>> run_timed() {
        detail::ModelMeta& m = *meta_;
        m.exec.init();
        for (auto& h : m.hooks) h.apply(m, *h.inst->var);
        std::vector<detail::LaunchRecord> recs;
        for (detail::Leaf* leaf : m.leaves) recs.push_back(leaf->prepare(m));
        std::vector<detail::Executor::TimedLaunch> t;
        m.exec.set_timing(&t);
        try {
            for (std::size_t k = 0; k < recs.size(); ++k) m.exec.submit(*m.leaves[k], recs[k]);
        } catch (...) {
            m.exec.set_timing(nullptr);
            throw;
        }
        m.exec.set_timing(nullptr);
        m.exec.synchronize();
        std::vector<std::pair<std::string, float>> out;
        for (auto& tl : t) {
            float ms = 0.0f;
            detail::hip_check(cs_hip_event_elapsed_ms(&ms, tl.start, tl.stop), "launch timing");
            out.emplace_back(tl.name, ms);
            cs_hip_event_destroy(tl.start);
            cs_hip_event_destroy(tl.stop);
        }
        ++stats_.runs;
        return out;
    }

    /// The steady-state duration in milliseconds of leaf `k`'s launches: the leaves before it run once, then leaf k runs
    /// `iters` times back to back on one stream between two device events (a forward stage is idempotent: it rewrites
    /// the same outputs from the same inputs). Eager; no graph is captured or replayed by it.
    float time_leaf(std::size_t k, int iters) {
        detail::ModelMeta& m = *meta_;
        if (iters <= 0) throw std::invalid_argument("naga::cospan: time_leaf: iters must be positive");
        m.exec.init();
        for (auto& h : m.hooks) h.apply(m, *h.inst->var);
        std::vector<detail::LaunchRecord> recs;
        for (detail::Leaf* leaf : m.leaves) recs.push_back(leaf->prepare(m));
        if (k >= recs.size())
            throw std::invalid_argument("naga::cospan: time_leaf: the model has " + std::to_string(recs.size()) + " leaves, not " +
                                        std::to_string(k + 1));
        for (std::size_t j = 0; j < k; ++j) m.exec.submit(*m.leaves[j], recs[j]);
        m.exec.synchronize();
        const cs_hip_stream s = m.exec.stream(0);
        cs_hip_event        a = nullptr, b = nullptr;
        detail::hip_check(cs_hip_event_create(&a, 1), "event");
        detail::hip_check(cs_hip_event_create(&b, 1), "event");
        float ms = 0.0f;
        try {
            recs[k].submit(s);  // warm
            detail::hip_check(cs_hip_event_record(a, s), "event record");
            for (int i = 0; i < iters; ++i) recs[k].submit(s);
            detail::hip_check(cs_hip_event_record(b, s), "event record");
            detail::hip_check(cs_hip_event_synchronize(b), "leaf timing");
            detail::hip_check(cs_hip_event_elapsed_ms(&ms, a, b), "leaf timing");
        } catch (...) {
            cs_hip_event_destroy(a);
            cs_hip_event_destroy(b);
            throw;
        }
        cs_hip_event_destroy(a);
        cs_hip_event_destroy(b);
        m.exec.synchronize();
        return ms / static_cast<float>(iters);
    }

    template <DType DT, int Rank>
    HostView<DT, Rank> view_output(std::string_view path) {
        return view<DT, Rank>(path, detail::IOKind::output);
    }
    template <DType DT, int Rank>
    HostView<DT, Rank> view_input(std::string_view path) {
        return view<DT, Rank>(path, detail::IOKind::input);
    }

    /// Graph capture and replay of steady-state runs (on by default).
    void set_graph_replay(bool on) {
        replay_enabled_ = on;
        if (!on) {
            meta_->exec.drop_graph();
            graph_sig_.clear();
            last_sig_.clear();
        }
    }
    ModelStats stats() const {
        ModelStats s = stats_;
        s.launches   = meta_->exec.launches();
        s.captures   = meta_->exec.captures();
        s.replays    = meta_->exec.replays();
        return s;
    }
    /// The stream the model's graph and first chain run on (for external timing with events).
    cs_hip_stream stream() const { return meta_->exec.stream(0); }
    /// The device pointer of a variable (for hand-off to native code; valid until the next run that resizes it).
    void* device_data(std::string_view path) { return meta_->at(path, "device_data")->var->data; }
    std::vector<std::size_t> shape_of(std::string_view path) {
        auto v = meta_->at(path, "shape_of")->var;
        return std::vector<std::size_t>(v->dims, v->dims + v->rank);
    }

private:
    template <DType DT, int Rank>
    HostView<DT, Rank> view(std::string_view path, detail::IOKind kind) {
        auto inst = meta_->at(path, kind == detail::IOKind::output ? "view_output" : "view_input");
        if (!dtype_compatible(inst->dtype, DT) || inst->rank != Rank)
            throw std::invalid_argument("naga::cospan: view: '" + std::string{path} + "' is " + dtype_name(inst->dtype) + " of rank " +
                                        std::to_string(inst->rank));
        if (inst->var->pending) meta_->materialize(*inst->var, "view");
        return HostView<DT, Rank>(inst->var, &meta_->exec);
    }

    /// After a graph launch on stream 0: every variable's last event is the graph's completion.
    /// After a replay on stream 0: the graph's tail (a deferred event) is every written variable's writer and every read
    /// variable's only reader (it subsumes earlier replays' reads, which ran before it on the same stream).
    void after_graph(detail::ModelMeta& m) {
        const detail::EventPtr e = m.exec.deferred(0);
        for (detail::Leaf* leaf : m.leaves) {
            for (detail::VarMeta* v : leaf->writes) {
                v->write = e;
                v->reads.clear();
                v->has_value = true;
            }
        }
        for (detail::Leaf* leaf : m.leaves)
            for (detail::VarMeta* v : leaf->reads)
                if (v->write != e) {
                    v->reads.clear();
                    v->reads.push_back(e);
                }
    }

    std::unique_ptr<detail::ModelMeta> meta_;
    std::unique_ptr<System>            root_;
    bool                               set_up_         = false;
    bool                               replay_enabled_ = true;
    std::vector<std::byte>             last_sig_, graph_sig_;
    ModelStats                         stats_;
};

/// A model whose root is an S constructed from `args` (reference model.hpp:974-988).
template <class S = System, class... A>
Model create_model(A&&... args) {
    return Model(std::make_unique<S>(std::forward<A>(args)...));
}

}  // namespace naga::cospan
