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
//
// Ragged shape hooks (counts_from / offsets_from, reference shape_hook.hpp:400-473) need the VALUES of their source: such a
// run goes leaf by leaf (prepare, launch), with each hook placed directly after the last writer of its source (else before
// the first holder of its target); the hook waits for the device once, builds the target's column with one launch and
// reads back [total, flags]. Those runs are eager: the column's total decides allocations, so no graph is captured.

#include <algorithm>
#include <cstdint>
#include <memory>
#include <span>
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
        if (!v.data && v.shape_known() && v.elements() > 0)
            throw std::logic_error("naga::cospan: view: '" + v.name + "' has no data: every subsystem holding it is out of scope");
        if (v.write) detail::hip_check(cs_hip_event_synchronize(v.write->get()), "view sync");
        for (int d = 0; d < Rank; ++d) shape_[d] = v.dims[d];
        data_.resize(v.elements());
        if (!data_.empty()) {
            detail::hip_check(cs_hip_memcpy_async(data_.data(), v.data, data_.size() * sizeof(value_type), ex_->stream(0)), "view sync");
            detail::hip_check(cs_hip_stream_synchronize(ex_->stream(0)), "view sync");
            detail::check_device_faults("view sync");
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

/// A host copy of a rank-2 variable as N blocks end to end (reference host_view.hpp:155-196): a ragged variable's blocks
/// have its per-index counts; a dense variable's blocks are its rows.
template <DType DT, int Rank>
class HostRagged {
public:
    static_assert(Rank == 2, "naga::cospan: a ragged host view is rank 2 only");
    using value_type = cpp_type_t<DT>;
    HostRagged(std::shared_ptr<detail::VarMeta> v, detail::Executor* ex) : var_{std::move(v)}, ex_{ex} {}

    void sync() {
        detail::VarMeta& v = *var_;
        if (v.write) detail::hip_check(cs_hip_event_synchronize(v.write->get()), "view sync");
        const cs_hip_stream s = ex_->stream(0);
        n_                    = v.dims[0];
        offs_.assign(n_ + 1, 0);
        if (v.per_index && !v.counts_on_host) {
            std::vector<std::uint32_t> o(n_ + 1);
            detail::hip_check(cs_hip_memcpy_async(o.data(), v.column->offs, (n_ + 1) * 4, s), "view sync");
            detail::hip_check(cs_hip_stream_synchronize(s), "view sync");
            std::copy(o.begin(), o.end(), offs_.begin());
        } else {
            for (std::size_t i = 0; i < n_; ++i) offs_[i + 1] = offs_[i] + (v.per_index ? v.counts[i] : v.dims[1]);
        }
        data_.resize(v.elements());
        if (!data_.empty()) {
            detail::hip_check(cs_hip_memcpy_async(data_.data(), v.data, data_.size() * sizeof(value_type), s), "view sync");
            detail::hip_check(cs_hip_stream_synchronize(s), "view sync");
        }
        detail::check_device_faults("view sync");
    }
    std::size_t                       size() const { return data_.size(); }
    const value_type*                 data() const { return data_.data(); }
    std::span<const value_type>       span() const { return {data_.data(), data_.size()}; }
    const value_type&                 operator[](std::size_t flat) const { return data_[flat]; }
    /// N, the number of blocks.
    std::size_t                       n() const { return n_; }
    /// The N + 1 block offsets: block i is data()[offsets[i], offsets[i + 1]).
    const std::vector<std::uint64_t>& block_offsets() const { return offs_; }
    std::span<const value_type>       block(std::size_t i) const {
        if (i + 1 >= offs_.size()) throw std::out_of_range("naga::cospan: HostRagged::block past the last block");
        return {data_.data() + offs_[i], static_cast<std::size_t>(offs_[i + 1] - offs_[i])};
    }

private:
    std::shared_ptr<detail::VarMeta> var_;
    detail::Executor*                ex_;
    std::size_t                      n_ = 0;
    std::vector<std::uint64_t>       offs_;
    std::vector<value_type>          data_;
};

/// Counters of a model's executor.
struct ModelStats {
    std::uint64_t runs = 0, launches = 0, captures = 0, replays = 0;
    std::uint64_t relaunches = 0;  ///< replays served as direct launches (ReplayMode::direct)
};

/// How a steady-state run is replayed once it has been captured. graph: one hipGraphLaunch. direct: when the captured
/// run stayed on stream 0, its recorded launches are re-issued in order on stream 0 with no events (a graph launch costs
/// a few microseconds more than a short chain of plain launches); a run that spans streams still replays the graph.
enum class ReplayMode { graph, direct };

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
        const std::vector<bool> active = begin_turn(m);
        ++stats_.runs;
        if (const auto steps = ragged_steps(m); !steps.empty()) {
            run_stepwise(m, steps, active);
            end_turn(m, active);
            return;
        }
        std::vector<detail::LaunchRecord> recs(m.leaves.size());
        std::vector<std::byte>            sig;
        for (std::size_t k = 0; k < m.leaves.size(); ++k) {
            if (!active[k]) continue;
            recs[k] = m.leaves[k]->prepare(m);
            const std::uint64_t tag = k;
            const auto*         b   = reinterpret_cast<const std::byte*>(&tag);
            sig.insert(sig.end(), b, b + sizeof tag);
            recs[k].append_signature(sig);
        }
        detail::Executor& ex = m.exec;
        if (replay_enabled_ && ex.has_graph() && sig == graph_sig_) {
            if (replay_mode_ == ReplayMode::direct && ex.graph_single_stream())
                ex.relaunch(recs);
            else
                ex.replay();
            after_graph(m, active);
            end_turn(m, active);
            return;
        }
        if (replay_enabled_ && !sig.empty() && sig == last_sig_) {
            // stable for two runs: capture this one, then launch it
            ex.begin_capture();
            for (auto& [path, inst] : m.instances) {
                inst->var->write.reset();
                inst->var->reads.clear();
            }
            submit_all(m, recs, active);
            ex.end_capture();
            graph_sig_ = sig;
            ex.replay();
            after_graph(m, active);
            end_turn(m, active);
            return;
        }
        if (ex.has_graph()) {
            ex.drop_graph();
            graph_sig_.clear();
        }
        submit_all(m, recs, active);
        last_sig_ = std::move(sig);
        end_turn(m, active);
    }

    /// Puts the subsystem at `path` (and everything under it) back in scope and resets their policies' counters
    /// (reference model.hpp:544-552). Legal between runs.
    void rearm_subsystem(std::string_view path) {
        detail::ModelMeta& m = *meta_;
        if (!m.has_node(path)) throw std::invalid_argument("naga::cospan: Model::rearm_subsystem: no subsystem '" + std::string{path} + "'");
        for (auto& [p, st] : m.node_state)
            if (detail::ModelMeta::under(p, path)) {
                st.in_scope = true;
                if (st.policy) st.policy->rearm();
            }
    }
    /// Whether the subsystem at `path` is in scope (it and every subsystem above it).
    bool in_scope(std::string_view path) const {
        const detail::ModelMeta& m = *meta_;
        if (!m.has_node(path)) throw std::invalid_argument("naga::cospan: Model::in_scope: no subsystem '" + std::string{path} + "'");
        for (const auto& [p, st] : m.node_state)
            if (!st.in_scope && detail::ModelMeta::under(path, p)) return false;
        return true;
    }
    /// The run policy of the subsystem at `path` (absolute), as System::set_run_policy.
    void set_run_policy(std::string_view path, const property_list& policy = {}) {
        detail::ModelMeta& m = *meta_;
        if (!m.has_node(path)) throw std::invalid_argument("naga::cospan: set_run_policy: no subsystem '" + std::string{path} + "'");
        m.node_state[std::string{path}].policy = detail::make_run_policy(policy, "set_run_policy('" + std::string{path} + "')");
    }
    void clear_run_policy(std::string_view path) {
        detail::ModelMeta& m = *meta_;
        if (!m.has_node(path)) throw std::invalid_argument("naga::cospan: clear_run_policy: no subsystem '" + std::string{path} + "'");
        m.node_state[std::string{path}].policy.reset();
    }

    /// Waits for every launch of the model.
    void wait() { meta_->exec.synchronize(); }

    /// One eager forward with device timing events around every launch: (launch name, milliseconds) in launch order.
    /// No graph is captured or replayed by it.
    std::vector<std::pair<std::string, float>> run_timed() {
        detail::ModelMeta& m = *meta_;
        m.exec.init();
        for (auto& h : m.hooks) h.apply(m, *h.inst->var);
        const std::vector<bool>           active = begin_turn(m);
        const auto                        steps  = ragged_steps(m);
        std::vector<detail::LaunchRecord> recs(m.leaves.size());
        if (steps.empty())
            for (std::size_t k = 0; k < m.leaves.size(); ++k)
                if (active[k]) recs[k] = m.leaves[k]->prepare(m);
        std::vector<detail::Executor::TimedLaunch> t;
        m.exec.set_timing(&t);
        try {
            if (!steps.empty())
                run_stepwise(m, steps, active);
            else
                submit_all(m, recs, active);
        } catch (...) {
            m.exec.set_timing(nullptr);
            throw;
        }
        m.exec.set_timing(nullptr);
        end_turn(m, active);
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
        if (!ragged_steps(m).empty()) throw std::invalid_argument("naga::cospan: time_leaf: a model with ragged shape hooks runs leaf by leaf; time it with run_timed");
        for (const auto& [p, st] : m.node_state)
            if (st.policy || !st.in_scope) throw std::invalid_argument("naga::cospan: time_leaf: a model with run policies or subsystems out of scope; time it with run_timed");
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

    /// A host view of a rank-2 variable as blocks (ragged: its per-index counts; dense: its rows).
    template <DType DT, int Rank>
    HostRagged<DT, Rank> view_ragged_output(std::string_view path) {
        return view_ragged<DT, Rank>(path, detail::IOKind::output);
    }
    template <DType DT, int Rank>
    HostRagged<DT, Rank> view_ragged_input(std::string_view path) {
        return view_ragged<DT, Rank>(path, detail::IOKind::input);
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
    /// How captured runs are replayed (graph by default).
    void       set_replay_mode(ReplayMode mode) { replay_mode_ = mode; }
    ReplayMode replay_mode() const { return replay_mode_; }
    ModelStats stats() const {
        ModelStats s = stats_;
        s.launches   = meta_->exec.launches();
        s.captures   = meta_->exec.captures();
        s.replays    = meta_->exec.replays();
        s.relaunches = meta_->exec.relaunches();
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
                                        std::to_string(inst->rank) + ": the wrong dtype or rank for this view");
        if (inst->var->per_index || inst->var->hooked_ragged)
            throw std::invalid_argument("naga::cospan: view: '" + std::string{path} +
                                        "' is ragged and has no rectangular host view; use view_ragged_input / view_ragged_output");
        if (inst->var->pending) meta_->materialize(*inst->var, "view");
        return HostView<DT, Rank>(inst->var, &meta_->exec);
    }
    template <DType DT, int Rank>
    HostRagged<DT, Rank> view_ragged(std::string_view path, detail::IOKind kind) {
        auto inst = meta_->at(path, kind == detail::IOKind::output ? "view_ragged_output" : "view_ragged_input");
        if (!dtype_compatible(inst->dtype, DT) || inst->rank != Rank)
            throw std::invalid_argument("naga::cospan: view: '" + std::string{path} + "' is " + dtype_name(inst->dtype) + " of rank " +
                                        std::to_string(inst->rank) + ": the wrong dtype or rank for this view");
        if (inst->var->pending) meta_->materialize(*inst->var, "view");
        return HostRagged<DT, Rank>(inst->var, &meta_->exec);
    }

    using RaggedSteps = std::vector<std::pair<std::size_t, detail::HookRecord*>>;

    /// Where each ragged hook runs in the walk: directly after the last leaf writing its source, else before the first leaf
    /// holding its target (else after every leaf). A holder of the target before that point is an error.
    static RaggedSteps ragged_steps(detail::ModelMeta& m) {
        RaggedSteps                                              out;
        std::vector<std::pair<detail::VarMeta*, detail::IOKind>> pv;
        for (detail::HookRecord& h : m.hooks) {
            if (!h.ragged) continue;
            const detail::VarMeta* t = h.inst->var.get();
            const detail::VarMeta* s = h.source->var.get();
            if (t == s) throw std::invalid_argument("naga::cospan: shape hook on '" + h.target + "': its source is its own target");
            constexpr std::size_t none = ~std::size_t{0};
            std::size_t           first_holder = none, last_writer = none;
            for (std::size_t k = 0; k < m.leaves.size(); ++k) {
                pv.clear();
                m.leaves[k]->port_vars(pv);
                for (const auto& [v, kind] : pv) {
                    if (v == t && first_holder == none) first_holder = k;
                    if (v == s && kind == detail::IOKind::output) last_writer = k;
                }
            }
            const std::size_t at = last_writer != none ? last_writer + 1 : first_holder != none ? first_holder : m.leaves.size();
            if (first_holder != none && first_holder < at)
                throw std::invalid_argument("naga::cospan: shape hook on '" + h.target + "': its source '" + h.source_path + "' is written by '" +
                                            m.leaves[last_writer]->path + "', which runs at or after '" + m.leaves[first_holder]->path +
                                            "', a holder of '" + h.target + "'");
            out.emplace_back(at, &h);
        }
        std::stable_sort(out.begin(), out.end(), [](const auto& a, const auto& b) { return a.first < b.first; });
        return out;
    }

    /// One eager forward leaf by leaf with the ragged hooks as steps between the leaves.
    void run_stepwise(detail::ModelMeta& m, const RaggedSteps& steps, const std::vector<bool>& active) {
        detail::Executor& ex = m.exec;
        if (ex.has_graph()) {
            ex.drop_graph();
            graph_sig_.clear();
        }
        last_sig_.clear();
        std::size_t                        si = 0;
        const auto                         deps = depends_of(m);
        std::vector<detail::EventPtr>      ev(m.leaves.size());
        for (std::size_t k = 0;; ++k) {
            while (si < steps.size() && steps[si].first == k) ragged_step(m, *steps[si++].second);
            if (k == m.leaves.size()) break;
            if (!active[k]) continue;
            const detail::LaunchRecord rec = m.leaves[k]->prepare(m);
            ev[k]                          = ex.submit(*m.leaves[k], rec, extra_deps(deps[k], ev));
        }
    }

    /// For every leaf, the leaves it waits on through depends_on.
    static std::vector<std::vector<std::size_t>> depends_of(const detail::ModelMeta& m) {
        std::vector<std::vector<std::size_t>> d(m.leaves.size());
        for (const auto& [b, a] : m.depends) {
            const auto la = m.leaves_under(a);
            for (std::size_t k : m.leaves_under(b)) d[k].insert(d[k].end(), la.begin(), la.end());
        }
        return d;
    }
    static std::vector<detail::EventPtr> extra_deps(const std::vector<std::size_t>& on, const std::vector<detail::EventPtr>& ev) {
        std::vector<detail::EventPtr> out;
        for (std::size_t j : on)
            if (ev[j]) out.push_back(ev[j]);
        return out;
    }
    /// Launches the prepared records of the active leaves in order, with their depends_on waits.
    static void submit_all(detail::ModelMeta& m, const std::vector<detail::LaunchRecord>& recs, const std::vector<bool>& active) {
        const auto                    deps = depends_of(m);
        std::vector<detail::EventPtr> ev(m.leaves.size());
        for (std::size_t k = 0; k < recs.size(); ++k)
            if (active[k]) ev[k] = m.exec.submit(*m.leaves[k], recs[k], extra_deps(deps[k], ev));
    }

    /// The start of a turn (reference run_policy.hpp, TransientPolicy::before): a transient subsystem out of scope whose
    /// parents are in scope is rearmed for this turn. Returns which leaves run.
    std::vector<bool> begin_turn(detail::ModelMeta& m) {
        for (auto& [path, st] : m.node_state)
            if (st.policy && st.policy->kind == detail::RunPolicy::Kind::transient && !st.in_scope) {
                bool parents = true;
                for (const auto& [p, o] : m.node_state)
                    if (p != path && !o.in_scope && detail::ModelMeta::under(path, p)) parents = false;
                if (parents) rearm_subsystem(path);
            }
        std::vector<bool> active(m.leaves.size());
        for (std::size_t k = 0; k < m.leaves.size(); ++k) active[k] = m.leaf_in_scope(*m.leaves[k]);
        return active;
    }

    /// The end of a turn (CountedPolicy / TransientPolicy::after): a counted subsystem that ran counts a turn and is
    /// dearmed after its last; a transient one is dearmed. Dearming waits for the device, then releases the memory of
    /// every variable that no leaf in scope holds (a free input with an initializer is initialized again when needed).
    void end_turn(detail::ModelMeta& m, const std::vector<bool>& active) {
        std::vector<std::string> dearm;
        for (auto& [path, st] : m.node_state) {
            if (!st.policy) continue;
            bool ran = false;
            for (std::size_t k : m.leaves_under(path)) ran = ran || active[k];
            bool visited = true;  // in scope this turn (it and its parents)
            for (const auto& [p, o] : m.node_state)
                if (!o.in_scope && detail::ModelMeta::under(path, p)) visited = false;
            if (!visited && !ran) continue;
            if (st.policy->kind == detail::RunPolicy::Kind::transient) {
                dearm.push_back(path);
            } else if (st.policy->left > 0 && --st.policy->left == 0) {
                dearm.push_back(path);
            }
        }
        if (dearm.empty()) return;
        for (const std::string& path : dearm) {
            for (const std::string& n : m.nodes)
                if (detail::ModelMeta::under(n, path)) m.node_state[n].in_scope = false;
        }
        m.exec.synchronize();
        std::vector<std::pair<detail::VarMeta*, detail::IOKind>> pv;
        std::vector<detail::VarMeta*>                            held;  // held by a leaf in scope
        for (detail::Leaf* leaf : m.leaves)
            if (m.leaf_in_scope(*leaf)) {
                pv.clear();
                leaf->port_vars(pv);
                for (const auto& [v, kind] : pv) held.push_back(v);
            }
        bool released = false;
        for (detail::Leaf* leaf : m.leaves) {
            if (m.leaf_in_scope(*leaf)) continue;
            pv.clear();
            leaf->port_vars(pv);
            for (const auto& [v, kind] : pv) {
                if (!v->data || std::find(held.begin(), held.end(), v) != held.end()) continue;
                v->release();
                v->has_value = false;
                v->write.reset();
                v->reads.clear();
                if (v->init && !v->has_producer) v->pending = true;
                released = true;
            }
        }
        if (released && m.exec.has_graph()) {
            m.exec.drop_graph();
            graph_sig_.clear();
            last_sig_.clear();
        }
    }

    /// A counts_from / offsets_from hook (reference shape_hook.hpp:215-348, :425-473): the target's device column from the
    /// source's values, one wait for [total, flags].
    static void ragged_step(detail::ModelMeta& m, detail::HookRecord& h) {
        detail::VarMeta&  src = *h.source->var;
        detail::VarMeta&  tgt = *h.inst->var;
        const std::string who = "the shape hook on '" + h.target + "'";
        if (src.pending) m.materialize(src, who);
        if (!src.shape_known() || (!src.has_value && !src.data))
            throw std::invalid_argument("naga::cospan: " + who + ": '" + h.source_path + "' has no value yet (no source, no initial value)");
        if (src.per_index) throw std::invalid_argument("naga::cospan: " + who + ": '" + h.source_path + "' is itself ragged");
        const std::size_t count = src.dims[0];
        const std::size_t k     = src.rank == 2 ? src.dims[1] : 1;
        if (h.offsets_mode && count == 0)
            throw std::invalid_argument("naga::cospan: " + who + ": '" + h.source_path + "' holds no offsets (it needs N + 1 >= 1 entries)");
        if (!h.offsets_mode && k == 0) throw std::invalid_argument("naga::cospan: " + who + ": '" + h.source_path + "' has no column 0");
        const std::size_t n = h.offsets_mode ? count - 1 : count;
        if (tgt.fixed[0] && tgt.known[0] && tgt.dims[0] != n)
            throw std::invalid_argument("naga::cospan: " + who + ": dimension 0 of '" + tgt.name + "' is fixed at " + std::to_string(tgt.dims[0]) +
                                        " but '" + h.source_path + "' gives N = " + std::to_string(n));
        int code = 0;
        switch (src.dtype) {
            case DType::i32: code = 0; break;
            case DType::u32: code = 1; break;
            case DType::i64: code = 2; break;
            case DType::u64: code = 3; break;
            default: throw std::invalid_argument("naga::cospan: " + who + ": '" + h.source_path + "' must have an integer dtype");
        }
        // every earlier launch (the source's writers, the readers of the last column) is complete
        m.exec.synchronize();
        const cs_hip_stream s = m.exec.stream(0);
        if (!h.column) h.column = std::make_shared<detail::DeviceColumn>();
        detail::DeviceColumn& c = *h.column;
        c.reserve(n);
        c.reserve_workspace(count, s);
        detail::hip_check(cs_hip_build_column(code, src.data, count, k, h.offsets_mode ? 1 : 0, c.extents, c.offs, c.result, c.workspace, s),
                          who);
        std::uint32_t res[2] = {0, 0};
        detail::hip_check(cs_hip_memcpy_async(res, c.result, sizeof res, s), who);
        detail::hip_check(cs_hip_stream_synchronize(s), who);
        if (res[1] & CS_HIP_FLAG_WAIT_TIMEOUT) {
            detail::check_device_faults(who);  // takes the device's fault word with the scan's bits
            throw std::runtime_error("naga::cospan: " + who + ": the column build's look-back gave up; its offsets are not valid");
        }
        if ((res[1] & CS_HIP_FLAG_NEGATIVE) == 0 && (res[1] & CS_HIP_FLAG_EXCEEDS32) != 0)
            throw std::invalid_argument("naga::cospan: " + who + ": '" + h.source_path +
                                        "' holds a value or a total past 2^32 - 1; ragged totals are limited to 2^32 - 1 elements");
        if (res[1] != 0)
            throw std::invalid_argument("naga::cospan: " + who + ": '" + h.source_path + "' holds " +
                                        (h.offsets_mode ? std::string{"offsets that are negative, do not start at 0, or decrease"}
                                                        : std::string{"a negative value"}) +
                                        "; a size cannot be negative");
        c.n     = n;
        c.total = res[0];
        ++c.version;
        tgt.set_device_counts(h.column);
        src.reads.clear();  // the build was the source's one read and it is complete
    }

    /// After a graph launch on stream 0: every variable's last event is the graph's completion.
    /// After a replay on stream 0: the graph's tail (a deferred event) is every written variable's writer and every read
    /// variable's only reader (it subsumes earlier replays' reads, which ran before it on the same stream).
    void after_graph(detail::ModelMeta& m, const std::vector<bool>& active) {
        const detail::EventPtr e = m.exec.deferred(0);
        for (std::size_t k = 0; k < m.leaves.size(); ++k) {
            if (!active[k]) continue;
            for (detail::VarMeta* v : m.leaves[k]->writes) {
                v->write = e;
                v->reads.clear();
                v->has_value = true;
            }
        }
        for (std::size_t k = 0; k < m.leaves.size(); ++k) {
            if (!active[k]) continue;
            for (detail::VarMeta* v : m.leaves[k]->reads)
                if (v->write != e) {
                    v->reads.clear();
                    v->reads.push_back(e);
                }
        }
    }

    std::unique_ptr<detail::ModelMeta> meta_;
    std::unique_ptr<System>            root_;
    bool                               set_up_         = false;
    bool                               replay_enabled_ = true;
    ReplayMode                         replay_mode_    = ReplayMode::graph;
    std::vector<std::byte>             last_sig_, graph_sig_;
    ModelStats                         stats_;
};

/// A model whose root is an S constructed from `args` (reference model.hpp:974-988).
template <class S = System, class... A>
Model create_model(A&&... args) {
    return Model(std::make_unique<S>(std::forward<A>(args)...));
}

}  // namespace naga::cospan
