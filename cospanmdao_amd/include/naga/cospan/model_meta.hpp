#pragma once

// =============================================================================
// naga/cospan/model_meta.hpp — the model's registry and its executor
// =============================================================================
//
// ModelMeta: every port and alias path (Instance -> VarMeta), the leaves in walk order (declaration order, reference
// model.hpp:904-941), the shape hooks.
//
// Executor: the reference submits each leaf on an unordered queue with explicit dependency events and records its own
// event on every container it touched (component.hpp:507-539). The MI355X executor keeps a small pool of HIP streams
// (default priority). A leaf's dependencies are the last write of each variable it reads and the last write plus the
// reads-since of each variable it writes (writes never wait on writes of OTHER variables). It launches on the stream
// of a dependency that is still that stream's latest work (so a chain stays on one stream, in order, with no waits),
// or else on the least recently used stream, where it waits (hipStreamWaitEvent) on the dependencies of other streams.
// Independent leaves therefore overlap (the reference's "parallel" style), dependent ones serialize.
//
// Graph capture: a run whose launches are byte-identical to the previous run's is captured on stream 0 (side streams
// join through fork events and are joined back before the capture ends) and every later identical run replays the
// graph: one hipGraphLaunch instead of one launch per leaf. A captured run that stayed on stream 0 can instead be re-issued
// as its recorded launches in order on stream 0 with no events (Model's ReplayMode::direct).

#include <cstddef>
#include <cstdint>
#include <map>
#include <memory>
#include <optional>
#include <stdexcept>
#include <string>
#include <string_view>
#include <vector>

#include "cospan_hip.h"
#include "naga/cospan/group_launch.hpp"
#include "naga/cospan/run_policy.hpp"
#include "naga/cospan/variable.hpp"

namespace naga::cospan::detail {

struct ModelMeta;

/// A leaf of the model's tree: a bare component (SoloDriver) or a fused group. One forward = host preparation (the
/// component steps and the description of its one launch), then the launch.
class Leaf {
public:
    virtual ~Leaf() = default;
    /// Materializes inputs, runs the user hooks, materializes outputs, points the wrappers and describes the launch.
    /// Fills `reads` and `writes` with the variables the launch touches.
    virtual LaunchRecord prepare(ModelMeta& m) = 0;
    /// The variables of every port of the leaf and their kinds (known before any forward).
    virtual void port_vars(std::vector<std::pair<VarMeta*, IOKind>>& out) const = 0;
    std::string           path;
    std::vector<VarMeta*> reads, writes;
};

class Executor {
public:
    static constexpr int kStreams = 4;

    Executor() = default;
    Executor(const Executor&)            = delete;
    Executor& operator=(const Executor&) = delete;
    ~Executor() {
        if (!ready_) return;
        for (int s = 0; s < kStreams; ++s) cs_hip_stream_synchronize(streams_[s]);
        for (int s = 0; s < kStreams; ++s) last_[s].reset();
        for (int s = 0; s < kStreams; ++s) cs_hip_stream_destroy(streams_[s]);
        if (graph_) cs_hip_graph_destroy(graph_);
    }

    void init() {
        if (ready_) return;
        hip_check(cs_hip_init(0), "device");
        for (int s = 0; s < kStreams; ++s) hip_check(cs_hip_stream_create(&streams_[s]), "stream");
        ready_ = true;
    }
    cs_hip_stream stream(int s = 0) const { return streams_[s]; }

    /// A new event recorded on stream s (the stream's latest work).
    EventPtr record(int s) {
        cs_hip_event e = nullptr;
        hip_check(cs_hip_event_create(&e, 0), "event");
        auto ev = std::make_shared<Event>(e, s);
        hip_check(cs_hip_event_record(e, streams_[s]), "event record");
        last_[s] = ev;
        return ev;
    }

    /// A deferred event standing for the work enqueued on stream s so far: no HIP event is created or recorded until a
    /// wait on another stream or the host needs one (a graph replay then costs one graph launch and nothing else).
    EventPtr deferred(int s) {
        auto ev   = std::make_shared<Event>(nullptr, s);
        ev->where = streams_[s];
        last_[s]  = ev;
        return ev;
    }

    /// Stream s waits for `e` (nothing when e is null or already in s's order).
    void wait(int s, const EventPtr& e) {
        if (!e || e->stream == s) return;
        join_capture(s);
        hip_check(cs_hip_stream_wait_event(streams_[s], e->get()), "stream wait");
    }

    /// The dependencies of a launch that reads `reads` and writes `writes`.
    static std::vector<EventPtr> deps_of(const std::vector<VarMeta*>& reads, const std::vector<VarMeta*>& writes) {
        std::vector<EventPtr> deps;
        for (VarMeta* v : reads)
            if (v->write) deps.push_back(v->write);
        for (VarMeta* v : writes) {
            if (v->write) deps.push_back(v->write);
            for (auto& r : v->reads) deps.push_back(r);
        }
        return deps;
    }

    /// The stream a launch with `deps` goes on: one whose latest work is a dependency, else the least recently used.
    int choose(const std::vector<EventPtr>& deps) {
        for (const EventPtr& d : deps)
            if (d && last_[d->stream] == d) return d->stream;
        if (deps.empty()) return 0;
        int best = 0;
        for (int s = 1; s < kStreams; ++s)
            if (use_[s] < use_[best]) best = s;
        return best;
    }

    /// A launch bracketed by timing events (eager runs with timing on).
    struct TimedLaunch {
        std::string  name;
        cs_hip_event start = nullptr, stop = nullptr;
    };
    /// While set, every submit records timing events around its launch into `sink`.
    void set_timing(std::vector<TimedLaunch>* sink) { timing_ = sink; }

    /// Launches one prepared leaf with its dependencies (and `extra` ones: depends_on) and records its event on every
    /// variable it touched; returns that event (null for an empty record).
    EventPtr submit(Leaf& leaf, const LaunchRecord& rec, const std::vector<EventPtr>& extra = {}) {
        if (rec.empty()) return nullptr;
        std::vector<EventPtr> deps = deps_of(leaf.reads, leaf.writes);
        deps.insert(deps.end(), extra.begin(), extra.end());
        const int                   s    = choose(deps);
        join_capture(s);
        for (const EventPtr& d : deps) wait(s, d);
        TimedLaunch t;
        if (timing_) {
            t.name = rec.name;
            hip_check(cs_hip_event_create(&t.start, 1), "event");
            hip_check(cs_hip_event_create(&t.stop, 1), "event");
            hip_check(cs_hip_event_record(t.start, streams_[s]), "event record");
        }
        rec.submit(streams_[s]);
        if (timing_) {
            hip_check(cs_hip_event_record(t.stop, streams_[s]), "event record");
            timing_->push_back(t);
        }
        use_[s] = ++clock_;
        ++launches_;
        const EventPtr ev = record(s);
        for (VarMeta* v : leaf.reads) {
            // a read on stream s subsumes the earlier reads on s (one stream runs in order): at most one per stream
            std::erase_if(v->reads, [s](const EventPtr& r) { return r->stream == s; });
            v->reads.push_back(ev);
        }
        for (VarMeta* v : leaf.writes) {
            v->write = ev;
            v->reads.clear();
            v->has_value = true;
        }
        return ev;
    }

    /// An initializer's work on stream 0: it overwrites the variable, so it waits like a writer.
    void before_init(VarMeta& v) {
        if (v.write) wait(0, v.write);
        for (auto& r : v.reads) wait(0, r);
    }
    void after_init(VarMeta& v) {
        v.write = record(0);
        v.reads.clear();
        v.has_value = true;
    }

    /// Stream s waits for the latest work of every other stream.
    void join_all_into(int s) {
        for (int k = 0; k < kStreams; ++k)
            if (k != s && last_[k]) wait(s, last_[k]);
    }

    /// Waits for every stream, then reports a bounded device wait of that work that gave up (std::runtime_error).
    void synchronize() {
        if (!ready_) return;
        for (int s = 0; s < kStreams; ++s) hip_check(cs_hip_stream_synchronize(streams_[s]), "synchronize");
        check_device_faults("synchronize");
    }

    // ---- graphs -------------------------------------------------------------------------------------------------
    bool capturing() const { return capturing_; }
    void begin_capture() {
        join_all_into(0);
        synchronize();  // everything before the capture is complete: no event of it is waited on inside the graph
        for (int k = 0; k < kStreams; ++k) last_[k].reset();
        hip_check(cs_hip_graph_begin(streams_[0]), "graph capture");
        capturing_ = true;
        for (int k = 0; k < kStreams; ++k) joined_[k] = (k == 0);
    }
    void end_capture() {
        for (int k = 1; k < kStreams; ++k)
            if (joined_[k]) {
                const EventPtr e = record(k);
                hip_check(cs_hip_stream_wait_event(streams_[0], e->e), "graph join");
            }
        cs_hip_graph g = nullptr;
        capturing_     = false;
        hip_check(cs_hip_graph_end(streams_[0], &g), "graph capture");
        if (graph_) cs_hip_graph_destroy(graph_);
        graph_         = g;
        single_stream_ = true;
        for (int k = 1; k < kStreams; ++k) single_stream_ = single_stream_ && !joined_[k];
        for (int k = 0; k < kStreams; ++k) last_[k].reset();
        ++captures_;
    }
    void replay() {
        join_all_into(0);
        hip_check(cs_hip_graph_launch(graph_, streams_[0]), "graph replay");
        ++replays_;
    }
    /// Whether the captured run used stream 0 alone (no side stream joined the capture).
    bool graph_single_stream() const { return graph_ != nullptr && single_stream_; }
    /// The captured run re-issued as its launches, in order, on stream 0 with no events: what the graph holds when it is
    /// single-stream (every dependency is stream order). `recs` are the run's records, identical to the captured ones.
    void relaunch(const std::vector<LaunchRecord>& recs) {
        join_all_into(0);
        for (const LaunchRecord& r : recs)
            if (!r.empty()) r.submit(streams_[0]);
        ++replays_;
        ++relaunches_;
    }
    bool has_graph() const { return graph_ != nullptr; }
    void drop_graph() {
        if (graph_) cs_hip_graph_destroy(graph_);
        graph_ = nullptr;
    }

    std::uint64_t launches() const { return launches_; }
    std::uint64_t captures() const { return captures_; }
    std::uint64_t replays() const { return replays_; }
    std::uint64_t relaunches() const { return relaunches_; }

private:
    /// During a capture a side stream joins it through an event recorded on the capturing stream 0.
    void join_capture(int s) {
        if (!capturing_ || joined_[s]) return;
        cs_hip_event e = nullptr;
        hip_check(cs_hip_event_create(&e, 0), "event");
        auto fork = std::make_shared<Event>(e, 0);
        hip_check(cs_hip_event_record(e, streams_[0]), "graph fork");
        hip_check(cs_hip_stream_wait_event(streams_[s], e), "graph fork");
        joined_[s] = true;
        forks_.push_back(fork);
    }

    bool                  ready_ = false;
    cs_hip_stream         streams_[kStreams] = {};
    EventPtr              last_[kStreams];
    std::uint64_t         use_[kStreams] = {};
    std::uint64_t         clock_ = 0, launches_ = 0, captures_ = 0, replays_ = 0, relaunches_ = 0;
    bool                  capturing_ = false, single_stream_ = false;
    bool                  joined_[kStreams] = {};
    std::vector<EventPtr> forks_;
    cs_hip_graph          graph_  = nullptr;
    std::vector<TimedLaunch>* timing_ = nullptr;
};

/// A shape hook's run-time step (reference shape_hook.hpp): a plain hook sets dimensions of its target before the walk;
/// a ragged hook (counts_from / offsets_from) builds its target's device column from the VALUES of its source, so it runs
/// as a step of the walk, after the source's producer and before the target's first user, with one host wait.
struct HookRecord {
    std::string                               target;
    std::shared_ptr<Instance>                 inst;
    std::function<void(ModelMeta&, VarMeta&)> apply;
    bool                                      ragged       = false;
    bool                                      offsets_mode = false;
    std::string                               source_path;
    std::shared_ptr<Instance>                 source;
    std::shared_ptr<DeviceColumn>             column;  ///< reused over forwards
};

/// A subsystem's run state: its policy and whether it is in scope.
struct NodeState {
    std::optional<RunPolicy> policy;
    bool                     in_scope = true;
};

struct ModelMeta {
    std::map<std::string, std::shared_ptr<Instance>, std::less<>> instances;
    std::vector<Leaf*>                                            leaves;
    std::vector<HookRecord>                                       hooks;
    std::vector<std::string>                                      nodes;  ///< every subsystem path, in order of addition
    std::map<std::string, NodeState, std::less<>>                 node_state;
    std::vector<std::pair<std::string, std::string>>              depends;  ///< (b, a): b's leaves wait on a's
    Executor                                                      exec;
    bool                                                          in_setup  = false;
    bool                                                          finalized = false;

    static std::string join(std::string_view base, std::string_view rel) {
        if (base.empty()) return std::string{rel};
        return std::string{base} + "." + std::string{rel};
    }
    /// `path` is `node` or lies under it.
    static bool under(std::string_view path, std::string_view node) {
        return path == node || (path.size() > node.size() && path.substr(0, node.size()) == node && path[node.size()] == '.');
    }
    bool has_node(std::string_view path) const {
        for (const std::string& n : nodes)
            if (n == path) return true;
        return false;
    }
    /// The indices of the leaves under `node` (empty when it has none).
    std::vector<std::size_t> leaves_under(std::string_view node) const {
        std::vector<std::size_t> out;
        for (std::size_t k = 0; k < leaves.size(); ++k)
            if (under(leaves[k]->path, node)) out.push_back(k);
        return out;
    }
    /// A leaf runs when neither it nor any subsystem above it is out of scope.
    bool leaf_in_scope(const Leaf& leaf) const {
        for (const auto& [path, st] : node_state)
            if (!st.in_scope && under(leaf.path, path)) return false;
        return true;
    }

    std::shared_ptr<Instance> find(std::string_view full) const {
        auto it = instances.find(full);
        return it == instances.end() ? nullptr : it->second;
    }
    std::shared_ptr<Instance> at(std::string_view full, std::string_view what) const {
        auto i = find(full);
        if (!i) throw std::invalid_argument("naga::cospan: " + std::string{what} + ": no port or alias named '" + std::string{full} + "'");
        return i;
    }
    void add(const std::shared_ptr<Instance>& inst) {
        if (instances.count(inst->path)) throw std::invalid_argument("naga::cospan: the path '" + inst->path + "' is already declared");
        instances.emplace(inst->path, inst);
    }

    /// Every instance of `gone` now names `keep`; definitions merge (two specified extents must agree). Both are taken by
    /// value: callers pass an instance's own `var`, which the loop below rewrites.
    void merge(const std::shared_ptr<VarMeta> keep, const std::shared_ptr<VarMeta> gone, const std::string& who) {
        if (keep == gone) return;
        if (!dtype_compatible(keep->dtype, gone->dtype) || keep->rank != gone->rank)
            throw std::invalid_argument("naga::cospan: " + who + ": '" + keep->name + "' (" + dtype_name(keep->dtype) + ", rank " +
                                        std::to_string(keep->rank) + ") and '" + gone->name + "' (" + dtype_name(gone->dtype) +
                                        ", rank " + std::to_string(gone->rank) + ") cannot share a variable");
        if (keep->data && gone->data)
            throw std::logic_error("naga::cospan: " + who + ": merging two materialized variables is not defined (yet)");
        // ragged definitions: declared counts are definitions and must agree
        if (gone->counts_fixed && gone->counts_on_host) {
            if (keep->counts_fixed && keep->counts_on_host && keep->counts != gone->counts) {
                const bool n_differs = keep->counts.size() != gone->counts.size();
                throw std::invalid_argument("naga::cospan: " + who + ": " +
                                            (n_differs ? "dimension 0: '" + keep->name + "' declares " + std::to_string(keep->counts.size()) +
                                                             " and '" + gone->name + "' declares " + std::to_string(gone->counts.size()) +
                                                             ": the extents must match; "
                                                       : std::string{}) +
                                            "dimension 1: the counts must match: '" + keep->name + "' " + keep->counts_text() + ", '" +
                                            gone->name + "' " + gone->counts_text());
            }
            if (!keep->counts_fixed) {
                keep->set_host_counts(gone->counts, who);
                keep->counts_fixed = true;
                keep->fixed[0]     = true;
            }
        }
        for (int d = 0; d < keep->rank; ++d) {
            if (gone->fixed[d]) {
                if (keep->fixed[d] && keep->dims[d] != gone->dims[d])
                    throw std::invalid_argument("naga::cospan: " + who + ": dimension " + std::to_string(d) + " of '" + keep->name + "' is " +
                                                std::to_string(keep->dims[d]) + " but '" + gone->name + "' specifies " +
                                                std::to_string(gone->dims[d]));
                keep->fixed[d] = true;
                keep->known[d] = true;
                keep->dims[d]  = gone->dims[d];
            }
        }
        keep->has_producer = keep->has_producer || gone->has_producer;
        keep->scratch      = keep->scratch || gone->scratch;
        keep->ragged_capable = keep->ragged_capable || gone->ragged_capable;
        keep->hooked_ragged  = keep->hooked_ragged || gone->hooked_ragged;
        if (!keep->init && gone->init) {
            keep->init    = gone->init;
            keep->pending = gone->pending;
        }
        for (auto& [path, inst] : instances)
            if (inst->var == gone) inst->var = keep;
    }

    /// Runs a pending initializer of `v` (allocation at the initializer's shape where it brings one).
    void materialize(VarMeta& v, const std::string& who) {
        if (!v.pending) return;
        exec.init();
        if (!v.init) throw std::logic_error("naga::cospan: '" + v.name + "' is pending without an initializer");
        if (v.init->shape) {
            const InitShape& s = *v.init->shape;
            if (s.rank != v.rank)
                throw std::invalid_argument("naga::cospan: " + who + ": the initializer of '" + v.name + "' has rank " +
                                            std::to_string(s.rank) + " but the variable has rank " + std::to_string(v.rank));
            for (int d = 0; d < v.rank; ++d) v.set_dim(d, s.extents[d], who);
        }
        if (!v.shape_known())
            throw std::invalid_argument("naga::cospan: " + who + ": '" + v.name + "' has an unknown shape " + v.shape_text() +
                                        "; give its initializer a shape or hook it");
        exec.before_init(v);
        v.ensure_allocation();
        InitCtx c{v, exec.stream(0)};
        v.init->apply(c);
        exec.after_init(v);
        v.pending = false;
    }
};

}  // namespace naga::cospan::detail
