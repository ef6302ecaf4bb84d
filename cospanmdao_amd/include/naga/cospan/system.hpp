#pragma once

// =============================================================================
// naga/cospan/system.hpp — System: children, connect, promote, tie_inputs, shape_hook, init_input
// =============================================================================
//
// The structural verbs of the reference (cospanmdao/include/naga/cospan/system.hpp:120-231), over this build's registry:
//   add_subsystem(subsystem_desc<T>(name, args...))  a component (its bare-component leaf), a System (its setup() runs
//                                                    now) or a fused group (cs::fuse_components)
//   connect(src, dst)                                dst's variable merges into src's (the source wins)
//   promote<DT, R>(child_path, name)                 an alias in this System's namespace for the child's variable
//   tie_inputs<DT, R>({members}, name, shape)        a NEW variable every member merges into
//   shape_hook(target, cs::dim::...)                 dimensions a run sets before the walk
//   init_input(target, init)                         the initial value of a free input
// Paths are relative to the System ("r2u.B", "fw.gtp.P").

#include <initializer_list>
#include <memory>
#include <stdexcept>
#include <string>
#include <string_view>
#include <tuple>
#include <type_traits>
#include <utility>
#include <vector>

#include "naga/cospan/component.hpp"
#include "naga/cospan/model_meta.hpp"
#include "naga/cospan/variable.hpp"

namespace naga::cospan {

/// A child to add: its type, its name and its constructor arguments (reference subsystem.hpp).
template <class T, class... A>
struct SubsystemDesc {
    using type = T;
    std::string      name;
    std::tuple<A...> args;
};
template <class T, class... A>
SubsystemDesc<T, std::decay_t<A>...> subsystem_desc(std::string name, A&&... args) {
    return SubsystemDesc<T, std::decay_t<A>...>{std::move(name), std::tuple<std::decay_t<A>...>{std::forward<A>(args)...}};
}

/// Connections inside a composed system or a fused group: {{"a.y", "b.x"}, ...}, source first.
struct ConnectsDesc {
    std::vector<std::pair<std::string, std::string>> pairs;
};
inline ConnectsDesc connects_desc(std::initializer_list<std::pair<std::string, std::string>> pairs) { return ConnectsDesc{pairs}; }

/// Whether tied outputs must be atomic (reference system.hpp:178-182): several writers of one atomic variable accumulate in
/// any order; plain writers (allow_plain) are sequenced by stage order in a fused group and by launch order across leaves.
enum class TieOutputsPolicy { require_atomic, allow_plain };

/// One tie as a value: the members (inputs, or outputs) become one variable that the alias names (reference
/// fused_component.hpp:500-515, cs::tie_inputs / cs::tie_outputs inside cs::ties_desc).
struct TieDesc {
    bool                     outputs = false;
    DType                    dtype   = DType::f32;
    int                      rank    = 1;
    std::vector<std::string> members;
    std::string              alias;
    ShapeSpec                shape;
    TieOutputsPolicy         policy = TieOutputsPolicy::require_atomic;
};
struct TiesDesc {
    std::vector<TieDesc> ties;
};
inline TiesDesc ties_desc(std::initializer_list<TieDesc> t) { return TiesDesc{t}; }
template <DType DT, int Rank>
TieDesc tie_inputs(std::initializer_list<std::string_view> members, std::string alias, ShapeSpec shape = {}) {
    TieDesc t;
    t.dtype = DT;
    t.rank  = Rank;
    for (std::string_view m : members) t.members.emplace_back(m);
    t.alias = std::move(alias);
    t.shape = shape;
    return t;
}
template <DType DT, int Rank>
TieDesc tie_outputs(std::initializer_list<std::string_view> members, std::string alias, ShapeSpec shape = {},
                    TieOutputsPolicy policy = TieOutputsPolicy::require_atomic) {
    TieDesc t = tie_inputs<DT, Rank>(members, std::move(alias), shape);
    t.outputs = true;
    t.policy  = policy;
    return t;
}

namespace detail {
/// Makes the alias's variable (base path + alias) and merges every member into it.
inline void apply_tie(ModelMeta& m, const std::string& base, const TieDesc& t) {
    const char* what = t.outputs ? "tie_outputs" : "tie_inputs";
    const std::string full = ModelMeta::join(base, t.alias);
    if (t.members.size() < 2) throw std::invalid_argument(std::string{"naga::cospan: "} + what + " '" + full + "' needs at least two members");
    if (t.outputs && t.policy == TieOutputsPolicy::require_atomic && !is_atomic(t.dtype))
        throw std::invalid_argument(std::string{"naga::cospan: tie_outputs '"} + full + "': " + dtype_name(t.dtype) +
                                    " is not atomic: plain writers of one variable need TieOutputsPolicy::allow_plain (they are then "
                                    "sequenced by stage order)");
    auto v   = std::make_shared<VarMeta>();
    v->name  = full;
    v->dtype = t.dtype;
    v->rank  = t.rank;
    for (int d = 0; d < t.shape.rank; ++d)
        if (t.shape.specified[d]) {
            v->dims[d]  = t.shape.extents[d];
            v->known[d] = v->fixed[d] = true;
        }
    auto a   = std::make_shared<Instance>();
    a->path  = full;
    a->kind  = t.outputs ? IOKind::output : IOKind::input;
    a->dtype = t.dtype;
    a->rank  = t.rank;
    a->var   = v;
    m.add(a);
    for (const std::string& mp : t.members) {
        auto inst = m.at(ModelMeta::join(base, mp), what);
        if (inst->kind != a->kind)
            throw std::invalid_argument(std::string{"naga::cospan: "} + what + " '" + full + "': '" + inst->path + "' is not an " +
                                        (t.outputs ? "output" : "input"));
        if (!t.outputs && inst->var->has_producer)
            throw std::invalid_argument(std::string{"naga::cospan: tie_inputs '"} + full + "': '" + inst->path + "' is fed by a connection");
        m.merge(a->var, inst->var, std::string{what} + " '" + full + "'");
    }
}
}  // namespace detail

/// Shape-hook dimension descriptors (reference shape_hook.hpp: cs::dim::fixed_at_run, as_is).
namespace dim {
struct FixedAtRun {
    std::size_t value;
};
inline FixedAtRun fixed_at_run(std::size_t v) { return FixedAtRun{v}; }
struct AsIs {};
inline constexpr AsIs as_is{};
}  // namespace dim

/// Ragged shape-hook descriptors (reference shape_hook.hpp:79-85): the target's N and counts come from the VALUES of an
/// integer variable: its counts (column 0 of a rank-2 [N, k] source) or its N + 1 offsets.
namespace ragged {
struct CountsFrom {
    std::string path;
};
struct OffsetsFrom {
    std::string path;
};
inline CountsFrom  counts_from(std::string path) { return CountsFrom{std::move(path)}; }
inline OffsetsFrom offsets_from(std::string path) { return OffsetsFrom{std::move(path)}; }
}  // namespace ragged

class System;

namespace detail {
/// A fused group: a leaf with its own namespace of stages.
class FusedBase : public Leaf {
public:
    virtual void build(ModelMeta& m, const std::string& path) = 0;
};

template <class T>
struct is_subsystem_desc : std::false_type {};
template <class T, class... A>
struct is_subsystem_desc<SubsystemDesc<T, A...>> : std::true_type {};

template <class T, class Tuple, std::size_t... I>
std::unique_ptr<T> make_from(Tuple& t, std::index_sequence<I...>) {
    return std::make_unique<T>(std::get<I>(t)...);
}
template <class T, class... A>
std::unique_ptr<T> make_subsystem(std::tuple<A...>& args) {
    return make_from<T>(args, std::index_sequence_for<A...>{});
}
struct SystemAccess;
}  // namespace detail

class System {
public:
    virtual ~System() = default;

    /// USER: structure and wiring.
    virtual void setup() {}

    const std::string& path() const { return path_; }

    /// A subsystem with a run policy (run_n, do_once or transient; reference system.hpp:125-128).
    template <class T, class... A>
    void add_subsystem(const property_list& props, SubsystemDesc<T, A...> desc) {
        const std::string full = detail::ModelMeta::join(path_, desc.name);
        auto              pol  = detail::make_run_policy(props, "add_subsystem('" + full + "')");
        add_subsystem(std::move(desc));
        if (pol) meta("add_subsystem").node_state[full].policy = pol;
    }

    template <class T, class... A>
    void add_subsystem(SubsystemDesc<T, A...> desc) {
        detail::ModelMeta& m    = meta("add_subsystem");
        const std::string  full = detail::ModelMeta::join(path_, desc.name);
        if (m.has_node(full)) throw std::invalid_argument("naga::cospan: add_subsystem: '" + full + "' already exists");
        m.nodes.push_back(full);
        if constexpr (std::is_base_of_v<CustomComponent<T>, T>) {
            std::unique_ptr<T> c = detail::make_subsystem<T>(desc.args);
            detail::ComponentAccess::bind(*c, full, &m);
            c->declare();
            auto drv  = std::make_unique<detail::SoloCustomDriver<T>>(std::move(c));
            drv->path = full;
            m.leaves.push_back(drv.get());
            leaves_.push_back(std::move(drv));
        } else if constexpr (std::is_base_of_v<ComponentBase, T>) {
            std::unique_ptr<T> c = detail::make_subsystem<T>(desc.args);
            detail::ComponentAccess::bind(*c, full, &m);
            c->declare();
            auto drv  = std::make_unique<detail::SoloDriver<T>>(std::move(c));
            drv->path = full;
            m.leaves.push_back(drv.get());
            leaves_.push_back(std::move(drv));
        } else if constexpr (std::is_base_of_v<detail::FusedBase, T>) {
            std::unique_ptr<T> g = detail::make_subsystem<T>(desc.args);
            g->path              = full;
            g->build(m, full);
            m.leaves.push_back(g.get());
            leaves_.push_back(std::move(g));
        } else {
            static_assert(std::is_base_of_v<System, T>, "cospan: add_subsystem takes a component, a System or a fused group");
            std::unique_ptr<T> s = detail::make_subsystem<T>(desc.args);
            s->path_             = full;
            s->meta_             = &m;
            s->setup();
            systems_.push_back(std::move(s));
        }
    }

    /// Feeds `dst` (an input) from `src`; the source's variable wins (reference wiring.hpp connect).
    void connect(std::string_view src, std::string_view dst) {
        detail::ModelMeta& m = meta("connect");
        auto               s = m.at(detail::ModelMeta::join(path_, src), "connect");
        auto               d = m.at(detail::ModelMeta::join(path_, dst), "connect");
        if (d->kind != detail::IOKind::input)
            throw std::invalid_argument("naga::cospan: connect: the destination '" + d->path + "' is not an input");
        if (d->var->has_producer && d->var != s->var)
            throw std::invalid_argument("naga::cospan: connect: '" + d->path + "' already has a source");
        if (!d->ragged && ((s->ragged && s->kind == detail::IOKind::output) || s->var->per_index || s->var->hooked_ragged))
            throw std::invalid_argument("naga::cospan: connect('" + s->path + "', '" + d->path + "'): '" + s->path +
                                        "' is ragged; a dense Input cannot see per-index extents: declare '" + d->path + "' a RaggedInput");
        m.merge(s->var, d->var, "connect('" + s->path + "', '" + d->path + "')");
    }

    /// An alias in this namespace for the child's variable (DT and Rank must match).
    template <DType DT, int Rank>
    void promote(std::string_view child_path, std::string promoted_name) {
        detail::ModelMeta& m      = meta("promote");
        auto               member = m.at(detail::ModelMeta::join(path_, child_path), "promote");
        if (!dtype_compatible(member->dtype, DT) || member->rank != Rank)
            throw std::invalid_argument("naga::cospan: promote: '" + member->path + "' is " + dtype_name(member->dtype) + " of rank " +
                                        std::to_string(member->rank));
        auto a   = std::make_shared<detail::Instance>(*member);
        a->path  = detail::ModelMeta::join(path_, promoted_name);
        m.add(a);
    }

    /// A new variable every member (inputs of dtype DT and rank Rank) merges into; the alias names it.
    template <DType DT, int Rank>
    void tie_inputs(std::initializer_list<std::string_view> members, std::string name, ShapeSpec shape = {}) {
        detail::apply_tie(meta("tie_inputs"), path_, naga::cospan::tie_inputs<DT, Rank>(members, std::move(name), shape));
    }
    /// The same for outputs (reference system.hpp:178-182): several writers of one variable. Atomic dtypes accumulate in any
    /// order; plain ones need TieOutputsPolicy::allow_plain and are sequenced by launch order.
    template <DType DT, int Rank>
    void tie_outputs(std::initializer_list<std::string_view> members, std::string name, ShapeSpec shape = {},
                     TieOutputsPolicy policy = TieOutputsPolicy::require_atomic) {
        detail::apply_tie(meta("tie_outputs"), path_, naga::cospan::tie_outputs<DT, Rank>(members, std::move(name), shape, policy));
    }

    /// Dimensions of `target` set at the start of every run (cs::dim::fixed_at_run(n), cs::dim::as_is).
    template <class... D>
    void shape_hook(std::string_view target, D... dims) {
        detail::ModelMeta& m    = meta("shape_hook");
        auto               inst = m.at(detail::ModelMeta::join(path_, target), "shape_hook");
        if (static_cast<int>(sizeof...(D)) != inst->rank)
            throw std::invalid_argument("naga::cospan: shape_hook: '" + inst->path + "' has rank " + std::to_string(inst->rank) +
                                        " but the hook gives " + std::to_string(sizeof...(D)) + " dimensions");
        std::vector<std::optional<std::size_t>> fixed;
        (fixed.push_back([](auto d) -> std::optional<std::size_t> {
             if constexpr (std::is_same_v<decltype(d), dim::FixedAtRun>)
                 return d.value;
             else
                 return std::nullopt;
         }(dims)),
         ...);
        // the hooked dimensions are owned by the hook
        for (int d = 0; d < inst->rank; ++d)
            if (fixed[d]) {
                if (inst->var->fixed[d] && inst->var->dims[d] != *fixed[d])
                    throw std::invalid_argument("naga::cospan: shape_hook: dimension " + std::to_string(d) + " of '" + inst->path +
                                                "' is specified by its definition");
                inst->var->dims[d]  = *fixed[d];
                inst->var->known[d] = inst->var->fixed[d] = true;
            }
        detail::HookRecord rec;
        rec.target = inst->path;
        rec.inst   = inst;
        rec.apply  = [fixed, path = inst->path](detail::ModelMeta&, detail::VarMeta& v) {
            for (int d = 0; d < v.rank; ++d)
                if (fixed[d]) {
                    v.dims[d]  = *fixed[d];
                    v.known[d] = true;
                }
        };
        m.hooks.push_back(std::move(rec));
    }

    /// A ragged hook: the target (a ragged port or its alias, rank 2) takes N and its per-index counts, every forward, from
    /// the values of `source` (an integer variable of rank 1, or column 0 of rank 2): counts_from reads counts, offsets_from
    /// N + 1 offsets. The hook runs after the source's producer, with one host wait for the total and the flags (negative
    /// counts, totals past 2^32 - 1 are errors), as the reference's (shape_hook.hpp:425-473).
    void shape_hook(std::string_view target, ragged::CountsFrom c) { ragged_hook(target, c.path, false); }
    void shape_hook(std::string_view target, ragged::OffsetsFrom o) { ragged_hook(target, o.path, true); }

    /// `b` depends on `a` (subsystems of this System added in that order; reference system.hpp:233-250): in every run, each
    /// leaf of b waits for every leaf of a that ran, with or without a variable between them. Kept in graph replay (the
    /// waits are captured with the launches). A repeated edge is ignored.
    void depends_on(std::string_view b, std::string_view a) {
        detail::ModelMeta& m  = meta("depends_on");
        const std::string  fb = detail::ModelMeta::join(path_, b), fa = detail::ModelMeta::join(path_, a);
        if (!m.has_node(fb)) throw std::invalid_argument("naga::cospan: depends_on: '" + std::string{b} + "' is not a subsystem of '" + path_ + "'");
        if (!m.has_node(fa)) throw std::invalid_argument("naga::cospan: depends_on: '" + std::string{a} + "' is not a subsystem of '" + path_ + "'");
        const auto la = m.leaves_under(fa), lb = m.leaves_under(fb);
        if (!la.empty() && !lb.empty() && la.back() >= lb.front())
            throw std::invalid_argument("naga::cospan: depends_on: '" + std::string{b} + "' can depend only on a subsystem added before it ('" +
                                        std::string{a} + "' was not)");
        for (const auto& e : m.depends)
            if (e.first == fb && e.second == fa) return;
        m.depends.emplace_back(fb, fa);
    }

    /// Sets the run policy of the subsystem at `rel` (relative to this System; a leaf or a System): {} none, {run_n{n}},
    /// {do_once} or {transient}. The last call wins and its counter starts over; the scope is not changed
    /// (Model::rearm_subsystem puts a subsystem back in scope).
    void set_run_policy(std::string_view rel, const property_list& policy = {}) {
        detail::ModelMeta& m    = meta("set_run_policy");
        const std::string  full = detail::ModelMeta::join(path_, rel);
        if (!m.has_node(full)) throw std::invalid_argument("naga::cospan: set_run_policy: no subsystem '" + full + "'");
        m.node_state[full].policy = detail::make_run_policy(policy, "set_run_policy('" + full + "')");
    }
    /// Removes the run policy of the subsystem at `rel`: it just runs.
    void clear_run_policy(std::string_view rel) {
        detail::ModelMeta& m    = meta("clear_run_policy");
        const std::string  full = detail::ModelMeta::join(path_, rel);
        if (!m.has_node(full)) throw std::invalid_argument("naga::cospan: clear_run_policy: no subsystem '" + full + "'");
        m.node_state[full].policy.reset();
    }

    /// The initial value of a free input (a child's input or an alias of this System).
    void init_input(std::string_view target, InitSpec init) {
        detail::ModelMeta& m    = meta("init_input");
        auto               inst = m.at(detail::ModelMeta::join(path_, target), "init_input");
        if (inst->kind != detail::IOKind::input)
            throw std::invalid_argument("naga::cospan: init_input: '" + inst->path + "' is an output");
        inst->var->init    = std::move(init);
        inst->var->pending = true;
    }

private:
    void ragged_hook(std::string_view target, const std::string& source, bool offsets_mode) {
        detail::ModelMeta& m    = meta("shape_hook");
        auto               inst = m.at(detail::ModelMeta::join(path_, target), "shape_hook");
        auto               src  = m.at(detail::ModelMeta::join(path_, source), "shape_hook");
        if (inst->rank != 2 || !inst->ragged)
            throw std::invalid_argument("naga::cospan: shape_hook: '" + inst->path + "' is not a ragged port (RaggedInput / RaggedOutput of rank 2)");
        const DType d = src->dtype;
        if (d != DType::u32 && d != DType::i32 && d != DType::u64 && d != DType::i64)
            throw std::invalid_argument("naga::cospan: shape_hook: '" + src->path + "' is " + dtype_name(d) + "; counts come from an integer variable");
        if (offsets_mode ? src->rank != 1 : (src->rank != 1 && src->rank != 2))
            throw std::invalid_argument("naga::cospan: shape_hook: '" + src->path + "' has rank " + std::to_string(src->rank) +
                                        (offsets_mode ? "; offsets come from a rank-1 variable" : "; counts come from rank 1 or column 0 of rank 2"));
        detail::VarMeta& v = *inst->var;
        if (v.counts_fixed)
            throw std::invalid_argument("naga::cospan: shape_hook: the counts of '" + inst->path + "' are declared; a hook cannot own them");
        v.hooked_ragged = v.counts_fixed = true;
        detail::HookRecord rec;
        rec.target       = inst->path;
        rec.inst         = inst;
        rec.apply        = [](detail::ModelMeta&, detail::VarMeta&) {};
        rec.ragged       = true;
        rec.offsets_mode = offsets_mode;
        rec.source_path  = src->path;
        rec.source       = src;
        m.hooks.push_back(std::move(rec));
    }

    friend struct detail::SystemAccess;
    detail::ModelMeta& meta(const char* what) const {
        if (!meta_) throw std::logic_error(std::string{"naga::cospan: "} + what + " is legal inside setup() only");
        return *meta_;
    }

protected:
    std::string        path_;
    detail::ModelMeta* meta_ = nullptr;

private:
    std::vector<std::unique_ptr<detail::Leaf>> leaves_;
    std::vector<std::unique_ptr<System>>       systems_;
};

namespace detail {
struct SystemAccess {
    static void bind(System& s, std::string path, ModelMeta* m) {
        s.path_ = std::move(path);
        s.meta_ = m;
    }
};

/// The System of cs::system_from(name, [connects], descs...): its children in order, then its connections.
template <class... Descs>
class ComposedSystem : public System {
public:
    ComposedSystem(ConnectsDesc c, Descs... d) : c_{std::move(c)}, d_{std::move(d)...} {}
    void setup() override {
        std::apply([this](auto&... d) { (this->add_subsystem(d), ...); }, d_);
        for (auto& [src, dst] : c_.pairs) this->connect(src, dst);
    }

private:
    ConnectsDesc          c_;
    std::tuple<Descs...>  d_;
};
}  // namespace detail

/// A plain System of the given children with the given connections (reference fused_component.hpp:573-582).
template <class... Descs>
auto system_from(std::string name, ConnectsDesc c, Descs... d) {
    return subsystem_desc<detail::ComposedSystem<Descs...>>(std::move(name), std::move(c), std::move(d)...);
}
template <class... Descs>
    requires(sizeof...(Descs) >= 1 && !(std::is_same_v<std::tuple_element_t<0, std::tuple<Descs...>>, ConnectsDesc>))
auto system_from(std::string name, Descs... d) {
    return subsystem_desc<detail::ComposedSystem<Descs...>>(std::move(name), ConnectsDesc{}, std::move(d)...);
}

}  // namespace naga::cospan
