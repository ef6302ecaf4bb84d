#pragma once

// =============================================================================
// naga/cospan/component.hpp — ComponentBase, Component<X>, launch descriptors, the bare-component driver
// =============================================================================
//
// The user's API of the reference (cospanmdao/include/naga/cospan/component.hpp:1445-1485): `class X : public
// cs::Component<X>` with nested Inputs/Outputs, a static compute_forward (its first parameter fixes the kind and rank:
// id<R> for an index-range body, GroupItem<R> for a group-range one, optionally a LocalMem fourth), public `inputs` and
// `outputs`, and the hooks declare() (ports), prepare_forward() (per-forward shapes) and define_forward(), which returns
// create_launch_config(range_desc(...), [wg_desc(...)], [local_desc(...)]).
//
// A forward of a component (reference component.hpp:372-431): (1) its input variables are materialized (pending
// initializers run), (2) prepare_forward() runs, (3) its outputs are allocated at their shapes, (4) the member wrappers are
// pointed at the data, (5) define_forward() describes the launch. A bare component launches as a fused group of one stage
// (reference component_driver.hpp:67-71); a fused group gathers the descriptions of its stages into one launch.

#include <cstddef>
#include <cstdint>
#include <cstring>
#include <memory>
#include <optional>
#include <stdexcept>
#include <string>
#include <string_view>
#include <tuple>
#include <type_traits>
#include <utility>
#include <vector>

#include "naga/cospan/group_launch.hpp"
#include "naga/cospan/kernel.hpp"
#include "naga/cospan/model_meta.hpp"
#include "naga/cospan/variable.hpp"
#include "naga/cospan/wrappers.hpp"

namespace naga::cospan {

// -----------------------------------------------------------------------------
// Launch descriptors (reference component.hpp:1106-1132)
// -----------------------------------------------------------------------------
template <int R>
struct RangeDesc {
    range<R> value;
};
template <int R>
struct WgDesc {
    range<R> value;
};
template <int... Rs>
struct LocalDesc {
    std::tuple<range<Rs>...> shapes;
};
template <int R>
RangeDesc<R> range_desc(const range<R>& r) {
    return RangeDesc<R>{r};
}
template <int R>
WgDesc<R> wg_desc(const range<R>& r) {
    return WgDesc<R>{r};
}
template <int... Rs>
    requires(sizeof...(Rs) >= 1)
LocalDesc<Rs...> local_desc(const range<Rs>&... shapes) {
    return LocalDesc<Rs...>{std::tuple<range<Rs>...>{shapes...}};
}

/// cs::as_is: in set_shape, "leave this dimension alone".
struct AsIsTag {};
inline constexpr AsIsTag as_is{};

/// What a component's define_forward returns: the facts its launcher needs.
class LaunchDescription {
public:
    virtual ~LaunchDescription() = default;
    detail::StagePlan plan;
};

namespace detail {

template <class X>
class KernelLaunch : public LaunchDescription {
public:
    StageArgs<X> args;
};

template <class D>
struct desc_kind : std::integral_constant<int, -1> {};
template <int R>
struct desc_kind<RangeDesc<R>> : std::integral_constant<int, 0> {};
template <int R>
struct desc_kind<WgDesc<R>> : std::integral_constant<int, 1> {};
template <int... Rs>
struct desc_kind<LocalDesc<Rs...>> : std::integral_constant<int, 2> {};

template <int R>
std::string range_text(const range<R>& r) {
    std::string s = "(";
    for (int d = 0; d < R; ++d) s += (d == 0 ? "" : ", ") + std::to_string(r[d]);
    return s + ")";
}

template <class LM, class D>
void apply_local_desc(LM& lm, const D& d) {
    if constexpr (desc_kind<D>::value == 2) lm = std::apply([](const auto&... r) { return LM{r...}; }, d.shapes);
}

struct ComponentAccess;
}  // namespace detail

// -----------------------------------------------------------------------------
// ComponentBase
// -----------------------------------------------------------------------------
// The component classes are host-and-device (a CS_COMPONENTS region): the user's classes, defined in such a region so
// that their static compute_forward is callable from the fused kernel, override these hooks with the same attributes.
// Nothing here is ever compiled for the device except what a kernel calls.
CS_COMPONENTS_BEGIN
class ComponentBase {
public:
    virtual ~ComponentBase() = default;

    /// USER: ports (declare_input / declare_output / init_input).
    virtual void declare() {}
    /// USER, every forward: output shapes (set_shape) from input shapes (shape_of).
    virtual void prepare_forward() {}

    const std::string& path() const { return path_; }
    std::string        describe() const { return "'" + path_ + "'"; }

protected:
    template <DType DT, int R>
    void declare_input(Input<DT, R>& w, std::string name, ShapeSpec shape = {}) {
        add_port(&w, std::move(name), detail::IOKind::input, DT, R, shape,
                 [&w](void* p, const std::size_t* s, const std::size_t* st) { detail::WrapperAccess::point(w, p, s, st); });
    }
    template <DType DT, int R>
    void declare_output(Output<DT, R>& w, std::string name, ShapeSpec shape = {}) {
        add_port(&w, std::move(name), detail::IOKind::output, DT, R, shape,
                 [&w](void* p, const std::size_t* s, const std::size_t* st) { detail::WrapperAccess::point(w, p, s, st); });
    }
    /// One spec per dimension, as the reference's declarations read (component.hpp:196-207): an extent or cs::unspecified
    /// for each of the port's R >= 2 dimensions (rank 1 takes its one extent or cs::unspecified through ShapeSpec).
    template <DType DT, int R, DimSpec... S>
        requires(sizeof...(S) == R && R >= 2)
    void declare_input(Input<DT, R>& w, std::string name, S... specs) {
        declare_input(w, std::move(name), shape_of_specs(specs...));
    }
    template <DType DT, int R, DimSpec... S>
        requires(sizeof...(S) == R && R >= 2)
    void declare_output(Output<DT, R>& w, std::string name, S... specs) {
        declare_output(w, std::move(name), shape_of_specs(specs...));
    }
    template <DType DT, int R, DimSpec... S>
        requires(sizeof...(S) == R && R >= 2)
    void declare_scratch(Output<DT, R>& w, std::string name, S... specs) {
        declare_scratch(w, std::move(name), shape_of_specs(specs...));
    }
    /// An output the component keeps as device state between forwards (work counters, look-back status words): zeroed
    /// whenever it is (re)allocated, and left zeroed by the kernels that use it, so every forward (and every replay of a
    /// captured graph) starts from zero without a memset.
    template <DType DT, int R>
    void declare_scratch(Output<DT, R>& w, std::string name, ShapeSpec shape = {}) {
        declare_output(w, name, shape);
        ports_.back().inst->var->scratch = true;
    }

    // ---- ragged ports (reference component.hpp:213-260): rank 2 only; a declaration gives nothing, N, or the counts ----
    template <DType DT, int R>
    void declare_input(RaggedInput<DT, R>& w, std::string name) {
        add_ragged_port(&w, std::move(name), detail::IOKind::input, DT, std::nullopt, nullptr, ragged_pointer(w));
    }
    template <DType DT, int R, std::integral N>
    void declare_input(RaggedInput<DT, R>& w, std::string name, N n) {
        add_ragged_port(&w, std::move(name), detail::IOKind::input, DT, leading_of(n), nullptr, ragged_pointer(w));
    }
    template <DType DT, int R>
    void declare_input(RaggedInput<DT, R>& w, std::string name, Counts c) {
        add_ragged_port(&w, std::move(name), detail::IOKind::input, DT, c.v.size(), &c.v, ragged_pointer(w));
    }
    template <DType DT, int R>
    void declare_input(RaggedInput<DT, R>& w, std::string name, const Offsets& o) {
        declare_input(w, std::move(name), counts_of(o));
    }
    template <DType DT, int R>
    void declare_output(RaggedOutput<DT, R>& w, std::string name) {
        add_ragged_port(&w, std::move(name), detail::IOKind::output, DT, std::nullopt, nullptr, ragged_pointer(w));
    }
    template <DType DT, int R, std::integral N>
    void declare_output(RaggedOutput<DT, R>& w, std::string name, N n) {
        add_ragged_port(&w, std::move(name), detail::IOKind::output, DT, leading_of(n), nullptr, ragged_pointer(w));
    }
    template <DType DT, int R>
    void declare_output(RaggedOutput<DT, R>& w, std::string name, Counts c) {
        add_ragged_port(&w, std::move(name), detail::IOKind::output, DT, c.v.size(), &c.v, ragged_pointer(w));
    }
    template <DType DT, int R>
    void declare_output(RaggedOutput<DT, R>& w, std::string name, const Offsets& o) {
        declare_output(w, std::move(name), counts_of(o));
    }

    /// The shape of a ragged port now: N and its counts (host or device); a dense variable reads as N rows of equal counts.
    template <class W>
        requires(W::ragged)
    RaggedShape shape_of(const W& w) const {
        const Port&            p = const_cast<ComponentBase*>(this)->port_of(&w, "shape_of");
        const detail::VarMeta& v = *p.inst->var;
        if (!v.shape_known())
            throw std::invalid_argument("naga::cospan: shape_of: " + describe() + ": '" + v.name + "' has an unknown shape " + v.shape_text());
        RaggedShape r;
        r.n = v.dims[0];
        if (v.per_index) {
            r.total = v.total;
            if (v.counts_on_host)
                r.counts = v.counts;
            else
                r.column = v.column;
        } else {
            const std::size_t e = v.rank == 2 ? v.dims[1] : 1;
            r.counts.assign(r.n, e);
            r.total = static_cast<std::uint64_t>(r.n) * e;
        }
        return r;
    }
    /// A ragged output's N and counts from a RaggedShape (shape_of of a ragged input, or built).
    template <DType DT, int R>
    void set_shape(RaggedOutput<DT, R>& w, const RaggedShape& shape) {
        detail::VarMeta& v = ragged_settable(&w);
        if (shape.column)
            v.set_device_counts(std::const_pointer_cast<detail::DeviceColumn>(shape.column));
        else
            v.set_host_counts(shape.counts, describe());
    }
    /// A ragged output's N (a number, or cs::as_is) and its host counts (counts.size() == N).
    template <DType DT, int R, class N>
    void set_shape(RaggedOutput<DT, R>& w, N n, std::vector<std::size_t> c) {
        detail::VarMeta& v = ragged_settable(&w);
        if constexpr (!std::is_same_v<N, AsIsTag>)
            if (static_cast<std::size_t>(n) != c.size())
                throw std::invalid_argument("naga::cospan: set_shape: " + describe() + ": N = " + std::to_string(static_cast<std::size_t>(n)) +
                                            " but " + std::to_string(c.size()) + " counts");
        v.set_host_counts(std::move(c), describe());
    }

    /// The initial value of one of this component's inputs (the last declaration wins; ignored when the input is fed).
    void init_input(std::string_view name, InitSpec init) {
        Port&                     p = port_named(name, "init_input");
        std::shared_ptr<detail::VarMeta> v = p.inst->var;
        if (p.kind != detail::IOKind::input)
            throw std::invalid_argument("naga::cospan: init_input: " + describe() + "." + std::string{name} + " is an output");
        v->init    = std::move(init);
        v->pending = true;
    }

    /// Sets the run-time shape of an output (or an unspecified input dimension); cs::as_is leaves a dimension alone.
    template <class W, class... E>
        requires(!requires { W::ragged; })
    void set_shape(W& w, E... dims) {
        static_assert(sizeof...(E) == W::rank, "cospan: set_shape takes one entry per dimension");
        Port&           p = port_of(&w, "set_shape");
        detail::VarMeta& v = *p.inst->var;
        int              d = 0;
        auto             one = [&](auto e) {
            if constexpr (!std::is_same_v<std::remove_cvref_t<decltype(e)>, AsIsTag>) {
                if constexpr (std::is_signed_v<decltype(e)>)
                    if (e < 0) throw std::invalid_argument("naga::cospan: set_shape: a negative extent for " + v.name);
                v.set_dim(d, static_cast<std::size_t>(e), describe());
            }
            ++d;
        };
        (one(dims), ...);
    }

    /// The run-time shape of a port (an error while it is unknown).
    template <class W>
        requires(!requires { W::ragged; })
    range<W::rank> shape_of(const W& w) const {
        const Port&            p = const_cast<ComponentBase*>(this)->port_of(&w, "shape_of");
        const detail::VarMeta& v = *p.inst->var;
        if (!v.shape_known())
            throw std::invalid_argument("naga::cospan: shape_of: " + describe() + ": '" + v.name + "' has an unknown shape " + v.shape_text());
        range<W::rank> r;
        for (int d = 0; d < W::rank; ++d) r[d] = v.dims[d];
        return r;
    }

private:
    friend struct detail::ComponentAccess;
    using RaggedPoint = std::function<void(void*, std::size_t, const std::uint32_t*, const std::uint32_t*, std::size_t)>;
    struct Port {
        std::string                                                        name;
        detail::IOKind                                                     kind;
        const void*                                                        wrapper;
        std::function<void(void*, const std::size_t*, const std::size_t*)> point;
        std::shared_ptr<detail::Instance>                                  inst;
        RaggedPoint                                                        point_ragged;  ///< set for a ragged port
    };

    template <class W>
    static RaggedPoint ragged_pointer(W& w) {
        return [&w](void* d, std::size_t n, const std::uint32_t* o, const std::uint32_t* per, std::size_t ext) {
            detail::WrapperAccess::point_ragged(w, d, n, o, per, ext);
        };
    }
    template <std::integral N>
    static std::size_t leading_of(N n) {
        if constexpr (std::is_signed_v<N>)
            if (n < 0) throw std::invalid_argument("naga::cospan: a ragged declaration: N cannot be negative");
        return static_cast<std::size_t>(n);
    }
    static Counts counts_of(const Offsets& o) {
        if (o.v.empty() || o.v[0] != 0) throw std::invalid_argument("naga::cospan: cs::offsets: the offsets must start at 0");
        Counts c;
        for (std::size_t i = 1; i < o.v.size(); ++i) {
            if (o.v[i] < o.v[i - 1]) throw std::invalid_argument("naga::cospan: cs::offsets: the offsets decrease at " + std::to_string(i));
            c.v.push_back(o.v[i] - o.v[i - 1]);
        }
        return c;
    }
    void add_ragged_port(const void* wrapper, std::string name, detail::IOKind kind, DType dt, std::optional<std::size_t> n,
                         const std::vector<std::size_t>* c, RaggedPoint point) {
        add_port(wrapper, std::move(name), kind, dt, 2, ShapeSpec{}, nullptr);
        Port&            p = ports_.back();
        detail::VarMeta& v = *p.inst->var;
        p.point_ragged     = std::move(point);
        p.inst->ragged     = true;
        v.ragged_capable   = true;
        if (n) {
            v.dims[0]  = *n;
            v.known[0] = v.fixed[0] = true;
        }
        if (c) {
            v.set_host_counts(*c, describe());
            v.counts_fixed = true;
            v.fixed[0]     = true;
        }
    }
    detail::VarMeta& ragged_settable(const void* w) {
        Port&            p = port_of(w, "set_shape");
        detail::VarMeta& v = *p.inst->var;
        if (p.kind != detail::IOKind::output)
            throw std::invalid_argument("naga::cospan: set_shape: " + describe() + "." + p.name + " is an input: inputs are never reshaped");
        if (v.counts_fixed)
            throw std::invalid_argument("naga::cospan: set_shape: " + describe() + "." + p.name + ": the counts of '" + v.name +
                                        "' come from " + (v.hooked_ragged ? "a shape hook" : "its declaration") + "; set_shape cannot change them");
        return v;
    }

    void add_port(const void* wrapper, std::string name, detail::IOKind kind, DType dt, int rank, const ShapeSpec& shape,
                  std::function<void(void*, const std::size_t*, const std::size_t*)> point) {
        if (!meta_) throw std::logic_error("naga::cospan: declare_input/declare_output are legal inside declare() only");
        if (shape.rank != 0 && shape.rank != rank)
            throw std::invalid_argument("naga::cospan: " + describe() + "." + name + ": a shape of rank " + std::to_string(shape.rank) +
                                        " for a port of rank " + std::to_string(rank));
        auto var  = std::make_shared<detail::VarMeta>();
        var->name = path_ + "." + name;
        var->dtype = dt;
        var->rank  = rank;
        for (int d = 0; d < shape.rank; ++d)
            if (shape.specified[d]) {
                var->dims[d]  = shape.extents[d];
                var->known[d] = var->fixed[d] = true;
            }
        var->has_producer = kind == detail::IOKind::output;
        auto inst         = std::make_shared<detail::Instance>();
        inst->path        = var->name;
        inst->kind        = kind;
        inst->dtype       = dt;
        inst->rank        = rank;
        inst->var         = var;
        meta_->add(inst);
        ports_.push_back(Port{std::move(name), kind, wrapper, std::move(point), inst});
    }
    Port& port_named(std::string_view name, const char* what) {
        for (Port& p : ports_)
            if (p.name == name) return p;
        throw std::invalid_argument("naga::cospan: " + std::string{what} + ": " + describe() + " has no port '" + std::string{name} + "'");
    }
    Port& port_of(const void* w, const char* what) {
        for (Port& p : ports_)
            if (p.wrapper == w) return p;
        throw std::invalid_argument("naga::cospan: " + std::string{what} + ": the wrapper is not a declared port of " + describe());
    }

protected:
    std::string        path_;
    detail::ModelMeta* meta_ = nullptr;

private:
    std::vector<Port> ports_;

public:
    /// The forward steps (framework): materialize inputs, prepare_forward, allocate outputs, point the wrappers. Fills the
    /// variables read and written.
    void precompute(detail::ModelMeta& m, std::vector<detail::VarMeta*>& reads, std::vector<detail::VarMeta*>& writes) {
        for (Port& p : ports_) {
            if (p.kind != detail::IOKind::input) continue;
            detail::VarMeta& v = *p.inst->var;
            if (v.has_producer) {
                v.pending = false;  // an initializer is ignored once the input has a source
            } else if (v.pending) {
                m.materialize(v, describe());
            } else if (v.hooked_ragged) {
                // a free ragged input whose shape a hook gives: it carries a shape; its data (zeroed) is never written
                if (v.ensure_allocation()) detail::hip_check(cs_hip_fill_u8_async(v.data, 0, v.capacity, m.exec.stream(0)), "zeroing of '" + v.name + "'");
                v.has_value = true;
            } else if (!v.has_value && v.shape_known() && v.elements() == 0) {
                v.has_value = true;  // an empty input needs no value
            } else if (!v.has_value && !v.data) {
                throw std::invalid_argument("naga::cospan: " + describe() + ": input '" + p.name + "' (" + v.name +
                                            ") has no source and no initial value");
            }
        }
        prepare_forward();
        for (Port& p : ports_) {
            detail::VarMeta& v = *p.inst->var;
            if (!v.shape_known())
                throw std::invalid_argument("naga::cospan: " + describe() + ": '" + p.name + "' has an unknown shape " + v.shape_text() +
                                            " after prepare_forward");
            if (p.kind == detail::IOKind::output && v.ensure_allocation() && v.scratch) {
                // scratch state starts at zero; the kernels that use it leave it zeroed (graph-safe, no memset node)
                m.exec.init();
                m.exec.before_init(v);
                detail::hip_check(cs_hip_fill_u8_async(v.data, 0, v.capacity, m.exec.stream(0)), "zeroing of '" + v.name + "'");
                m.exec.after_init(v);
            }
            if (p.point_ragged) {
                if (v.per_index) {
                    v.ensure_device_counts(m.exec.stream(0));
                    p.point_ragged(v.data, v.dims[0], v.column->offs, v.column->extents, 0);
                } else {
                    p.point_ragged(v.data, v.dims[0], nullptr, nullptr, v.rank == 2 ? v.dims[1] : 1);  // a dense variable: rows
                }
            } else {
                if (v.per_index)
                    throw std::invalid_argument("naga::cospan: " + describe() + ": '" + v.name +
                                                "' is ragged; a dense Input / Output cannot see per-index extents: declare '" + p.name +
                                                "' with RaggedInput / RaggedOutput");
                std::size_t stride[3];
                detail::row_major_strides(v, stride);
                p.point(v.data, v.dims, stride);
            }
            (p.kind == detail::IOKind::input ? reads : writes).push_back(&v);
        }
    }

    virtual std::unique_ptr<LaunchDescription> describe_launch() = 0;

    /// The variables of the ports and their kinds (where a ragged shape hook runs relative to the leaves).
    void port_vars(std::vector<std::pair<detail::VarMeta*, detail::IOKind>>& out) const {
        for (const Port& p : ports_) out.emplace_back(p.inst->var.get(), p.kind);
    }
};

namespace detail {
struct ComponentAccess {
    static void bind(ComponentBase& c, std::string path, ModelMeta* m) {
        c.path_ = std::move(path);
        c.meta_ = m;
    }
};
}  // namespace detail

// -----------------------------------------------------------------------------
// Component<X>
// -----------------------------------------------------------------------------
template <class X>
class Component : public ComponentBase {
public:
    /// USER, every forward: the launch, from create_launch_config(descriptors...).
    virtual std::unique_ptr<LaunchDescription> define_forward() = 0;

    std::unique_ptr<LaunchDescription> describe_launch() override { return define_forward(); }

protected:
    /// range_desc (required), wg_desc (required for a group-range body, rejected for an index-range one), local_desc
    /// (optional, group-range bodies with a LocalMem parameter). A group-range range must be a multiple of the work-group
    /// shape in every dimension (reference component.hpp:1224-1235 message).
    template <class... Ds>
    std::unique_ptr<LaunchDescription> create_launch_config(Ds... ds) {
        static_assert(KernelStage<X>, "cospan: compute_forward must take id<R> or GroupItem<R>, Inputs, Outputs[, LocalMem]");
        constexpr int  R  = detail::kernel_rank_v<X>;
        constexpr int  nr = ((detail::desc_kind<Ds>::value == 0 ? 1 : 0) + ... + 0);
        constexpr int  nw = ((detail::desc_kind<Ds>::value == 1 ? 1 : 0) + ... + 0);
        constexpr int  nl = ((detail::desc_kind<Ds>::value == 2 ? 1 : 0) + ... + 0);
        static_assert(nr == 1, "cospan: create_launch_config needs exactly one range_desc");
        static_assert(GroupRangeStage<X> ? nw == 1 : nw == 0, "cospan: wg_desc is required for a group-range body and rejected for an index-range one");
        static_assert(nl == 0 || detail::has_local_mem_v<X>, "cospan: local_desc needs a LocalMem parameter");
        X&                      self = static_cast<X&>(*this);
        const std::tuple<Ds...> t{ds...};
        const range<R>          r = std::get<RangeDesc<R>>(t).value;
        auto                    kl = std::make_unique<detail::KernelLaunch<X>>();
        detail::StagePlan&      pl = kl->plan;
        pl.who                     = describe();
        const range<3>          r3 = detail::lift_range<R>(r);
        if constexpr (GroupRangeStage<X>) {
            const range<R> wg  = std::get<WgDesc<R>>(t).value;
            const range<3> wg3 = detail::lift_range<R>(wg);
            for (int d = 0; d < R; ++d)
                if (wg[d] == 0)
                    throw std::invalid_argument("naga::cospan: " + describe() + ": the work-group shape " + detail::range_text<R>(wg) +
                                                " has a zero dimension");
            if (r.size() != 0)
                for (int d = 0; d < R; ++d)
                    if (r[d] % wg[d] != 0)
                        throw std::invalid_argument("naga::cospan: " + describe() + ": the range " + detail::range_text<R>(r) +
                                                    " is not a multiple of the work-group shape " + detail::range_text<R>(wg) +
                                                    " in dimension " + std::to_string(d) +
                                                    "; round the range up to whole work-groups and guard the tail in the body");
            for (int d = 0; d < 3; ++d) {
                pl.units[d] = r3[d] / wg3[d];
                pl.wg[d]    = static_cast<std::uint32_t>(wg3[d]);
            }
            pl.group_range = true;
            if constexpr (detail::has_local_mem_v<X>) {
                detail::local_mem_t<X> lm{};
                (detail::apply_local_desc(lm, ds), ...);
                kl->args.local = lm;
                pl.local_bytes = lm.bytes();
                pl.local_align = 16;
            }
        } else {
            for (int d = 0; d < 3; ++d) pl.units[d] = r3[d];
            pl.group_range = false;
        }
        for (int d = 0; d < 3; ++d) kl->args.units[d] = pl.units[d];
        kl->args.in  = self.inputs;
        kl->args.out = self.outputs;
        return kl;
    }
};

/// What a custom-op body receives (reference kernel.hpp:56-68, "native backend work"): the stream its work goes on, and
/// the device pointers of its ports.
struct CustomOp {
    cs_hip_stream stream;
    template <class W>
    auto device_ptr(const W& w) const {
        return raw::data(w);
    }
};

namespace detail {
template <class X>
class CustomLaunch : public LaunchDescription {
public:
    typename X::Inputs  in;
    typename X::Outputs out;
};
}  // namespace detail

/// A CUSTOM-OP component (reference component.hpp:1490-1503): `static void compute_forward(cs::CustomOp&, const Inputs&,
/// const Outputs&)` enqueues native work (a C-boundary call, a library kernel) on op.stream. No define_forward; not fusible.
template <class X>
class CustomComponent : public ComponentBase {
public:
    std::unique_ptr<LaunchDescription> describe_launch() override {
        X&   self = static_cast<X&>(*this);
        auto d    = std::make_unique<detail::CustomLaunch<X>>();
        d->in     = self.inputs;
        d->out    = self.outputs;
        return d;
    }
};
CS_COMPONENTS_END

namespace detail {

/// The segment layout of a bare component: one segment of one stage (the native launch).
inline std::vector<SegmentPlan> single_segment() { return {SegmentPlan{0, 1}}; }

/// The leaf of a bare component (reference component_driver.hpp:67-71).
template <class X>
class SoloDriver : public Leaf {
public:
    explicit SoloDriver(std::unique_ptr<X> c) : comp_{std::move(c)} {}
    X&           component() { return *comp_; }
    LaunchRecord prepare(ModelMeta& m) override {
        reads.clear();
        writes.clear();
        comp_->precompute(m, reads, writes);
        std::unique_ptr<LaunchDescription> d = comp_->describe_launch();
        auto* kl = dynamic_cast<KernelLaunch<X>*>(d.get());
        if (!kl) throw std::logic_error("naga::cospan: " + comp_->describe() + ": define_forward must return create_launch_config(...)");
        const StagePlan plans[1] = {kl->plan};
        ArgPack<X>      args{};
        args.head = kl->args;
        return launcher_.prepare(Split::pad, false, single_segment(), plans, args, m.exec.stream(0), path);
    }
    void port_vars(std::vector<std::pair<VarMeta*, IOKind>>& out) const override { comp_->port_vars(out); }

private:
    std::unique_ptr<X> comp_;
    GroupLauncher<X>   launcher_;
};

/// The leaf of a custom-op component: its launch is the body itself, run on the executor's chosen stream.
template <class X>
class SoloCustomDriver : public Leaf {
public:
    explicit SoloCustomDriver(std::unique_ptr<X> c) : comp_{std::move(c)} {}
    LaunchRecord prepare(ModelMeta& m) override {
        reads.clear();
        writes.clear();
        comp_->precompute(m, reads, writes);
        std::unique_ptr<LaunchDescription> d  = comp_->describe_launch();
        auto*                              cl = dynamic_cast<CustomLaunch<X>*>(d.get());
        LaunchRecord                       rec;
        rec.name = path;
        static_assert(std::is_trivially_copyable_v<typename X::Inputs> && std::is_trivially_copyable_v<typename X::Outputs>);
        rec.args.resize(sizeof(typename X::Inputs) + sizeof(typename X::Outputs));
        std::memcpy(rec.args.data(), &cl->in, sizeof cl->in);
        std::memcpy(rec.args.data() + sizeof cl->in, &cl->out, sizeof cl->out);
        rec.custom = [in = cl->in, out = cl->out](cs_hip_stream s) {
            CustomOp op{s};
            X::compute_forward(op, in, out);
        };
        return rec;
    }
    void port_vars(std::vector<std::pair<VarMeta*, IOKind>>& out) const override { comp_->port_vars(out); }

private:
    std::unique_ptr<X> comp_;
};

}  // namespace detail
}  // namespace naga::cospan
