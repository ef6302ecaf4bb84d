#pragma once

// =============================================================================
// naga/cospan/fused_component.hpp — fuse_components, segment, global_barrier: one kernel per fused group
// =============================================================================
//
// The reference's helpers (cospanmdao/include/naga/cospan/fused_component.hpp:455-569) with the same meaning:
//   cs::fuse_components<opts>(name, [connects_desc(...)], [ties_desc(...)], items...)   one kernel of stages
//   cs::segment<cs::local_mem::shared|partitioned, cs::range_mode::whole|own>(descs...)
//   cs::global_barrier()                                               every group of the next segment waits for every
//                                                                      group of the earlier segments
//   options: cs::group_waits_enabled, cs::range_split::{flatten_and_split, pad_and_split, padded}
// How a global barrier is kept: with group waits enabled the whole group runs as ONE launch in ticket order and a group
// of a later segment waits on the earlier segments' counters. Without group waits each barrier-free run of segments is
// one launch of the same kernel, in order on one stream: the launch boundary is the barrier. On CDNA the second form is
// the fast one, because a waiting protocol needs every group to make its writes visible at agent scope (an L2
// write-back per group) where a launch boundary does it once.
// A stage outside a segment is a segment of its own. The group's stage types are a compile-time list (the item types),
// so its kernel is detail::group_kernel<MaxThreads, MinWaves, X...>, instantiated by hipcc in the translation unit that builds the
// model: the compile-time counterpart of the reference's JIT-assembled indexed call sequence (kernel.hpp:911-1004).

#include <memory>
#include <stdexcept>
#include <string>
#include <tuple>
#include <type_traits>
#include <utility>
#include <vector>

#include "naga/cospan/component.hpp"
#include "naga/cospan/system.hpp"

namespace naga::cospan {

struct group_waits_enabled {};
namespace range_split {
struct flatten_and_split {};
struct pad_and_split {};
struct padded {};
}  // namespace range_split

namespace detail {

template <LocalMode LM, RangeMode RM, class... Descs>
struct SegmentDesc {
    std::tuple<Descs...> descs;
};
template <Boundary B>
struct BoundaryItem {};

template <class... T>
struct TypeList {};
template <class... L>
struct concat;
template <>
struct concat<> {
    using type = TypeList<>;
};
template <class... A>
struct concat<TypeList<A...>> {
    using type = TypeList<A...>;
};
template <class... A, class... B, class... Rest>
struct concat<TypeList<A...>, TypeList<B...>, Rest...> : concat<TypeList<A..., B...>, Rest...> {};

template <class Item>
struct item_stages {
    using type = TypeList<>;
};
template <class X, class... A>
struct item_stages<SubsystemDesc<X, A...>> {
    using type = TypeList<X>;
};
template <LocalMode LM, RangeMode RM, class... D>
struct item_stages<SegmentDesc<LM, RM, D...>> {
    using type = TypeList<typename D::type...>;
};

template <class O, class... Opts>
inline constexpr int option_count_v = (static_cast<int>(std::is_same_v<O, Opts>) + ... + 0);

template <class T>
struct is_boundary : std::false_type {};
template <Boundary B>
struct is_boundary<BoundaryItem<B>> : std::true_type {};

template <class L>
struct launcher_of;
template <class... X>
struct launcher_of<TypeList<X...>> {
    using type = GroupLauncher<X...>;
    using pack = ArgPack<X...>;
    using comps = std::tuple<std::unique_ptr<X>...>;
    static constexpr int N = sizeof...(X);
};

/// The fused group (reference FusedGroup, fused_component.hpp:372-453).
template <bool Waits, Split Sp, class... Items>
class FusedGroup : public FusedBase {
    using Stages = typename concat<typename item_stages<Items>::type...>::type;
    using L      = launcher_of<Stages>;
    static constexpr int N = L::N;

public:
    FusedGroup(ConnectsDesc c, TiesDesc t, Items... items) : connects_{std::move(c)}, ties_{std::move(t)}, items_{std::move(items)...} {}

    void build(ModelMeta& m, const std::string& path) override {
        meta_ = &m;
        path_ = path;
        int      k        = 0;
        Boundary boundary = Boundary::none;
        std::apply([&](auto&... item) { (add_item(item, k, boundary), ...); }, items_);
        if (k != N) throw std::logic_error("naga::cospan: fused group '" + path + "': stage count mismatch");
        // ties first: a connection may name a tie alias as its source
        for (const TieDesc& t : ties_.ties) apply_tie(m, path, t);
        for (auto& [src, dst] : connects_.pairs) {
            auto s = m.at(ModelMeta::join(path, src), "connects_desc");
            auto d = m.at(ModelMeta::join(path, dst), "connects_desc");
            if (d->kind != IOKind::input)
                throw std::invalid_argument("naga::cospan: connects_desc: the destination '" + d->path + "' is not an input");
            m.merge(s->var, d->var, "fused group '" + path + "' connect");
        }
    }

    LaunchRecord prepare(ModelMeta& m) override {
        reads.clear();
        writes.clear();
        StagePlan           plans[N];
        typename L::pack    args{};
        prepare_stages(m, plans, args, std::make_index_sequence<N>{});
        // global barriers without group waits become launch boundaries (GroupLauncher::prepare)
        return launcher_.prepare(Sp, Waits, segs_, plans, args, m.exec.stream(0), path_);
    }
    void port_vars(std::vector<std::pair<VarMeta*, IOKind>>& out) const override {
        std::apply([&](const auto&... c) { ((c ? c->port_vars(out) : void()), ...); }, comps_);
    }

private:
    template <class X, class... A>
    void add_stage(SubsystemDesc<X, A...>& d, int& k) {
        std::unique_ptr<X> c = make_subsystem<X>(d.args);
        ComponentAccess::bind(*c, ModelMeta::join(path_, d.name), meta_);
        c->declare();
        place<X>(std::move(c), k, std::make_index_sequence<N>{});
        ++k;
    }
    template <class X, std::size_t... I>
    void place(std::unique_ptr<X> c, int k, std::index_sequence<I...>) {
        (void)((I == static_cast<std::size_t>(k) ? (assign<I>(c), true) : false) || ...);
    }
    template <std::size_t I, class X>
    void assign(std::unique_ptr<X>& c) {
        using Slot = std::tuple_element_t<I, typename L::comps>;
        if constexpr (std::is_same_v<Slot, std::unique_ptr<X>>) std::get<I>(comps_) = std::move(c);
    }

    template <class X, class... A>
    void add_item(SubsystemDesc<X, A...>& d, int& k, Boundary& boundary) {
        segs_.push_back(SegmentPlan{k, 1, RangeMode::own, LocalMode::partitioned, boundary});
        boundary = Boundary::none;
        add_stage(d, k);
    }
    template <LocalMode LM, RangeMode RM, class... D>
    void add_item(SegmentDesc<LM, RM, D...>& s, int& k, Boundary& boundary) {
        segs_.push_back(SegmentPlan{k, static_cast<int>(sizeof...(D)), RM, LM, boundary});
        boundary = Boundary::none;
        std::apply([&](auto&... d) { (add_stage(d, k), ...); }, s.descs);
    }
    template <Boundary B>
    void add_item(BoundaryItem<B>&, int& k, Boundary& boundary) {
        if (k == 0) throw std::invalid_argument("naga::cospan: fused group '" + path_ + "': a global_barrier cannot come first");
        boundary = B;
    }

    template <std::size_t... I>
    void prepare_stages(ModelMeta& m, StagePlan (&plans)[N], typename L::pack& args, std::index_sequence<I...>) {
        (prepare_one<I>(m, plans, args), ...);
    }
    template <std::size_t I>
    void prepare_one(ModelMeta& m, StagePlan (&plans)[N], typename L::pack& args) {
        auto& c = *std::get<I>(comps_);
        using X = std::remove_reference_t<decltype(c)>;
        c.precompute(m, reads, writes);
        std::unique_ptr<LaunchDescription> d  = c.describe_launch();
        auto*                              kl = dynamic_cast<KernelLaunch<X>*>(d.get());
        if (!kl) throw std::logic_error("naga::cospan: " + c.describe() + ": define_forward must return create_launch_config(...)");
        plans[I] = kl->plan;
        pack_get<I>(args) = kl->args;
    }
    template <std::size_t I>
    static auto& pack_get(typename L::pack& p) {
        return pack_slot<I>(p, Stages{});
    }
    template <std::size_t I, class... X>
    static auto& pack_slot(ArgPack<X...>& p, TypeList<X...>) {
        return detail::pack_get<I, X...>::get(p);
    }

    ConnectsDesc             connects_;
    TiesDesc                 ties_;
    std::tuple<Items...>     items_;
    typename L::comps        comps_;
    std::vector<SegmentPlan> segs_;
    typename L::type         launcher_;
    ModelMeta*               meta_ = nullptr;
    std::string              path_;
};

template <class... Opts>
constexpr Split split_of() {
    if constexpr (option_count_v<range_split::pad_and_split, Opts...> > 0)
        return Split::pad;
    else if constexpr (option_count_v<range_split::padded, Opts...> > 0)
        return Split::padded;
    else
        return Split::flatten;
}
template <class O>
inline constexpr bool is_fuse_option_v = std::is_same_v<O, group_waits_enabled> || std::is_same_v<O, range_split::flatten_and_split> ||
                                         std::is_same_v<O, range_split::pad_and_split> || std::is_same_v<O, range_split::padded>;
}  // namespace detail

/// A segment of consecutive stages that share groups (options: cs::local_mem::*, cs::range_mode::*).
template <class... Opts, class... Descs>
auto segment(Descs... d) {
    static_assert(sizeof...(Descs) >= 1, "naga::cospan: a segment needs at least one stage");
    constexpr detail::LocalMode lm = detail::option_count_v<local_mem::shared, Opts...> ? detail::LocalMode::shared : detail::LocalMode::partitioned;
    constexpr detail::RangeMode rm = detail::option_count_v<range_mode::whole, Opts...> ? detail::RangeMode::whole : detail::RangeMode::own;
    return detail::SegmentDesc<lm, rm, Descs...>{std::tuple<Descs...>{std::move(d)...}};
}

inline detail::BoundaryItem<detail::Boundary::global_barrier> global_barrier() { return {}; }

/// A fused group of the given items, for add_subsystem (reference fused_component.hpp:528-569).
/// Optional descriptors, in this order: cs::connects_desc(...), cs::ties_desc(...).
template <class... Opts, class... Items>
    requires(sizeof...(Items) >= 1)
auto fuse_components(std::string name, ConnectsDesc c, TiesDesc t, Items... items) {
    static_assert((detail::is_fuse_option_v<Opts> && ...), "naga::cospan: the fuse_components options are cs::group_waits_enabled and cs::range_split::*");
    constexpr bool          waits = detail::option_count_v<group_waits_enabled, Opts...> > 0;
    constexpr detail::Split sp    = detail::split_of<Opts...>();
    return subsystem_desc<detail::FusedGroup<waits, sp, Items...>>(std::move(name), std::move(c), std::move(t), std::move(items)...);
}
template <class... Opts, class... Items>
    requires(sizeof...(Items) >= 1 && !std::is_same_v<std::tuple_element_t<0, std::tuple<Items...>>, TiesDesc>)
auto fuse_components(std::string name, ConnectsDesc c, Items... items) {
    return fuse_components<Opts...>(std::move(name), std::move(c), TiesDesc{}, std::move(items)...);
}
template <class... Opts, class... Items>
    requires(sizeof...(Items) >= 1)
auto fuse_components(std::string name, TiesDesc t, Items... items) {
    return fuse_components<Opts...>(std::move(name), ConnectsDesc{}, std::move(t), std::move(items)...);
}
template <class... Opts, class... Items>
    requires(sizeof...(Items) >= 1 && !std::is_same_v<std::tuple_element_t<0, std::tuple<Items...>>, ConnectsDesc> &&
             !std::is_same_v<std::tuple_element_t<0, std::tuple<Items...>>, TiesDesc>)
auto fuse_components(std::string name, Items... items) {
    return fuse_components<Opts...>(std::move(name), ConnectsDesc{}, TiesDesc{}, std::move(items)...);
}

}  // namespace naga::cospan
