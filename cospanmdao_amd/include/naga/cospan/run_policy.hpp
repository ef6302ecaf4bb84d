#pragma once

// =============================================================================
// naga/cospan/run_policy.hpp — run policies: run_n, do_once, transient (reference run_policy.hpp)
// =============================================================================
//
// A run policy belongs to a subsystem (a leaf or a System) and is applied around its turn in a run:
//   run_n{n} / do_once   the subsystem runs on its first n turns; after the n-th it is dearmed
//   transient            rearmed before its turn, dearmed after it (its variables hold memory only while it runs)
// DEARM takes the subsystem out of scope: its leaves no longer run, and every variable that no leaf in scope holds
// releases its device memory (after the device has finished with it; the variable's description stays). REARM puts it
// back in scope, resets its policies' counters, and a free input whose memory was released is initialized again from its
// initializer in the next run. Model::rearm_subsystem(path) rearms by hand. The last set_run_policy wins.

#include <cstddef>
#include <optional>
#include <stdexcept>
#include <string>
#include <tuple>
#include <type_traits>

namespace naga::cospan {

namespace property {
struct run_n {
    std::size_t n;  ///< the subsystem runs on its first n turns (n >= 1)
};
struct do_once {};    ///< run_n{1}
struct transient {};  ///< in scope for its turn only
}  // namespace property

namespace detail {
template <class P>
inline constexpr bool is_property_v = std::is_same_v<std::remove_cvref_t<P>, property::transient> ||
                                      std::is_same_v<std::remove_cvref_t<P>, property::run_n> ||
                                      std::is_same_v<std::remove_cvref_t<P>, property::do_once>;
}  // namespace detail

/// Any subset of the properties, each at most once (reference run_policy.hpp property_list).
class property_list {
public:
    property_list() = default;
    template <class... Ps>
        requires(sizeof...(Ps) >= 1 && (detail::is_property_v<Ps> && ...))
    property_list(Ps... ps) {  // NOLINT: implicit, as a SYCL property list
        (set(std::move(ps)), ...);
        if (has_property<property::run_n>() && get_property<property::run_n>().n == 0)
            throw std::invalid_argument("naga::cospan: run_n needs n >= 1");
        if (has_property<property::run_n>() && has_property<property::do_once>())
            throw std::invalid_argument("naga::cospan: run_n and do_once are two run policies; give one");
    }
    template <class P>
    bool has_property() const {
        return std::get<std::optional<P>>(props_).has_value();
    }
    template <class P>
    const P& get_property() const {
        const auto& o = std::get<std::optional<P>>(props_);
        if (!o) throw std::invalid_argument("naga::cospan: the property_list does not hold that property");
        return *o;
    }

private:
    template <class P>
    void set(P p) {
        auto& o = std::get<std::optional<P>>(props_);
        if (o) throw std::invalid_argument("naga::cospan: a property was given twice");
        o = std::move(p);
    }
    std::tuple<std::optional<property::transient>, std::optional<property::run_n>, std::optional<property::do_once>> props_;
};

namespace detail {

/// A subsystem's policy: counted (run_n / do_once: `left` turns remain) or transient.
struct RunPolicy {
    enum class Kind { counted, transient } kind = Kind::counted;
    std::size_t n = 0, left = 0;
    void        rearm() { left = n; }
};

/// The policy a property_list asks for, or none. `where` prefixes the errors.
inline std::optional<RunPolicy> make_run_policy(const property_list& props, const std::string& where) {
    const bool tr    = props.has_property<property::transient>();
    const bool has_n = props.has_property<property::run_n>();
    const bool once  = props.has_property<property::do_once>();
    if (tr && (has_n || once)) throw std::invalid_argument("naga::cospan: " + where + ": transient cannot be combined with run_n or do_once");
    if (tr) return RunPolicy{RunPolicy::Kind::transient, 0, 0};
    if (has_n || once) {
        const std::size_t n = has_n ? props.get_property<property::run_n>().n : 1;
        return RunPolicy{RunPolicy::Kind::counted, n, n};
    }
    return std::nullopt;
}

}  // namespace detail
}  // namespace naga::cospan
