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

/// The run-policy properties given to add_subsystem / set_run_policy: a set of flags (which properties were given) plus
/// run_n's count. Converts implicitly from one or more properties, so `{cs::property::transient{}}` reads as the
/// reference's property_list does; a property given twice, run_n{0}, or run_n with do_once is an std::invalid_argument.
class property_list {
public:
    property_list() = default;
    template <class... Ps>
        requires(sizeof...(Ps) >= 1 && (detail::is_property_v<Ps> && ...))
    property_list(Ps... ps) {  // implicit on purpose: the reference API takes braced property lists
        (add(ps), ...);
        if ((given_ & kRunN) && turns_ == 0) throw std::invalid_argument("naga::cospan: run_n needs n >= 1");
        if ((given_ & kRunN) && (given_ & kOnce))
            throw std::invalid_argument("naga::cospan: run_n and do_once are two run policies; give one");
    }

    template <class P>
    bool has_property() const {
        return (given_ & flag<P>()) != 0;
    }
    /// The property as given (run_n with its count).
    template <class P>
    P get_property() const {
        if (!has_property<P>()) throw std::invalid_argument("naga::cospan: the property_list does not hold that property");
        if constexpr (std::is_same_v<P, property::run_n>)
            return property::run_n{turns_};
        else
            return P{};
    }

private:
    static constexpr unsigned kTransient = 1u, kRunN = 2u, kOnce = 4u;
    template <class P>
    static constexpr unsigned flag() {
        if constexpr (std::is_same_v<P, property::transient>)
            return kTransient;
        else if constexpr (std::is_same_v<P, property::run_n>)
            return kRunN;
        else
            return kOnce;
    }
    void mark(unsigned f) {
        if (given_ & f) throw std::invalid_argument("naga::cospan: a property was given twice");
        given_ |= f;
    }
    void add(property::transient) { mark(kTransient); }
    void add(property::do_once) { mark(kOnce); }
    void add(property::run_n r) {
        mark(kRunN);
        turns_ = r.n;
    }

    unsigned    given_ = 0;
    std::size_t turns_ = 0;
};

namespace detail {

/// A subsystem's policy: counted (run_n / do_once: `left` turns remain) or transient.
struct RunPolicy {
    enum class Kind { counted, transient } kind = Kind::counted;
    std::size_t n = 0, left = 0;
    void        rearm() { left = n; }
};

/// The policy `props` asks for (none when it names no property); `where` names the subsystem in the error.
inline std::optional<RunPolicy> make_run_policy(const property_list& props, const std::string& where) {
    // do_once is run_n{1}; a count of 0 means no counted policy was asked for (run_n{0} never gets this far)
    const std::size_t turns     = props.has_property<property::run_n>()  ? props.get_property<property::run_n>().n
                                  : props.has_property<property::do_once>() ? 1
                                                                            : 0;
    const bool        transient = props.has_property<property::transient>();
    if (transient && turns != 0)
        throw std::invalid_argument("naga::cospan: " + where + ": transient cannot be combined with run_n or do_once");
    if (transient) return RunPolicy{RunPolicy::Kind::transient, 0, 0};
    if (turns != 0) return RunPolicy{RunPolicy::Kind::counted, turns, turns};
    return std::nullopt;
}

}  // namespace detail
}  // namespace naga::cospan
