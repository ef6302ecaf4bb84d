#pragma once

// =============================================================================
// naga/cospan/vocab.hpp — the index vocabulary of component bodies, native to HIP
// =============================================================================
//
// Component bodies of the reference name a small vocabulary of index and function-object types (sycl::id<R>,
// sycl::range<R>, sycl::plus<T>, ...; e.g. examples/gemver-chain/components.hpp:85-97, :237-250). This header defines
// those types natively (plain aggregates of std::size_t, usable on host and device) in naga::cospan, and makes the same
// names visible under `sycl::` so that existing bodies compile unchanged. Nothing else of SYCL exists here: there is no
// queue, handler or accessor, and no runtime behind these names.
//
// CS_COMPONENTS_BEGIN / CS_COMPONENTS_END: HIP needs every function a kernel calls to be a device function. A model's
// component definitions placed between the two macros are host-and-device without per-function annotations (clang's
// force_cuda_host_device region); functions in the region that are never called from a kernel are never compiled for
// the device.

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

#define CS_HD __host__ __device__
#define CS_COMPONENTS_BEGIN _Pragma("clang force_cuda_host_device begin")
#define CS_COMPONENTS_END _Pragma("clang force_cuda_host_device end")

namespace naga::cospan {

template <int R>
struct id {
    static_assert(R >= 1 && R <= 3, "cospan: rank is 1..3");
    std::size_t v[R] = {};

    CS_HD constexpr id() = default;
    template <class... T>
        requires(sizeof...(T) == R && R > 0)
    CS_HD constexpr id(T... xs) : v{static_cast<std::size_t>(xs)...} {}

    CS_HD constexpr std::size_t&       operator[](int d) { return v[d]; }
    CS_HD constexpr const std::size_t& operator[](int d) const { return v[d]; }
    CS_HD constexpr std::size_t        get(int d) const { return v[d]; }
    CS_HD constexpr operator std::size_t() const
        requires(R == 1)
    {
        return v[0];
    }
    CS_HD friend constexpr bool operator==(const id& a, const id& b) {
        for (int d = 0; d < R; ++d)
            if (a.v[d] != b.v[d]) return false;
        return true;
    }
};

template <int R>
struct range {
    static_assert(R >= 1 && R <= 3, "cospan: rank is 1..3");
    std::size_t v[R] = {};

    CS_HD constexpr range() = default;
    template <class... T>
        requires(sizeof...(T) == R && R > 0)
    CS_HD constexpr range(T... xs) : v{static_cast<std::size_t>(xs)...} {}

    CS_HD constexpr std::size_t&       operator[](int d) { return v[d]; }
    CS_HD constexpr const std::size_t& operator[](int d) const { return v[d]; }
    CS_HD constexpr std::size_t        get(int d) const { return v[d]; }
    CS_HD constexpr std::size_t        size() const {
        std::size_t n = 1;
        for (int d = 0; d < R; ++d) n *= v[d];
        return n;
    }
    CS_HD friend constexpr bool operator==(const range& a, const range& b) {
        for (int d = 0; d < R; ++d)
            if (a.v[d] != b.v[d]) return false;
        return true;
    }
};

template <class T = void>
struct plus {
    CS_HD constexpr T operator()(const T& a, const T& b) const { return a + b; }
    static constexpr CS_HD T identity() { return T(0); }
};
template <class T = void>
struct multiplies {
    CS_HD constexpr T operator()(const T& a, const T& b) const { return a * b; }
    static constexpr CS_HD T identity() { return T(1); }
};
template <class T = void>
struct minimum {
    CS_HD constexpr T operator()(const T& a, const T& b) const { return b < a ? b : a; }
};
template <class T = void>
struct maximum {
    CS_HD constexpr T operator()(const T& a, const T& b) const { return a < b ? b : a; }
};
template <class T = void>
struct bit_and {
    CS_HD constexpr T operator()(const T& a, const T& b) const { return a & b; }
};
template <class T = void>
struct bit_or {
    CS_HD constexpr T operator()(const T& a, const T& b) const { return a | b; }
};
template <class T = void>
struct bit_xor {
    CS_HD constexpr T operator()(const T& a, const T& b) const { return a ^ b; }
};

/// The memory scopes a group barrier names (accepted for source compatibility; a work-group barrier on CDNA4 is
/// s_barrier with the fences the scope asks for).
enum class memory_scope { work_item, sub_group, work_group, device, system };

}  // namespace naga::cospan

/// The vocabulary names under sycl:: (types only; see the header comment).
namespace sycl {
using naga::cospan::bit_and;
using naga::cospan::bit_or;
using naga::cospan::bit_xor;
using naga::cospan::id;
using naga::cospan::maximum;
using naga::cospan::memory_scope;
using naga::cospan::minimum;
using naga::cospan::multiplies;
using naga::cospan::plus;
using naga::cospan::range;
}  // namespace sycl
