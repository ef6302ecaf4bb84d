#pragma once

// =============================================================================
// naga/cospan/wrappers.hpp — the device-side port views
// =============================================================================
//
// Input, Output: {shape, stride, data}, trivially copyable, indexed by id<R> or by a flat row-major index over shape()
// (reference: cospanmdao/include/naga/cospan/wrappers.hpp:82-154). An Output of an atomic dtype hands out atomic
// references (relaxed, agent scope) instead of plain ones. LocalArray declares an array of a group-range stage's local
// memory (LDS) and LocalView is the stage's view of it.
//
// Additions for CDNA4 kernels: vec<N>() loads / stores N consecutive elements of the last dimension in one access (the
// caller guarantees alignment of N * sizeof(T) bytes), and raw::data() exposes the device pointer for hand-written
// inner loops (nontemporal loads, buffer addressing) that a generic accessor cannot express.

#include <cstddef>
#include <cstdint>
#include <type_traits>

#include "naga/cospan/dtype.hpp"
#include "naga/cospan/vocab.hpp"

namespace naga::cospan {

namespace detail {
struct WrapperAccess;

template <int R>
CS_HD constexpr range<R> filled_range(std::size_t v) {
    range<R> r;
    for (int d = 0; d < R; ++d) r[d] = v;
    return r;
}

template <int R>
CS_HD constexpr std::size_t offset_of(const id<R>& idx, const range<R>& stride) {
    std::size_t o = 0;
    for (int d = 0; d < R; ++d) o += idx[d] * stride[d];
    return o;
}

template <int R>
CS_HD constexpr std::size_t offset_of_flat(std::size_t flat, const range<R>& shape, const range<R>& stride) {
    std::size_t o = 0;
    for (int d = R - 1; d >= 0; --d) {
        o += (flat % shape[d]) * stride[d];
        flat /= shape[d];
    }
    return o;
}
}  // namespace detail

/// N elements of T in one register group, loaded or stored with one access.
template <class T, int N>
struct alignas(sizeof(T) * N) vec {
    T v[N];
    CS_HD T&       operator[](int i) { return v[i]; }
    CS_HD const T& operator[](int i) const { return v[i]; }
};

/// Relaxed, agent-scope atomic reference (the reference's Output discipline for atomic dtypes: relaxed, device scope).
template <class T>
struct atomic_ref {
    T* p;
    __device__ T load() const { return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT); }
    __device__ void store(T v) const { __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT); }
    __device__ T fetch_add(T v) const { return __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT); }
    __device__ T fetch_sub(T v) const { return __hip_atomic_fetch_add(p, T(0) - v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT); }
    __device__ T fetch_min(T v) const { return __hip_atomic_fetch_min(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT); }
    __device__ T fetch_max(T v) const { return __hip_atomic_fetch_max(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT); }
    __device__ T fetch_or(T v) const
        requires std::is_integral_v<T>
    {
        return __hip_atomic_fetch_or(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    __device__ T exchange(T v) const { return __hip_atomic_exchange(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT); }
    __device__ bool compare_exchange_strong(T& expected, T desired) const {
        return __hip_atomic_compare_exchange_strong(p, &expected, desired, __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                                    __HIP_MEMORY_SCOPE_AGENT);
    }
    __device__ T operator+=(T v) const { return fetch_add(v) + v; }
    __device__ operator T() const { return load(); }
};

/// Input<DT, Rank>: read-only view (reference wrappers.hpp:82-107).
template <DType DT, int Rank>
    requires(!is_atomic_v<DT>)
struct Input {
    static_assert(Rank >= 1 && Rank <= 3, "cospan: rank is 1..3");
    using value_type          = cpp_type_t<DT>;
    static constexpr DType dtype = DT;
    static constexpr int   rank  = Rank;

    CS_HD const value_type& operator[](id<Rank> idx) const { return data_[detail::offset_of<Rank>(idx, stride_)]; }
    CS_HD const value_type& operator[](std::size_t flat) const {
        return data_[detail::offset_of_flat<Rank>(flat, shape_, stride_)];
    }
    /// N consecutive elements along the last dimension, from `idx`, in one load.
    template <int N>
    CS_HD vec<value_type, N> load_vec(id<Rank> idx) const {
        return *reinterpret_cast<const vec<value_type, N>*>(data_ + detail::offset_of<Rank>(idx, stride_));
    }

    CS_HD range<Rank>  shape() const { return shape_; }
    CS_HD range<Rank>  stride() const { return stride_; }
    CS_HD std::size_t  size() const { return shape_.size(); }

private:
    range<Rank>       shape_  = detail::filled_range<Rank>(0);
    range<Rank>       stride_ = detail::filled_range<Rank>(0);
    const value_type* data_   = nullptr;
    friend struct detail::WrapperAccess;
};

/// Output<DT, Rank>: read-write view; atomic references for atomic dtypes (reference wrappers.hpp:113-154).
template <DType DT, int Rank>
struct Output {
    static_assert(Rank >= 1 && Rank <= 3, "cospan: rank is 1..3");
    using value_type          = cpp_type_t<DT>;
    using atomic_type         = atomic_ref<value_type>;
    static constexpr DType dtype = DT;
    static constexpr int   rank  = Rank;

    CS_HD value_type& operator[](id<Rank> idx) const
        requires(!is_atomic_v<DT>)
    {
        return data_[detail::offset_of<Rank>(idx, stride_)];
    }
    CS_HD value_type& operator[](std::size_t flat) const
        requires(!is_atomic_v<DT>)
    {
        return data_[detail::offset_of_flat<Rank>(flat, shape_, stride_)];
    }
    CS_HD atomic_type operator[](id<Rank> idx) const
        requires(is_atomic_v<DT>)
    {
        return atomic_type{data_ + detail::offset_of<Rank>(idx, stride_)};
    }
    CS_HD atomic_type operator[](std::size_t flat) const
        requires(is_atomic_v<DT>)
    {
        return atomic_type{data_ + detail::offset_of_flat<Rank>(flat, shape_, stride_)};
    }
    template <int N>
    CS_HD void store_vec(id<Rank> idx, const vec<value_type, N>& v) const
        requires(!is_atomic_v<DT>)
    {
        *reinterpret_cast<vec<value_type, N>*>(data_ + detail::offset_of<Rank>(idx, stride_)) = v;
    }
    template <int N>
    CS_HD vec<value_type, N> load_vec(id<Rank> idx) const
        requires(!is_atomic_v<DT>)
    {
        return *reinterpret_cast<const vec<value_type, N>*>(data_ + detail::offset_of<Rank>(idx, stride_));
    }

    CS_HD range<Rank> shape() const { return shape_; }
    CS_HD range<Rank> stride() const { return stride_; }
    CS_HD std::size_t size() const { return shape_.size(); }

private:
    range<Rank> shape_  = detail::filled_range<Rank>(0);
    range<Rank> stride_ = detail::filled_range<Rank>(0);
    value_type* data_   = nullptr;
    friend struct detail::WrapperAccess;
};

// -----------------------------------------------------------------------------
// Ragged ports (reference wrappers.hpp:244-354): rank 2 only, N blocks, block i of extent_1(i) elements
// -----------------------------------------------------------------------------

/// One block of a ragged port: a rank-R view of contiguous elements.
template <class T, int R>
struct Block {
    T*       data_  = nullptr;
    range<R> shape_ = detail::filled_range<R>(0);

    CS_HD T& operator[](id<R> idx) const {
        std::size_t o = 0;
        for (int d = 0; d < R; ++d) o = o * shape_[d] + idx[d];
        return data_[o];
    }
    CS_HD T&          operator[](std::size_t flat) const { return data_[flat]; }
    CS_HD range<R>    shape() const { return shape_; }
    CS_HD std::size_t size() const { return shape_.size(); }
    CS_HD T*          get_pointer() const { return data_; }
};

/// The block of an ATOMIC ragged output: operator[] gives the relaxed, agent-scope atomic reference a dense atomic Output
/// gives; no bare reference.
template <class T, int R>
struct AtomicBlock {
    Block<T, R>                 b_{};
    CS_HD atomic_ref<T>         operator[](id<R> idx) const { return atomic_ref<T>{&b_[idx]}; }
    CS_HD atomic_ref<T>         operator[](std::size_t flat) const { return atomic_ref<T>{&b_[flat]}; }
    CS_HD range<R>              shape() const { return b_.shape(); }
    CS_HD std::size_t           size() const { return b_.size(); }
};

namespace detail {
/// Block i of a ragged port: with per-index extents (`per`, `offsets` on the device) block i holds per[i] elements from
/// offsets[i]; a dense variable read through a ragged port has no offsets and block i is its row i (`ext` elements).
template <class T>
CS_HD Block<T, 1> ragged_block(T* data, const std::uint32_t* offsets, const std::uint32_t* per, std::size_t ext, std::size_t i) {
    Block<T, 1> b;
    if (per != nullptr) {
        b.shape_[0] = per[i];
        b.data_     = data + offsets[i];
    } else {
        b.shape_[0] = ext;
        b.data_     = data + i * ext;
    }
    return b;
}
}  // namespace detail

/// RaggedInput<DT, 2>: a read-only ragged port; also reads a DENSE rank-2 variable (block i = row i).
template <DType DT, int Rank>
    requires(!is_atomic_v<DT>)
struct RaggedInput {
    static_assert(Rank == 2, "cospan: a ragged port is rank 2 only");
    using value_type              = cpp_type_t<DT>;
    static constexpr DType dtype  = DT;
    static constexpr int   rank   = Rank;
    static constexpr bool  ragged = true;

    CS_HD std::size_t size() const { return n_; }
    CS_HD Block<const value_type, 1> block(std::size_t i) const {
        return detail::ragged_block<const value_type>(data_, offsets_, per_, ext_, i);
    }

private:
    std::size_t          n_       = 0;
    const value_type*    data_    = nullptr;
    const std::uint32_t* offsets_ = nullptr;
    const std::uint32_t* per_     = nullptr;
    std::size_t          ext_     = 0;
    friend struct detail::WrapperAccess;
};

/// RaggedOutput<DT, 2>: the writable twin; with an atomic DT, block(i) is an AtomicBlock.
template <DType DT, int Rank>
struct RaggedOutput {
    static_assert(Rank == 2, "cospan: a ragged port is rank 2 only");
    using value_type              = cpp_type_t<DT>;
    static constexpr DType dtype  = DT;
    static constexpr int   rank   = Rank;
    static constexpr bool  ragged = true;

    CS_HD std::size_t size() const { return n_; }
    CS_HD Block<value_type, 1> block(std::size_t i) const
        requires(!is_atomic_v<DT>)
    {
        return detail::ragged_block<value_type>(data_, offsets_, per_, ext_, i);
    }
    CS_HD AtomicBlock<value_type, 1> block(std::size_t i) const
        requires(is_atomic_v<DT>)
    {
        return AtomicBlock<value_type, 1>{detail::ragged_block<value_type>(data_, offsets_, per_, ext_, i)};
    }

private:
    std::size_t          n_       = 0;
    value_type*          data_    = nullptr;
    const std::uint32_t* offsets_ = nullptr;
    const std::uint32_t* per_     = nullptr;
    std::size_t          ext_     = 0;
    friend struct detail::WrapperAccess;
};

static_assert(std::is_trivially_copyable_v<RaggedInput<DType::f32, 2>>);
static_assert(std::is_trivially_copyable_v<RaggedOutput<DType::f32, 2>>);

namespace detail {
/// Points a wrapper at its data (the executor's side; user code never does this).
struct WrapperAccess {
    /// A ragged port: n blocks; per-index extents and offsets on the device, or (both null) rows of `ext` elements.
    template <class W>
    static void point_ragged(W& w, void* data, std::size_t n, const std::uint32_t* offsets, const std::uint32_t* per, std::size_t ext) {
        using T     = typename W::value_type;
        w.n_        = n;
        w.data_     = static_cast<T*>(data);
        w.offsets_  = offsets;
        w.per_      = per;
        w.ext_      = ext;
    }
    template <class W>
    static void point(W& w, void* data, const std::size_t* shape, const std::size_t* stride) {
        using T = typename W::value_type;
        for (int d = 0; d < W::rank; ++d) {
            w.shape_[d]  = shape[d];
            w.stride_[d] = stride[d];
        }
        w.data_ = static_cast<T*>(data);
    }
    template <class W>
    static CS_HD auto data(const W& w) {
        return w.data_;
    }
};

template <class T>
struct is_input_wrapper : std::false_type {};
template <DType DT, int R>
struct is_input_wrapper<Input<DT, R>> : std::true_type {};
template <class T>
struct is_output_wrapper : std::false_type {};
template <DType DT, int R>
struct is_output_wrapper<Output<DT, R>> : std::true_type {};
}  // namespace detail

/// The raw device pointer of a port view, for hand-written inner loops.
namespace raw {
template <class W>
CS_HD auto data(const W& w) {
    return detail::WrapperAccess::data(w);
}
}  // namespace raw

/// LocalArray<DT, R>: one array of a group-range stage's local memory; its shape comes from cs::local_desc.
template <DType DT, int Rank>
struct LocalView {
    using value_type = cpp_type_t<DT>;
    value_type* data_;
    range<Rank> shape_;
    CS_HD value_type& operator[](id<Rank> idx) const {
        std::size_t o = 0;
        for (int d = 0; d < Rank; ++d) o = o * shape_[d] + idx[d];
        return data_[o];
    }
    CS_HD value_type& operator[](std::size_t flat) const { return data_[flat]; }
    CS_HD range<Rank> shape() const { return shape_; }
    CS_HD value_type* get_pointer() const { return data_; }
};

template <DType DT, int Rank>
struct LocalArray {
    static_assert(!is_atomic_v<DT>, "cospan: a LocalArray holds plain values");
    using value_type          = cpp_type_t<DT>;
    using view_type           = LocalView<DT, Rank>;
    static constexpr int rank = Rank;
};

}  // namespace naga::cospan
