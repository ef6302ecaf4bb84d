#pragma once

// =============================================================================
// naga/cospan/variable.hpp — shapes, initializers, variables and the executor's event tracking
// =============================================================================
//
// A variable is the data behind one or more ports (connected, promoted and tied ports share one). It holds its dtype,
// rank and shape, its device allocation, its pending initializer and the executor's dependency state: the event of its
// last writer and the events of its readers since (the reference keeps per-container read and write event lists,
// cospanmdao/include/naga/cospan/variable.hpp:98-172; a writer waits on both lists, a reader on the write list:
// component.hpp:507-539). Here an event is a HIP event recorded on one of the model's streams, and "waiting" is
// hipStreamWaitEvent on the launching stream (nothing at all when the event is on the same stream).

#include <cstddef>
#include <cstdint>
#include <cstring>
#include <functional>
#include <initializer_list>
#include <memory>
#include <optional>
#include <concepts>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <vector>

#include "cospan_hip.h"
#include "naga/cospan/dtype.hpp"
#include "naga/cospan/group_launch.hpp"
#include "naga/cospan/wrappers.hpp"

namespace naga::cospan {

// -----------------------------------------------------------------------------
// Shapes of declarations and initializers
// -----------------------------------------------------------------------------

/// cs::unspecified: every dimension of the port's rank is left open.
struct UnspecifiedTag {};
inline constexpr UnspecifiedTag unspecified{};

/// A declaration's shape: per dimension specified (an extent) or not. rank 0: the port's rank, all open.
struct ShapeSpec {
    int         rank = 0;
    std::size_t extents[3] = {0, 0, 0};
    bool        specified[3] = {false, false, false};

    ShapeSpec() = default;
    ShapeSpec(UnspecifiedTag) {}
    ShapeSpec(std::size_t e0) : rank{1} {
        extents[0]   = e0;
        specified[0] = true;
    }
    ShapeSpec(int e0) : ShapeSpec(static_cast<std::size_t>(e0)) {}
    ShapeSpec(std::initializer_list<std::size_t> dims) : rank{static_cast<int>(dims.size())} {
        int d = 0;
        for (std::size_t e : dims) {
            extents[d]     = e;
            specified[d++] = true;
        }
    }
};

/// All `R` dimensions open (cs::unspecified_shape<2>()).
template <int R>
ShapeSpec unspecified_shape() {
    ShapeSpec s;
    s.rank = R;
    return s;
}

/// What one dimension of a declaration may say (reference metadata.hpp:54): an extent, or cs::unspecified.
template <class T>
concept DimSpec = std::integral<T> || std::same_as<T, UnspecifiedTag>;

/// The shape a declaration with one spec per dimension gives (declare_output(y, "y", 4, ny, cs::unspecified)).
template <DimSpec... S>
ShapeSpec shape_of_specs(S... specs) {
    ShapeSpec out;
    out.rank = static_cast<int>(sizeof...(S));
    int  d   = 0;
    auto put = [&](auto v) {
        if constexpr (std::integral<decltype(v)>) {
            if constexpr (std::is_signed_v<decltype(v)>)
                if (v < 0) throw std::invalid_argument("naga::cospan: a declared extent is negative");
            out.extents[d]   = static_cast<std::size_t>(v);
            out.specified[d] = true;
        }
        ++d;
    };
    (put(specs), ...);
    return out;
}

/// A fully specified shape (cs::shape(1), cs::shape(n, m)).
template <class... E>
ShapeSpec shape(E... e) {
    return ShapeSpec{static_cast<std::size_t>(e)...};
}

/// The shape an initializer brings (init::fill(v, shape), init::from_host(values, shape)).
struct InitShape {
    int         rank       = 0;
    std::size_t extents[3] = {0, 0, 0};
    std::size_t elements() const {
        std::size_t n = 1;
        for (int d = 0; d < rank; ++d) n *= extents[d];
        return rank ? n : 0;
    }
};

class Shape {
public:
    Shape(std::initializer_list<std::size_t> dims) {
        for (std::size_t e : dims) s_.extents[s_.rank++] = e;
    }
    const InitShape& get() const { return s_; }

private:
    InitShape s_;
};

namespace detail {
struct VarMeta;
/// What an initializer gets: its variable (sized, allocated) and the stream to enqueue on.
struct InitCtx {
    VarMeta&      var;
    cs_hip_stream stream;
};
}  // namespace detail

/// The initial value of a variable: applied when the variable is materialized with a pending initializer.
struct InitSpec {
    std::function<void(detail::InitCtx&)> apply;
    std::optional<InitShape>              shape;
};

/// The writable view an init::custom function gets (the same access discipline as an Output).
template <DType DT, int Rank>
using Initializer = Output<DT, Rank>;

/// The host buffer an init::custom_host function fills.
template <DType DT, int Rank>
class HostInit {
public:
    using value_type = cpp_type_t<DT>;
    HostInit(const std::size_t* dims) {
        std::size_t n = 1;
        for (int d = 0; d < Rank; ++d) {
            shape_[d] = dims[d];
            n *= dims[d];
        }
        data_.resize(n);
    }
    value_type& operator[](std::size_t flat) { return data_[flat]; }
    value_type& operator[](id<Rank> idx) {
        std::size_t o = 0;
        for (int d = 0; d < Rank; ++d) o = o * shape_[d] + idx[d];
        return data_[o];
    }
    range<Rank>  shape() const { return shape_; }
    std::size_t  size() const { return data_.size(); }
    value_type*  data() { return data_.data(); }

private:
    range<Rank>             shape_;
    std::vector<value_type> data_;
};

namespace detail {

// -----------------------------------------------------------------------------
// Events and streams
// -----------------------------------------------------------------------------

/// A recorded HIP event and the executor stream it was recorded on; destroyed with its last holder.
struct Event {
    cs_hip_event  e      = nullptr;
    int           stream = 0;
    cs_hip_stream where  = nullptr;  ///< a deferred event: recorded on this stream when first needed
    Event(cs_hip_event ev, int s) : e{ev}, stream{s} {}
    Event(const Event&)            = delete;
    Event& operator=(const Event&) = delete;
    ~Event() {
        if (e) cs_hip_event_destroy(e);
    }
    /// The HIP event. A deferred event (the tail of a graph replay) is recorded on its stream now: it then covers at least
    /// the work it stands for, since that stream runs in order.
    cs_hip_event get() {
        if (!e) {
            hip_check(cs_hip_event_create(&e, 0), "event");
            hip_check(cs_hip_event_record(e, where), "event record");
        }
        return e;
    }
};
using EventPtr = std::shared_ptr<Event>;

// -----------------------------------------------------------------------------
// Ragged extents (reference variable.hpp:257-267, metadata.hpp:67-80)
// -----------------------------------------------------------------------------

/// The per-index extents of a ragged variable on the device: extents[n] and their exclusive scan offs[n + 1] (u32, as
/// the reference's device column), and the host copy of the total. Built by a ragged shape hook (cs_hip_build_column) or
/// uploaded from host counts.
struct DeviceColumn {
    std::uint32_t* extents   = nullptr;
    std::uint32_t* offs      = nullptr;
    std::uint32_t* result    = nullptr;  ///< [total, flag] of a column build
    void*          workspace = nullptr;  ///< the scan's look-back words (zeroed once; every scan leaves it zeroed)
    std::size_t    cap_n = 0, ws_bytes = 0;
    std::size_t    n     = 0;
    std::uint64_t  total = 0, version = 0;

    DeviceColumn() = default;
    DeviceColumn(const DeviceColumn&)            = delete;
    DeviceColumn& operator=(const DeviceColumn&) = delete;
    ~DeviceColumn() {
        for (void* p : {static_cast<void*>(extents), static_cast<void*>(offs), static_cast<void*>(result), workspace})
            if (p) cs_hip_free(p);
    }
    /// Room for n indices (the contents are not kept).
    void reserve(std::size_t want) {
        if (want <= cap_n && extents) return;
        for (void** p : {reinterpret_cast<void**>(&extents), reinterpret_cast<void**>(&offs), reinterpret_cast<void**>(&result)})
            if (*p) {
                cs_hip_free(*p);
                *p = nullptr;
            }
        const std::size_t c = want < 64 ? 64 : want;
        hip_check(cs_hip_malloc(reinterpret_cast<void**>(&extents), c * 4), "ragged column");
        hip_check(cs_hip_malloc(reinterpret_cast<void**>(&offs), (c + 1) * 4), "ragged column");
        hip_check(cs_hip_malloc(reinterpret_cast<void**>(&result), 16), "ragged column");
        cap_n = c;
    }
    /// A zeroed scan workspace for a column build over `count` source entries.
    void reserve_workspace(std::size_t count, cs_hip_stream s) {
        const std::size_t need = cs_hip_scan_workspace_bytes(count);
        if (need <= ws_bytes && workspace) return;
        if (workspace) cs_hip_free(workspace);
        workspace = nullptr;
        hip_check(cs_hip_malloc(&workspace, need), "ragged column workspace");
        hip_check(cs_hip_scan_workspace_init(workspace, need, s), "ragged column workspace");
        ws_bytes = need;
    }
};

}  // namespace detail

/// A host array of per-index counts for a ragged declaration (its length is N): cs::counts(v).
struct Counts {
    std::vector<std::size_t> v;
};
inline Counts counts(std::vector<std::size_t> v) { return Counts{std::move(v)}; }
/// The same as N + 1 offsets: cs::offsets(o), o[0] == 0, non-decreasing.
struct Offsets {
    std::vector<std::size_t> v;
};
inline Offsets offsets(std::vector<std::size_t> v) { return Offsets{std::move(v)}; }

/// The shape of a ragged port (N indices and one count per index): what shape_of returns for a RaggedInput /
/// RaggedOutput and what set_shape takes. The counts are on the host, or, for a hooked ragged variable, on the device
/// (`column`): they come down to the host only when asked for (counts_on_host).
struct RaggedShape {
    std::size_t                                 n     = 0;
    std::uint64_t                               total = 0;
    std::vector<std::size_t>                    counts;  ///< the host counts (empty when `column` is set)
    std::shared_ptr<const detail::DeviceColumn> column;
    bool                                        on_device() const noexcept { return column != nullptr; }
    /// The counts on the host (the device column copied down, waiting for it).
    std::vector<std::size_t> counts_on_host() const {
        if (!column) return counts;
        std::vector<std::uint32_t> e(n);
        if (n) detail::hip_check(cs_hip_memcpy_async(e.data(), column->extents, n * 4, nullptr), "ragged counts");
        detail::hip_check(cs_hip_stream_synchronize(nullptr), "ragged counts");
        return std::vector<std::size_t>(e.begin(), e.end());
    }
};

namespace detail {

// -----------------------------------------------------------------------------
// Variables
// -----------------------------------------------------------------------------

enum class IOKind : std::uint8_t { input, output };

struct VarMeta {
    std::string name;  ///< the first declaring path, for messages
    DType       dtype = DType::f32;
    int         rank  = 1;
    std::size_t dims[3]  = {0, 0, 0};
    bool        known[3] = {false, false, false};
    bool        fixed[3] = {false, false, false};  ///< specified by a definition, a tie or a hook: nobody else sets it

    std::optional<InitSpec> init;
    bool                    pending      = false;  ///< the initializer has to (re)run before the next reader
    bool                    has_producer = false;  ///< an output writes it
    bool                    has_value    = false;  ///< written at least once (by an initializer or a producer)
    bool                    scratch      = false;  ///< a declare_scratch output: zeroed at every (re)allocation

    void*       data     = nullptr;
    std::size_t capacity = 0;  ///< bytes allocated

    // ---- ragged (rank 2: dimension 1 holds one extent per index) ----
    bool                          ragged_capable = false;  ///< a RaggedInput / RaggedOutput names it
    bool                          per_index      = false;  ///< dimension 1 is per index: the variable is ragged
    bool                          counts_fixed   = false;  ///< the counts come from a definition or a hook
    bool                          hooked_ragged  = false;  ///< a counts_from / offsets_from hook owns N and the counts
    bool                          counts_on_host = false;  ///< `counts` holds them (else only `column` does)
    std::vector<std::size_t>      counts;
    std::shared_ptr<DeviceColumn> column;
    std::uint64_t                 total          = 0;
    std::uint64_t                 counts_version = 0;  ///< bumped at every change of the host counts
    std::uint64_t                 uploaded       = ~std::uint64_t{0};  ///< the counts_version `column` was uploaded from

    EventPtr              write;  ///< the last writer
    std::vector<EventPtr> reads;  ///< readers since the last write

    VarMeta() = default;
    VarMeta(const VarMeta&)            = delete;
    VarMeta& operator=(const VarMeta&) = delete;
    ~VarMeta() {
        if (data) cs_hip_free(data);
    }

    bool shape_known() const {
        if (per_index) return known[0] && (counts_on_host || column != nullptr);
        for (int d = 0; d < rank; ++d)
            if (!known[d]) return false;
        return true;
    }
    std::size_t elements() const {
        if (per_index) return static_cast<std::size_t>(total);
        std::size_t n = 1;
        for (int d = 0; d < rank; ++d) n *= dims[d];
        return n;
    }
    /// Makes the variable ragged with host counts (N = counts.size()).
    void set_host_counts(std::vector<std::size_t> c, const std::string& who) {
        if (rank != 2) throw std::invalid_argument("naga::cospan: " + who + ": '" + name + "' is not rank 2; a ragged variable is");
        std::uint64_t t = 0;
        for (std::size_t e : c) t += e;
        if (t > 0xffffffffull)
            throw std::invalid_argument("naga::cospan: " + who + ": '" + name + "': ragged totals are limited to 2^32 - 1 elements");
        if (fixed[0] && known[0] && dims[0] != c.size())
            throw std::invalid_argument("naga::cospan: " + who + ": dimension 0 of '" + name + "' is fixed at " + std::to_string(dims[0]) +
                                        "; counts of length " + std::to_string(c.size()) + " do not fit");
        const bool same = per_index && counts_on_host && counts == c;
        per_index      = true;
        counts_on_host = true;
        dims[0]        = c.size();
        dims[1]        = 0;
        known[0] = known[1] = true;
        total               = t;
        if (!same) {
            counts = std::move(c);
            ++counts_version;
            column.reset();
        }
    }
    /// Makes the variable ragged with the counts of a device column (a hook's or another variable's).
    void set_device_counts(std::shared_ptr<DeviceColumn> col) {
        per_index      = true;
        counts_on_host = false;
        counts.clear();
        dims[0]  = col->n;
        dims[1]  = 0;
        known[0] = known[1] = true;
        total               = col->total;
        column              = std::move(col);
        uploaded            = counts_version;
    }
    /// The device column of a ragged variable with host counts: uploaded when the counts changed (waits for the copy).
    void ensure_device_counts(cs_hip_stream s) {
        if (!per_index || !counts_on_host || (column && uploaded == counts_version)) return;
        const std::size_t          n = counts.size();
        std::vector<std::uint32_t> e(n), o(n + 1);
        std::uint64_t              run = 0;
        for (std::size_t i = 0; i < n; ++i) {
            e[i] = static_cast<std::uint32_t>(counts[i]);
            o[i] = static_cast<std::uint32_t>(run);
            run += counts[i];
        }
        o[n] = static_cast<std::uint32_t>(run);
        // a new column: kernels of earlier forwards may still read the old one (it lives until its last holder lets go)
        column = std::make_shared<DeviceColumn>();
        column->reserve(n);
        column->n     = n;
        column->total = run;
        if (n) hip_check(cs_hip_memcpy_async(column->extents, e.data(), n * 4, s), "ragged counts upload");
        hip_check(cs_hip_memcpy_async(column->offs, o.data(), (n + 1) * 4, s), "ragged counts upload");
        hip_check(cs_hip_stream_synchronize(s), "ragged counts upload");
        uploaded = counts_version;
    }
    std::string counts_text() const {
        std::string t = "[";
        for (std::size_t i = 0; i < counts.size(); ++i) t += (i ? "," : "") + std::to_string(counts[i]);
        return t + "]";
    }
    std::size_t bytes() const { return elements() * dtype_size(dtype); }
    std::string shape_text() const {
        if (per_index) return "(" + std::to_string(dims[0]) + ", per index)";
        std::string s = "(";
        for (int d = 0; d < rank; ++d) s += (d ? ", " : "") + (known[d] ? std::to_string(dims[d]) : std::string{"?"});
        return s + ")";
    }
    /// Sets dimension d; a fixed dimension may only be set to its own value.
    void set_dim(int d, std::size_t e, const std::string& who) {
        if (fixed[d] && known[d] && dims[d] != e)
            throw std::invalid_argument("naga::cospan: " + who + ": dimension " + std::to_string(d) + " of '" + name + "' is fixed at " +
                                        std::to_string(dims[d]) + "; it cannot become " + std::to_string(e));
        dims[d]  = e;
        known[d] = true;
    }
    /// Waits (on the host) for every event, then frees the allocation.
    void release() {
        if (write) cs_hip_event_synchronize(write->get());
        for (auto& r : reads) cs_hip_event_synchronize(r->get());
        if (data) hip_check(cs_hip_free(data), "free of '" + name + "'");
        data     = nullptr;
        capacity = 0;
    }
    /// Allocation of at least bytes() (the old contents are not kept: sizes change only at materialization). True when
    /// a new allocation was made.
    bool ensure_allocation() {
        const std::size_t need = bytes();
        if (need <= capacity && data) return false;
        release();
        void* p = nullptr;
        hip_check(cs_hip_malloc(&p, need < 256 ? 256 : need), "allocation of '" + name + "' (" + std::to_string(need) + " bytes)");
        data     = p;
        capacity = need < 256 ? 256 : need;
        return true;
    }
};

/// A port or an alias: a path that names a variable.
struct Instance {
    std::string              path;
    IOKind                   kind  = IOKind::input;
    DType                    dtype = DType::f32;
    int                      rank  = 1;
    bool                     ragged = false;  ///< a RaggedInput / RaggedOutput port (or an alias of one)
    std::shared_ptr<VarMeta> var;
};

/// The row-major strides of the variable's shape.
inline void row_major_strides(const VarMeta& v, std::size_t* stride) {
    std::size_t s = 1;
    for (int d = v.rank - 1; d >= 0; --d) {
        stride[d] = s;
        s *= v.dims[d];
    }
}

/// The byte pattern of `value` as `dt` (fills convert like the reference's static_cast in init::fill).
template <class T>
std::vector<std::byte> converted_bytes(DType dt, T value) {
    std::vector<std::byte> b(dtype_size(dt));
    auto put = [&](auto v) { std::memcpy(b.data(), &v, sizeof v); };
    switch (value_dtype(dt)) {
        case DType::f32: put(static_cast<float>(value)); break;
        case DType::f64: put(static_cast<double>(value)); break;
        case DType::i32: put(static_cast<std::int32_t>(value)); break;
        case DType::u32: put(static_cast<std::uint32_t>(value)); break;
        case DType::i64: put(static_cast<std::int64_t>(value)); break;
        case DType::u64: put(static_cast<std::uint64_t>(value)); break;
        case DType::u8: put(static_cast<std::uint8_t>(value)); break;
        case DType::bf16: put(bf16::from_float(static_cast<float>(value)).bits); break;
        default: throw std::invalid_argument("naga::cospan: init::fill cannot write an e4m3 value; use init::custom");
    }
    return b;
}

/// Fills `count` elements of `elem` bytes with a pattern, with kernels (a graph-safe fill) when the pattern allows.
inline void fill_pattern(void* data, const std::vector<std::byte>& pat, std::size_t count, cs_hip_stream s, const std::string& who) {
    if (count == 0) return;
    const std::size_t elem = pat.size();
    bool              uniform = true;
    for (std::byte b : pat) uniform = uniform && b == pat[0];
    if (uniform) {
        hip_check(cs_hip_fill_u8_async(data, static_cast<std::uint8_t>(pat[0]), count * elem, s), who);
        return;
    }
    if (elem == 4) {
        std::uint32_t v;
        std::memcpy(&v, pat.data(), 4);
        hip_check(cs_hip_fill_u32_async(data, v, count, s), who);
        return;
    }
    // 2- and 8-byte patterns: from a pinned host copy
    std::vector<std::byte> host(count * elem);
    for (std::size_t i = 0; i < count; ++i) std::memcpy(host.data() + i * elem, pat.data(), elem);
    void* pinned = nullptr;
    hip_check(cs_hip_host_malloc(&pinned, host.size()), who);
    std::memcpy(pinned, host.data(), host.size());
    hip_check(cs_hip_memcpy_async(data, pinned, host.size(), s), who);
    hip_check(cs_hip_stream_synchronize(s), who);
    cs_hip_host_free(pinned);
}

}  // namespace detail

// -----------------------------------------------------------------------------
// cs::init (reference init.hpp:203-329)
// -----------------------------------------------------------------------------
namespace init {

namespace impl {
template <class T>
InitSpec fill_impl(T value, std::optional<InitShape> shape) {
    return InitSpec{[value](detail::InitCtx& c) {
                        const std::vector<std::byte> pat = detail::converted_bytes(c.var.dtype, value);
                        detail::fill_pattern(c.var.data, pat, c.var.elements(), c.stream, "init::fill of '" + c.var.name + "'");
                    },
                    std::move(shape)};
}
}  // namespace impl

template <class T>
InitSpec fill(T value) {
    return impl::fill_impl(value, std::nullopt);
}
template <class T>
InitSpec fill(T value, const Shape& shape) {
    return impl::fill_impl(value, shape.get());
}
/// init::fill(value, {extents}) as the reference writes it (init::fill(1.0f, {1})).
template <class T>
InitSpec fill(T value, std::initializer_list<std::size_t> dims) {
    return impl::fill_impl(value, Shape(dims).get());
}
inline InitSpec zero() { return impl::fill_impl(0, std::nullopt); }
inline InitSpec zero(const Shape& shape) { return impl::fill_impl(0, shape.get()); }

namespace impl {
template <class T>
InitSpec from_host_impl(std::vector<T> values, InitShape shape) {
    return InitSpec{[values = std::move(values)](detail::InitCtx& c) {
                        const std::size_t n = c.var.elements();
                        if (values.size() != n)
                            throw std::invalid_argument("naga::cospan: init::from_host has " + std::to_string(values.size()) +
                                                        " values but '" + c.var.name + "' has " + std::to_string(n) + " elements");
                        const std::size_t elem = dtype_size(c.var.dtype);
                        std::vector<std::byte> host(n * elem);
                        for (std::size_t i = 0; i < n; ++i) {
                            const std::vector<std::byte> b = detail::converted_bytes(c.var.dtype, values[i]);
                            std::memcpy(host.data() + i * elem, b.data(), elem);
                        }
                        detail::hip_check(cs_hip_memcpy_async(c.var.data, host.data(), host.size(), c.stream),
                                          "init::from_host of '" + c.var.name + "'");
                        // the host vector dies here: the copy must have read it
                        detail::hip_check(cs_hip_stream_synchronize(c.stream), "init::from_host of '" + c.var.name + "'");
                    },
                    shape};
}
}  // namespace impl

template <class T>
InitSpec from_host(std::vector<T> values) {
    InitShape s;
    s.rank       = 1;
    s.extents[0] = values.size();
    return impl::from_host_impl(std::move(values), s);
}
template <class T>
InitSpec from_host(std::vector<T> values, const Shape& shape) {
    if (shape.get().elements() != values.size())
        throw std::invalid_argument("naga::cospan: init::from_host has " + std::to_string(values.size()) + " values but its shape has " +
                                    std::to_string(shape.get().elements()) + " elements");
    return impl::from_host_impl(std::move(values), shape.get());
}

/// A device initializer: fn enqueues work on the stream that writes the Initializer view (e.g. a hand-written kernel).
template <DType DT, int Rank>
using CustomFn = std::function<void(cs_hip_stream, Initializer<DT, Rank>)>;

template <DType DT, int Rank>
InitSpec custom(CustomFn<DT, Rank> fn, const Shape& shape) {
    if (shape.get().rank != Rank)
        throw std::invalid_argument("naga::cospan: init::custom<..., " + std::to_string(Rank) + "> was given a shape of rank " +
                                    std::to_string(shape.get().rank));
    return InitSpec{[fn = std::move(fn)](detail::InitCtx& c) {
                        if (!dtype_compatible(c.var.dtype, DT) || c.var.rank != Rank)
                            throw std::invalid_argument("naga::cospan: init::custom for '" + c.var.name +
                                                        "' was declared with the wrong dtype or rank");
                        Initializer<DT, Rank> w;
                        std::size_t           stride[3];
                        detail::row_major_strides(c.var, stride);
                        detail::WrapperAccess::point(w, c.var.data, c.var.dims, stride);
                        fn(c.stream, w);
                    },
                    shape.get()};
}
template <DType DT, int Rank>
InitSpec custom(CustomFn<DT, Rank> fn) {
    return InitSpec{[fn = std::move(fn)](detail::InitCtx& c) {
                        Initializer<DT, Rank> w;
                        std::size_t           stride[3];
                        detail::row_major_strides(c.var, stride);
                        detail::WrapperAccess::point(w, c.var.data, c.var.dims, stride);
                        fn(c.stream, w);
                    },
                    std::nullopt};
}

template <DType DT, int Rank>
InitSpec custom_host(std::function<void(HostInit<DT, Rank>&)> fn, const Shape& shape) {
    return InitSpec{[fn = std::move(fn)](detail::InitCtx& c) {
                        HostInit<DT, Rank> h(c.var.dims);
                        fn(h);
                        detail::hip_check(cs_hip_memcpy_async(c.var.data, h.data(), h.size() * sizeof(cpp_type_t<DT>), c.stream),
                                          "init::custom_host of '" + c.var.name + "'");
                        detail::hip_check(cs_hip_stream_synchronize(c.stream), "init::custom_host");
                    },
                    shape.get()};
}

}  // namespace init
}  // namespace naga::cospan
