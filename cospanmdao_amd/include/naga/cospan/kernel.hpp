#pragma once

// =============================================================================
// naga/cospan/kernel.hpp — the fused-group kernel: stage ABI, tickets, GroupItem, LocalMem, group functions
// =============================================================================
//
// The reference assembles a fused group at RUN TIME: one generic kernel calls a placeholder that AdaptiveCpp's JIT
// replaces with an indexed call sequence of per-stage functions (cospanmdao/include/naga/cospan/kernel.hpp:911-1004).
// Here the assembly happens at COMPILE TIME: a fused group's stage types are known to the C++ type system
// (FusedGroup<..., Items...>), so hipcc instantiates ONE kernel per group,
//
//     template <int MaxThreads, class... X> __global__ void group_kernel(GroupArgs<X...>)
//
// whose body locates the work-group (ticket or hardware id -> segment and group coordinate), runs every stage
// X_0 .. X_n-1 in order as a fold over an index sequence (each stage_run<X_k> guarded by its segment and own range, as
// stage_fn<KB> is in the reference, kernel.hpp:919-950) and signals completion. There are no indirect calls and no
// device-side argument arrays: the geometry (one entry per stage, the segment table) and every stage's arguments
// (its Inputs and Outputs views, its units, its LocalMem shapes) travel as kernel arguments, so a launch needs no
// host-to-device upload and can be captured in a graph as it is.
//
// Ticket mode (a group with cs::group_waits_enabled or a cs::global_barrier): the state is a u32 array holding the
// ticket counter, the finished counter, S segment-done counters and per-XCD write-back counters, each on a 256-byte line
// of its own, then T per-group done flags. Each work-group's leader takes the next ticket (so a group only ever waits on
// groups that have started: no co-residency needed, reference kernel.hpp:197-205), waits are agent-scope acquire spins by
// the leader plus a barrier, completion is a barrier, the group's stores published by an L2 write-back that the
// work-groups of one XCD share, and the leader's atomics. The last group to finish zeroes the whole state, so the next
// launch (or graph replay) starts from zero without a memset.

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <limits>
#include <type_traits>
#include <utility>

#include "cospan_hip.h"
#include "naga/cospan/device.hpp"
#include "naga/cospan/dtype.hpp"
#include "naga/cospan/vocab.hpp"
#include "naga/cospan/wrappers.hpp"

namespace naga::cospan {

/// Local memory of a segment: partitioned (default, a slice per stage) or shared (one block, the max).
namespace local_mem {
struct partitioned {};
struct shared {};
}  // namespace local_mem

/// Range of a segment: own (default, a stage runs for its own units only) or whole (every stage runs every unit).
namespace range_mode {
struct own {};
struct whole {};
}  // namespace range_mode

namespace detail {

enum class LocalMode : std::uint8_t { partitioned = 0, shared = 1 };
enum class RangeMode : std::uint8_t { own = 0, whole = 1 };
enum class Boundary : std::uint8_t { none = 0, global_barrier = 1 };
enum class Split : std::uint8_t { flatten = 0, pad = 1, padded = 2 };

inline constexpr int kMaxStages = 16;

/// One entry per stage (reference kernel.hpp:109-120).
struct StageGeometry {
    std::uint32_t segment      = 0;
    std::uint32_t local_offset = 0;  ///< bytes into the launch's LDS block
    std::uint32_t local_bytes  = 0;
    std::uint32_t range_mode   = 0;
};

/// One entry per segment (reference kernel.hpp:127-133): its block of groups as a rank-3 range, the logical count, the first
/// ticket, the offset along the last dimension (pad split) and the boundary before it.
struct SegmentInfo {
    std::uint64_t range[3] = {0, 0, 0};
    std::uint64_t count    = 0;
    std::uint64_t first    = 0;
    std::uint64_t offset   = 0;
    std::uint64_t boundary = 0;
};

template <int N>
struct Geometry {
    std::uint32_t* ticket_state = nullptr;  ///< null: no ticket mode
    std::uint32_t* fault        = nullptr;  ///< the device's fault word (cs_hip_fault_word): where a wait that gave up reports
    std::uint64_t  total_groups = 0;
    std::uint32_t  segments     = 0;
    std::uint32_t  split        = 0;
    std::uint32_t  seg_lo       = 0;  ///< the segments this launch runs: [seg_lo, seg_hi) (all of them but for a group
    std::uint32_t  seg_hi       = 0;  ///< whose global barriers are launch boundaries)
    std::uint32_t  spin_limit   = CS_HIP_DEFAULT_SPIN_LIMIT;  ///< polls per wait before it gives up
    std::uint32_t  unused       = 0;  ///< keeps the struct free of padding: a launch's signature compares its bytes
    StageGeometry  stage[N];
    SegmentInfo    seg[N];
};

// ---- rank-R values in the trailing physical dimensions ------------------------------------------------------------------
template <int R>
CS_HD inline id<R> tail_id(std::size_t c0, std::size_t c1, std::size_t c2) {
    if constexpr (R == 1)
        return id<1>{c2};
    else if constexpr (R == 2)
        return id<2>{c1, c2};
    else
        return id<3>{c0, c1, c2};
}
template <int R>
CS_HD inline range<R> tail_range(std::size_t c0, std::size_t c1, std::size_t c2) {
    if constexpr (R == 1)
        return range<1>{c2};
    else if constexpr (R == 2)
        return range<2>{c1, c2};
    else
        return range<3>{c0, c1, c2};
}
template <int R>
CS_HD inline range<3> lift_range(const range<R>& r) {
    if constexpr (R == 1)
        return range<3>{1, 1, r[0]};
    else if constexpr (R == 2)
        return range<3>{1, r[0], r[1]};
    else
        return r;
}
template <int R>
CS_HD inline id<3> lift_id(const id<R>& r) {
    if constexpr (R == 1)
        return id<3>{0, 0, r[0]};
    else if constexpr (R == 2)
        return id<3>{0, r[0], r[1]};
    else
        return r;
}

/// Where a work-group sits logically.
struct LogicalGroup {
    std::uint32_t segment = 0;
    std::uint64_t coord[3] = {0, 0, 0};
    bool          valid   = false;
};

__device__ inline std::uint64_t segment_linear(const SegmentInfo& s, const std::uint64_t* c) {
    return (c[0] * s.range[1] + c[1]) * s.range[2] + c[2];
}

/// The segment that holds group number `lin` (a ticket, or a hardware id flattened along x) and the group's coordinate in
/// it. Numbers run segment after segment (segment s owns [first_s, first_s + count_s), the segments back to back), and
/// inside a segment over its rank-3 range with the last coordinate fastest. A number at or past `total` is no group.
__device__ inline LogicalGroup locate_linear(const SegmentInfo* tbl, std::uint32_t segments, std::uint64_t lin, std::uint64_t total) {
    LogicalGroup where;
    if (lin >= total) return where;
    // the owner is the segment before the first one that starts past `lin` (an empty segment starts where the next one
    // does, so it never owns a number)
    std::uint32_t seg = segments - 1;
    for (std::uint32_t s = 1; s < segments; ++s)
        if (lin < tbl[s].first) {
            seg = s - 1;
            break;
        }
    const SegmentInfo&  info  = tbl[seg];
    const std::uint64_t rel   = lin - info.first;
    const std::uint64_t plane = info.range[1] * info.range[2];
    const std::uint64_t c0    = rel / plane;
    const std::uint64_t rest  = rel - c0 * plane;
    const std::uint64_t c1    = rest / info.range[2];
    where.segment  = seg;
    where.coord[0] = c0;
    where.coord[1] = c1;
    where.coord[2] = rest - c1 * info.range[2];
    where.valid    = true;
    return where;
}

/// Hardware-direct mapping (no tickets), per split strategy (reference kernel.hpp:266-288). Physical dimension 2 is
/// HIP's x, 1 is y, 0 is z.
/// Segments [lo, hi) are the ones this launch runs; the physical ids count from segment lo.
__device__ inline LogicalGroup locate_physical(std::uint32_t split, const SegmentInfo* tbl, std::uint32_t lo, std::uint32_t hi,
                                               std::uint64_t total) {
    const std::uint64_t d0 = blockIdx.z, d1 = blockIdx.y, d2 = blockIdx.x;
    if (split == static_cast<std::uint32_t>(Split::flatten)) return locate_linear(tbl, hi, tbl[lo].first + d2, total);
    LogicalGroup r;
    if (split == static_cast<std::uint32_t>(Split::pad)) {
        const std::uint64_t x = tbl[lo].offset + d2;
        for (std::uint32_t s = lo; s < hi; ++s) {
            if (x >= tbl[s].offset && x < tbl[s].offset + tbl[s].range[2]) {
                if (d0 < tbl[s].range[0] && d1 < tbl[s].range[1]) {
                    r.segment  = s;
                    r.coord[0] = d0;
                    r.coord[1] = d1;
                    r.coord[2] = x - tbl[s].offset;
                    r.valid    = true;
                }
                break;
            }
        }
    } else {
        const std::uint64_t s = lo + d0;
        if (s < hi && d1 < tbl[s].range[1] && d2 < tbl[s].range[2]) {
            r.segment  = static_cast<std::uint32_t>(s);
            r.coord[0] = 0;
            r.coord[1] = d1;
            r.coord[2] = d2;
            r.valid    = true;
        }
    }
    return r;
}

// ---- ticket state ---------------------------------------------------------------------------------------------------
// Word 0 the ticket counter, word kStateLine the finished counter, word (2 + s) kStateLine segment s's finished count,
// then per XCD x the started and completed counts of its shared L2 write-backs (words (2 + S + 2x) kStateLine and
// (3 + S + 2x) kStateLine), then one done flag per ticket. Every work-group makes one atomic on each of the three
// counters it touches: on one line they were served one after another by the same atomic unit (the reference's fused NS
// residual at 4096^2, 16384 groups: 1.94 ms with them packed, 0.85 ms with a line each).
inline constexpr std::size_t   kStateLine = 64;  // u32 words: 256 bytes
inline constexpr std::uint32_t kXcds      = 8;   // XCDs of an MI355X (one L2 each)
CS_HD inline std::size_t finished_word() { return kStateLine; }
CS_HD inline std::size_t segment_word(std::size_t s) { return (2 + s) * kStateLine; }
CS_HD inline std::size_t write_back_word(std::size_t segments, std::uint32_t xcd, std::uint32_t completed) {
    return (2 + segments + 2 * xcd + completed) * kStateLine;
}
CS_HD inline std::size_t flag_word(std::size_t segments, std::uint64_t ticket) { return (2 + segments + 2 * kXcds) * kStateLine + ticket; }
CS_HD inline std::size_t ticket_words(std::size_t segments, std::uint64_t total) { return flag_word(segments, total); }

/// A spin-wait load: relaxed at agent scope (coherent across the XCDs' L2s without invalidating them on every poll); the
/// waits below order what follows with ONE acquire fence once the condition holds, made by the leader before the
/// work-group barrier (the barrier carries it to the other waves: one cache invalidation per work-group, not per wave).
__device__ inline std::uint32_t poll_load(std::uint32_t* p) { return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT); }

/// Waits only ever target groups with earlier tickets, which have started and finish without waiting on later ones, so
/// they end. Each wait still polls at most `limit` times (Geometry::spin_limit, set from cs_hip_spin_limit at launch): one
/// that gives up reports CS_HIP_FAULT_TICKET_WAIT to the fault word and lets its group run on, so corrupted ticket state
/// can neither hang the device nor pass unnoticed.

/// COLLECTIVE: wait for group `c` of segment `s` (reference kernel.hpp:215-224).
__device__ inline void ticket_wait_group(std::uint32_t* ts, std::uint32_t segments, const SegmentInfo* tbl, std::uint32_t s,
                                         const std::uint64_t* c, std::uint32_t* fault, std::uint32_t limit) {
    if (threadIdx.x + threadIdx.y + threadIdx.z == 0) {
        std::uint32_t* const flag = &ts[flag_word(segments, tbl[s].first + segment_linear(tbl[s], c))];
        if (!bounded::poll([&] { return poll_load(flag) != 0; }, limit)) bounded::report(fault, CS_HIP_FAULT_TICKET_WAIT);
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    }
    __syncthreads();
}

/// COLLECTIVE: wait until every group of segments 0 .. upto-1 has finished (reference kernel.hpp:227-236).
__device__ inline void ticket_wait_segments(std::uint32_t* ts, const SegmentInfo* tbl, std::uint32_t upto, std::uint32_t* fault,
                                            std::uint32_t limit) {
    if (threadIdx.x + threadIdx.y + threadIdx.z == 0) {
        for (std::uint32_t i = 0; i < upto; ++i) {
            std::uint32_t* const done = &ts[segment_word(i)];
            if (!bounded::poll([&] { return static_cast<std::uint64_t>(poll_load(done)) == tbl[i].count; }, limit)) {
                bounded::report(fault, CS_HIP_FAULT_TICKET_WAIT);
                break;
            }
        }
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    }
    __syncthreads();
}

/// One poll of a shared write-back (publish_stores) by a work-group that read s0 = started after its stores had reached
/// its XCD's L2, with d = completed and s = started read now: covered once a write-back numbered above s0 has completed
/// (it started after those stores arrived), issue when none is in flight (the group then claims number s + 1 by a
/// compare-and-swap on started), else wait. Host-and-device: the CPU simulation of the protocol runs this function.
enum class WriteBackStep : std::uint8_t { covered, wait, issue };
CS_HD inline WriteBackStep write_back_step(std::uint32_t s0, std::uint32_t d, std::uint32_t s) {
    if (d > s0) return WriteBackStep::covered;
    return s == d ? WriteBackStep::issue : WriteBackStep::wait;
}

/// The XCD this wave runs on (hardware register XCC_ID; work-group i of a launch lands on XCD i % 8).
__device__ inline std::uint32_t xcd_id() { return __builtin_amdgcn_s_getreg((20) | (0 << 6) | ((4 - 1) << 11)); }

/// Publishes the work-group's stores to the other XCDs with a write-back of its XCD's L2 that it may share: a write-back
/// of an L2 takes every dirty line in it, so one that STARTS after this group's stores have reached the L2 publishes them
/// too, whichever work-group issues it. Per XCD, write-backs are numbered (`started`, then `completed` once done; one in
/// flight at a time). The group reads s0 = started after its stores are in the L2 (every wave waited for them before
/// the barrier), so write-back s0 + 1 starts after them; it returns once `completed` > s0, issuing that write-back itself
/// when none is in flight. A release fence per work-group serialised them instead: the fused NS residual at 4096^2 with
/// 16384 groups took 0.85 ms, 0.47 ms when only one group in eight wrote back (a timing probe, not a protocol). An XCD
/// id past the table, or a poll that gives up, falls back to the group's own write-back (false: the caller reports).
/// `wb`: the 2 kXcds counters (XCD x's started at wb[2 x kStateLine], completed at wb[(2 x + 1) kStateLine]), zero at
/// the launch. Every counter update of the call has been performed when it returns, so the caller's later atomics (its
/// finished count, after which the last work-group zeroes the counters) come after them. Leader only.
__device__ inline bool publish_stores(std::uint32_t* wb, std::uint32_t limit) {
    const std::uint32_t x = xcd_id();
    if (x >= kXcds) {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
        return true;
    }
    std::uint32_t* const started   = &wb[2 * x * kStateLine];
    std::uint32_t* const completed = &wb[(2 * x + 1) * kStateLine];
    const std::uint32_t  s0        = poll_load(started);
    const bool           done      = bounded::poll(
        [&] {
            const std::uint32_t d    = poll_load(completed);
            std::uint32_t       s    = d > s0 ? d : poll_load(started);  // covered needs no second load
            const WriteBackStep step = write_back_step(s0, d, s);
            if (step != WriteBackStep::issue) return step == WriteBackStep::covered;  // wait: one in flight may predate the stores
            if (!__hip_atomic_compare_exchange_strong(started, &s, s + 1, __ATOMIC_RELAXED, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))
                return false;
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");  // the write-back, ordered before the count that follows
            __hip_atomic_exchange(completed, s + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            return true;  // s == d <= s0 <= s: this is write-back s0 + 1
        },
        limit);
    if (!done) __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
    asm volatile("s_waitcnt vmcnt(0)" ::: "memory");  // the counter updates above are performed
    return done;
}

__device__ inline std::uint32_t local_linear_id() { return threadIdx.x + blockDim.x * (threadIdx.y + blockDim.y * threadIdx.z); }
__device__ inline std::uint32_t local_linear_range() { return blockDim.x * blockDim.y * blockDim.z; }

struct ItemFixup;
}  // namespace detail

template <int R>
struct GroupItem;
template <class... Arrays>
struct LocalMem;

// -----------------------------------------------------------------------------
// Group<R>: the work-group a stage belongs to, with rank-R ids (reference kernel.hpp:317-360)
// -----------------------------------------------------------------------------
template <int R>
struct Group {
    __device__ id<R>       get_group_id() const { return id_; }
    __device__ std::size_t get_group_id(int d) const { return id_[d]; }
    __device__ range<R>    get_group_range() const { return range_; }
    __device__ std::size_t get_group_range(int d) const { return range_[d]; }
    __device__ std::size_t get_group_linear_id() const {
        std::size_t l = 0;
        for (int d = 0; d < R; ++d) l = l * range_[d] + id_[d];
        return l;
    }
    __device__ std::size_t get_group_linear_range() const { return range_.size(); }
    __device__ id<R>       get_local_id() const { return detail::tail_id<R>(threadIdx.z, threadIdx.y, threadIdx.x); }
    __device__ std::size_t get_local_id(int d) const { return get_local_id()[d]; }
    __device__ range<R>    get_local_range() const { return detail::tail_range<R>(blockDim.z, blockDim.y, blockDim.x); }
    __device__ std::size_t get_local_range(int d) const { return get_local_range()[d]; }
    __device__ std::size_t get_local_linear_id() const { return detail::local_linear_id(); }
    __device__ std::size_t get_local_linear_range() const { return detail::local_linear_range(); }
    __device__ bool        leader() const { return detail::local_linear_id() == 0; }

    id<R>    id_;
    range<R> range_;
};

// -----------------------------------------------------------------------------
// sub_group: one wavefront of 64 lanes (reference kernel.hpp:306 aliases sycl::sub_group)
// -----------------------------------------------------------------------------
/// The work-items of one wave, in local linear order: wave k of a work-group holds local linear ids 64k .. 64k+63 (the
/// last wave of a group whose size is not a multiple of 64 holds fewer).
struct sub_group {
    using id_type        = id<1>;
    using range_type     = range<1>;
    using linear_id_type = std::uint32_t;

    __device__ id<1>          get_local_id() const { return id<1>{static_cast<std::size_t>(get_local_linear_id())}; }
    __device__ linear_id_type get_local_linear_id() const { return detail::local_linear_id() % wave::kSize; }
    __device__ range<1>       get_local_range() const { return range<1>{static_cast<std::size_t>(get_local_linear_range())}; }
    __device__ linear_id_type get_local_linear_range() const {
        const std::uint32_t rest = detail::local_linear_range() - get_group_linear_id() * wave::kSize;
        return rest < wave::kSize ? rest : wave::kSize;
    }
    __device__ range<1>       get_max_local_range() const { return range<1>{static_cast<std::size_t>(wave::kSize)}; }
    __device__ id<1>          get_group_id() const { return id<1>{static_cast<std::size_t>(get_group_linear_id())}; }
    __device__ linear_id_type get_group_linear_id() const { return detail::local_linear_id() / wave::kSize; }
    __device__ range<1>       get_group_range() const {
        return range<1>{static_cast<std::size_t>((detail::local_linear_range() + wave::kSize - 1) / wave::kSize)};
    }
    __device__ bool leader() const { return get_local_linear_id() == 0; }
};

// -----------------------------------------------------------------------------
// GroupItem<R>: what a group-range body receives (reference kernel.hpp:364-446)
// -----------------------------------------------------------------------------
template <int R>
struct GroupItem {
    __device__ id<R>       get_local_id() const { return detail::tail_id<R>(threadIdx.z, threadIdx.y, threadIdx.x); }
    __device__ std::size_t get_local_id(int d) const { return get_local_id()[d]; }
    __device__ range<R>    get_local_range() const { return detail::tail_range<R>(blockDim.z, blockDim.y, blockDim.x); }
    __device__ std::size_t get_local_range(int d) const { return get_local_range()[d]; }
    __device__ std::size_t get_local_linear_id() const { return detail::local_linear_id(); }

    __device__ id<R>       get_group_id() const { return group_id_; }
    __device__ std::size_t get_group(int d) const { return group_id_[d]; }
    __device__ range<R>    get_group_range() const { return group_range_; }
    __device__ std::size_t get_group_range(int d) const { return group_range_[d]; }
    __device__ std::size_t get_group_linear_id() const {
        std::size_t l = 0;
        for (int d = 0; d < R; ++d) l = l * group_range_[d] + group_id_[d];
        return l;
    }
    __device__ id<R> get_global_id() const {
        id<R>          g;
        const range<R> lr = get_local_range();
        const id<R>    li = get_local_id();
        for (int d = 0; d < R; ++d) g[d] = group_id_[d] * lr[d] + li[d];
        return g;
    }
    __device__ std::size_t get_global_id(int d) const { return get_global_id()[d]; }
    __device__ range<R>    get_global_range() const {
        range<R>       g;
        const range<R> lr = get_local_range();
        for (int d = 0; d < R; ++d) g[d] = group_range_[d] * lr[d];
        return g;
    }
    __device__ std::size_t get_global_linear_id() const {
        const id<R>    g = get_global_id();
        const range<R> r = get_global_range();
        std::size_t    l = 0;
        for (int d = 0; d < R; ++d) l = l * r[d] + g[d];
        return l;
    }

    __device__ std::size_t get_segment() const { return segment_; }
    __device__ std::size_t get_segment_count() const { return segments_; }
    __device__ range<R>    get_segment_group_count(std::size_t s) const {
        return detail::tail_range<R>(table_[s].range[0], table_[s].range[1], table_[s].range[2]);
    }

    /// COLLECTIVE: wait until group `g` of the EARLIER segment `s` has finished all its stages. Without wait state (a
    /// group whose global barriers are launch boundaries) an earlier segment behind a barrier ran in an earlier launch and
    /// has finished; waits between groups of one launch need cs::group_waits_enabled.
    __device__ void wait_group(std::size_t s, const id<R>& g) const {
        if (ticket_state_ == nullptr) return;
        const id<3>         l = detail::lift_id<R>(g);
        const std::uint64_t c[3] = {l[0], l[1], l[2]};
        detail::ticket_wait_group(ticket_state_, segments_, table_, static_cast<std::uint32_t>(s), c, fault_, spin_limit_);
    }
    /// COLLECTIVE: wait until every group of every earlier segment has finished (what cs::global_barrier() does).
    __device__ void wait_earlier_segments() const {
        if (ticket_state_ != nullptr) detail::ticket_wait_segments(ticket_state_, table_, segment_, fault_, spin_limit_);
    }

    /// The poll bound of this launch's waits; a stage with a bounded wait of its own polls at most this often and, when it
    /// gives up, reports through report_fault (CS_HIP_FAULT_* bits, read by Model::wait).
    __device__ std::uint32_t spin_limit() const { return spin_limit_; }
    __device__ void          report_fault(std::uint32_t bits) const { bounded::report(fault_, bits); }

    __device__ Group<R>  get_group() const { return Group<R>{group_id_, group_range_}; }
    __device__ sub_group get_sub_group() const { return sub_group{}; }
    __device__ void      barrier() const { __syncthreads(); }

    id<R>                      group_id_;
    range<R>                   group_range_;
    std::uint32_t              segment_;
    std::uint32_t              segments_;
    const detail::SegmentInfo* table_;
    std::uint32_t*             ticket_state_;
    std::byte*                 base_;  ///< this stage's slice of the launch's LDS block
    std::uint32_t*             fault_;
    std::uint32_t              spin_limit_;
};

// -----------------------------------------------------------------------------
// LocalMem (reference kernel.hpp:515-570)
// -----------------------------------------------------------------------------
namespace detail {
template <class T>
struct is_local_array : std::false_type {};
template <DType DT, int R>
struct is_local_array<LocalArray<DT, R>> : std::true_type {};

CS_HD constexpr std::size_t align_up(std::size_t v, std::size_t a) { return (v + a - 1) / a * a; }

template <class... Vs>
struct LocalViews;
template <class A>
struct LocalViews<A> { A v0; };
template <class A, class B>
struct LocalViews<A, B> { A v0; B v1; };
template <class A, class B, class C>
struct LocalViews<A, B, C> { A v0; B v1; C v2; };
template <class A, class B, class C, class D>
struct LocalViews<A, B, C, D> { A v0; B v1; C v2; D v3; };
template <class A, class B, class C, class D, class E>
struct LocalViews<A, B, C, D, E> { A v0; B v1; C v2; D v3; E v4; };
template <class A, class B, class C, class D, class E, class F>
struct LocalViews<A, B, C, D, E, F> { A v0; B v1; C v2; D v3; E v4; F v5; };

template <std::size_t I, class... As>
struct nth_array;
template <class A, class... As>
struct nth_array<0, A, As...> { using type = A; };
template <std::size_t I, class A, class... As>
struct nth_array<I, A, As...> : nth_array<I - 1, As...> {};

/// Where the arrays of a LocalMem slice sit: walking the declarations in order, each array starts at the first multiple
/// of its element's alignment at or after the end of the one before. place() returns the end of the last array (the
/// slice's size) and, when `starts` is given, writes each array's first byte there.
template <class... As>
struct LocalLayout {
    static CS_HD std::size_t place(const range<3>*, std::size_t end, std::size_t*) { return end; }
};
template <class A, class... Rest>
struct LocalLayout<A, Rest...> {
    static CS_HD std::size_t place(const range<3>* shapes, std::size_t end, std::size_t* starts) {
        using T                 = typename A::value_type;
        const std::size_t first = align_up(end, alignof(T));
        if (starts != nullptr) *starts++ = first;
        return LocalLayout<Rest...>::place(shapes + 1, first + shapes->size() * sizeof(T), starts);
    }
};
}  // namespace detail

/// The local-memory arrays a group-range stage asks for (its fourth compute_forward parameter). It carries one shape per
/// array; the arrays themselves live in the stage's slice of the launch's LDS block, laid out by detail::LocalLayout,
/// and get() hands a work-item typed views of them in declaration order. Trivially copyable (it travels as a kernel
/// argument).
template <class... Arrays>
struct LocalMem {
    static_assert(sizeof...(Arrays) >= 1 && sizeof...(Arrays) <= 6, "cospan: LocalMem takes between 1 and 6 LocalArray declarations");
    static_assert((detail::is_local_array<Arrays>::value && ...), "cospan: every LocalMem parameter must be a cs::LocalArray");
    static constexpr std::size_t count = sizeof...(Arrays);
    template <std::size_t I>
    using array_t = typename detail::nth_array<I, Arrays...>::type;

    /// All shapes zero: a stage that asks for no local memory this launch.
    CS_HD LocalMem() {
        for (range<3>& r : shapes_) r = range<3>{0, 0, 0};
    }
    /// The shapes, the k-th for the k-th LocalArray parameter.
    CS_HD explicit LocalMem(range<Arrays::rank>... shapes) : shapes_{detail::lift_range<Arrays::rank>(shapes)...} {}

    template <std::size_t I>
    CS_HD range<array_t<I>::rank> shape() const {
        return detail::tail_range<array_t<I>::rank>(shapes_[I][0], shapes_[I][1], shapes_[I][2]);
    }

    /// The size of the stage's slice in bytes.
    CS_HD std::size_t bytes() const { return detail::LocalLayout<Arrays...>::place(shapes_, 0, nullptr); }

    /// The views of this stage's arrays for the calling work-item (no synchronisation).
    template <int R>
    __device__ auto get(const GroupItem<R>& it) const {
        std::size_t starts[count];
        detail::LocalLayout<Arrays...>::place(shapes_, 0, starts);
        return views(it.base_, starts, std::make_index_sequence<count>{});
    }

    range<3> shapes_[count];

private:
    template <std::size_t... I>
    __device__ auto views(std::byte* slice, const std::size_t* starts, std::index_sequence<I...>) const {
        using Views = detail::LocalViews<typename array_t<I>::view_type...>;
        return Views{typename array_t<I>::view_type{reinterpret_cast<typename array_t<I>::value_type*>(slice + starts[I]), shape<I>()}...};
    }
};

// -----------------------------------------------------------------------------
// Group functions (reference kernel.hpp:581-694), over Group<R>
// -----------------------------------------------------------------------------
namespace detail {
template <class Op, class T>
CS_HD constexpr T identity_of() {
    if constexpr (std::is_same_v<Op, minimum<T>>)
        return std::numeric_limits<T>::has_infinity ? std::numeric_limits<T>::infinity() : std::numeric_limits<T>::max();
    else if constexpr (std::is_same_v<Op, maximum<T>>)
        return std::numeric_limits<T>::has_infinity ? -std::numeric_limits<T>::infinity() : std::numeric_limits<T>::lowest();
    else if constexpr (std::is_same_v<Op, multiplies<T>>)
        return T(1);
    else if constexpr (std::is_same_v<Op, bit_and<T>>)
        return static_cast<T>(~T(0));
    else
        return T(0);
}
}  // namespace detail

template <int R>
__device__ inline void group_barrier(const Group<R>&, memory_scope = memory_scope::work_group) {
    __syncthreads();
}

template <int R, class T>
__device__ inline T group_broadcast(const Group<R>&, T x, std::size_t local_linear_id = 0) {
    __shared__ T slot;
    if (detail::local_linear_id() == local_linear_id) slot = x;
    __syncthreads();
    const T v = slot;
    __syncthreads();
    return v;
}
/// From the work-item at local id `local_id` (row-major over the local range, last dimension fastest).
template <int R, class T>
__device__ inline T group_broadcast(const Group<R>& g, T x, const id<R>& local_id) {
    const range<R> lr = g.get_local_range();
    std::size_t    l  = 0;
    for (int d = 0; d < R; ++d) l = l * lr[d] + local_id[d];
    return group_broadcast(g, x, l);
}

template <int R>
__device__ inline bool any_of_group(const Group<R>&, bool pred) {
    return __syncthreads_or(pred ? 1 : 0) != 0;
}
template <int R>
__device__ inline bool all_of_group(const Group<R>&, bool pred) {
    return __syncthreads_and(pred ? 1 : 0) != 0;
}
template <int R>
__device__ inline bool none_of_group(const Group<R>& g, bool pred) {
    return !any_of_group(g, pred);
}

/// Every work-item gets the reduction: wave butterflies then a pairwise tree over the waves (deterministic).
template <int R, class T, class Op>
__device__ inline T reduce_over_group(const Group<R>&, T x, Op op) {
    return block::reduce(x, op);
}
template <int R, class V, class T, class Op>
__device__ inline T reduce_over_group(const Group<R>& g, V x, T init, Op op) {
    return op(init, reduce_over_group(g, static_cast<T>(x), op));
}

template <int R, class T, class Op>
__device__ inline T exclusive_scan_over_group(const Group<R>&, T x, Op op) {
    return block::exclusive_scan(x, op, detail::identity_of<Op, T>());
}
template <int R, class V, class T, class Op>
__device__ inline T exclusive_scan_over_group(const Group<R>&, V x, T init, Op op) {
    return op(init, block::exclusive_scan(static_cast<T>(x), op, detail::identity_of<Op, T>()));
}
template <int R, class T, class Op>
__device__ inline T inclusive_scan_over_group(const Group<R>&, T x, Op op) {
    return op(block::exclusive_scan(x, op, detail::identity_of<Op, T>()), x);
}
template <int R, class V, class Op, class T>
__device__ inline T inclusive_scan_over_group(const Group<R>&, V x, Op op, T init) {
    return op(init, op(block::exclusive_scan(static_cast<T>(x), op, detail::identity_of<Op, T>()), static_cast<T>(x)));
}

// ---- joint algorithms over a range [first, last) that the whole group reads (reference kernel.hpp:614-676) ----
// Work-item t takes elements t, t + size, t + 2 size, ... (size = the group's work-item count); the range and the
// results are the same for every work-item. All are COLLECTIVE.

/// Reduction of *first .. *(last-1); an empty range gives the operator's identity.
template <int R, class Ptr, class Op>
__device__ inline auto joint_reduce(const Group<R>&, Ptr first, Ptr last, Op op) {
    using T              = std::remove_cv_t<std::remove_reference_t<decltype(*first)>>;
    const std::ptrdiff_t n = last - first;
    if (n <= 0) return detail::identity_of<Op, T>();
    const std::ptrdiff_t size = static_cast<std::ptrdiff_t>(detail::local_linear_range());
    const std::ptrdiff_t t    = static_cast<std::ptrdiff_t>(detail::local_linear_id());
    T                    acc  = t < n ? static_cast<T>(first[t]) : detail::identity_of<Op, T>();
    for (std::ptrdiff_t i = t + size; i < n; i += size) acc = op(acc, static_cast<T>(first[i]));
    return block::reduce_first(acc, op, static_cast<int>(n < size ? n : size));
}
template <int R, class Ptr, class T, class Op>
__device__ inline T joint_reduce(const Group<R>& g, Ptr first, Ptr last, T init, Op op) {
    if (last - first <= 0) return init;
    return op(init, static_cast<T>(joint_reduce(g, first, last, op)));
}

template <int R, class Ptr, class Pred>
__device__ inline bool joint_any_of(const Group<R>& g, Ptr first, Ptr last, Pred pred) {
    const std::ptrdiff_t n = last - first, size = static_cast<std::ptrdiff_t>(detail::local_linear_range());
    bool                 hit = false;
    for (std::ptrdiff_t i = static_cast<std::ptrdiff_t>(detail::local_linear_id()); i < n && !hit; i += size) hit = pred(first[i]);
    return any_of_group(g, hit);
}
template <int R, class Ptr, class Pred>
__device__ inline bool joint_all_of(const Group<R>& g, Ptr first, Ptr last, Pred pred) {
    const std::ptrdiff_t n = last - first, size = static_cast<std::ptrdiff_t>(detail::local_linear_range());
    bool                 ok = true;
    for (std::ptrdiff_t i = static_cast<std::ptrdiff_t>(detail::local_linear_id()); i < n && ok; i += size) ok = pred(first[i]);
    return all_of_group(g, ok);
}
template <int R, class Ptr, class Pred>
__device__ inline bool joint_none_of(const Group<R>& g, Ptr first, Ptr last, Pred pred) {
    return !joint_any_of(g, first, last, pred);
}

namespace detail {
/// The joint scans: the range in chunks of one element per work-item, each chunk scanned across the group and offset by
/// the running total of the chunks before it. `carry` starts as the init (or the identity). Returns result + n.
template <bool Inclusive, class InPtr, class OutPtr, class T, class Op>
__device__ inline OutPtr joint_scan(InPtr first, InPtr last, OutPtr result, T carry, Op op) {
    const std::ptrdiff_t n = last - first, size = static_cast<std::ptrdiff_t>(local_linear_range());
    const std::ptrdiff_t t  = static_cast<std::ptrdiff_t>(local_linear_id());
    const T              id = identity_of<Op, T>();
    for (std::ptrdiff_t base = 0; base < n; base += size) {
        const bool in    = base + t < n;
        const T    x     = in ? static_cast<T>(first[base + t]) : id;
        T          total = id;
        const T    ex    = block::exclusive_scan(x, op, id, &total);   // ends with a barrier: every read above is done
        if (in) result[base + t] = Inclusive ? op(carry, op(ex, x)) : op(carry, ex);
        carry = op(carry, total);
    }
    return result + (n > 0 ? n : 0);
}
}  // namespace detail

template <int R, class InPtr, class OutPtr, class Op>
__device__ inline OutPtr joint_exclusive_scan(const Group<R>&, InPtr first, InPtr last, OutPtr result, Op op) {
    using T = std::remove_cv_t<std::remove_reference_t<decltype(*first)>>;
    return detail::joint_scan<false>(first, last, result, detail::identity_of<Op, T>(), op);
}
template <int R, class InPtr, class OutPtr, class T, class Op>
__device__ inline OutPtr joint_exclusive_scan(const Group<R>&, InPtr first, InPtr last, OutPtr result, T init, Op op) {
    return detail::joint_scan<false>(first, last, result, init, op);
}
template <int R, class InPtr, class OutPtr, class Op>
__device__ inline OutPtr joint_inclusive_scan(const Group<R>&, InPtr first, InPtr last, OutPtr result, Op op) {
    using T = std::remove_cv_t<std::remove_reference_t<decltype(*first)>>;
    return detail::joint_scan<true>(first, last, result, detail::identity_of<Op, T>(), op);
}
template <int R, class InPtr, class OutPtr, class Op, class T>
__device__ inline OutPtr joint_inclusive_scan(const Group<R>&, InPtr first, InPtr last, OutPtr result, Op op, T init) {
    return detail::joint_scan<true>(first, last, result, init, op);
}

// ---- the sub-group (wave) shuffles (reference kernel.hpp:679-693) ----
// A source lane outside the sub-group's populated lanes gives the lane's own value (the result is unspecified there).

template <class T>
__device__ inline T select_from_group(sub_group g, T x, sub_group::id_type remote_local_id) {
    const std::uint32_t src = static_cast<std::uint32_t>(remote_local_id[0]);
    return wave::shfl(x, static_cast<int>(src < g.get_local_linear_range() ? src : g.get_local_linear_id()));
}
template <class T>
__device__ inline T shift_group_left(sub_group g, T x, sub_group::linear_id_type delta = 1) {
    const std::uint32_t src = g.get_local_linear_id() + delta;
    return wave::shfl(x, static_cast<int>(src < g.get_local_linear_range() ? src : g.get_local_linear_id()));
}
template <class T>
__device__ inline T shift_group_right(sub_group g, T x, sub_group::linear_id_type delta = 1) {
    const std::uint32_t l = g.get_local_linear_id();
    return wave::shfl(x, static_cast<int>(l >= delta ? l - delta : l));
}
template <class T>
__device__ inline T permute_group_by_xor(sub_group g, T x, sub_group::linear_id_type mask) {
    const std::uint32_t src = g.get_local_linear_id() ^ mask;
    return wave::shfl(x, static_cast<int>(src < g.get_local_linear_range() ? src : g.get_local_linear_id()));
}

// -----------------------------------------------------------------------------
// The body concepts, read off X::compute_forward (reference kernel.hpp:712-835, component.hpp:1245-1290)
// -----------------------------------------------------------------------------
namespace detail {
struct None {};

template <class F>
struct impl_call {
    static constexpr bool        valid = false;
    static constexpr std::size_t nargs = 0;
    using a0 = None;
    using a3 = None;
};
template <class Ret, class... A>
struct impl_call<Ret (*)(A...)> {
    static constexpr bool        valid = true;
    static constexpr std::size_t nargs = sizeof...(A);
    template <std::size_t I, class Def, class... B>
    struct at {
        using type = Def;
    };
    template <std::size_t I, class Def, class H, class... B>
    struct at<I, Def, H, B...> : at<I - 1, Def, B...> {};
    template <class Def, class H, class... B>
    struct at<0, Def, H, B...> {
        using type = H;
    };
    using a0 = std::remove_cvref_t<typename at<0, None, A...>::type>;
    using a3 = std::remove_cvref_t<typename at<3, None, A...>::type>;
};

template <class X, class = void>
struct impl_traits : impl_call<void> {};
template <class X>
struct impl_traits<X, std::void_t<decltype(&X::compute_forward)>> : impl_call<decltype(&X::compute_forward)> {};

/// kind: 0 index-range (id<R>), 1 group-range (GroupItem<R>).
template <class T>
struct first_traits {
    static constexpr int kind = -1;
    static constexpr int rank = 0;
};
template <int R>
struct first_traits<id<R>> {
    static constexpr int kind = 0;
    static constexpr int rank = R;
};
template <int R>
struct first_traits<GroupItem<R>> {
    static constexpr int kind = 1;
    static constexpr int rank = R;
};

template <class T>
struct is_local_mem : std::false_type {};
template <class... As>
struct is_local_mem<LocalMem<As...>> : std::true_type {};

template <class X>
inline constexpr int body_kind_v = first_traits<typename impl_traits<X>::a0>::kind;
template <class X>
inline constexpr int kernel_rank_v = first_traits<typename impl_traits<X>::a0>::rank;
template <class X>
inline constexpr bool has_local_mem_v = impl_traits<X>::nargs == 4 && is_local_mem<typename impl_traits<X>::a3>::value;

template <class X>
struct local_mem_of {
    using type = None;
};
template <class X>
    requires has_local_mem_v<X>
struct local_mem_of<X> {
    using type = typename impl_traits<X>::a3;
};
template <class X>
using local_mem_t = typename local_mem_of<X>::type;

/// The most work-items a stage's work-group may have, as a compile-time bound (X::max_work_group_size, else 1024): the
/// fused kernel's launch bound, so that register allocation fits the real work-group.
template <class X, class = void>
struct max_wg_of : std::integral_constant<int, 1024> {};
template <class X>
struct max_wg_of<X, std::void_t<decltype(X::max_work_group_size)>> : std::integral_constant<int, X::max_work_group_size> {};

/// The fewest waves per SIMD a stage asks the fused kernel to keep resident (X::min_waves_per_eu, else 1): the kernel's
/// register budget is then at most 512 / waves VGPRs, so a stage that keeps many loads in flight per wave does not trade
/// them for occupancy.
template <class X, class = void>
struct min_waves_of : std::integral_constant<int, 1> {};
template <class X>
struct min_waves_of<X, std::void_t<decltype(X::min_waves_per_eu)>> : std::integral_constant<int, X::min_waves_per_eu> {};
}  // namespace detail

template <class X>
concept IndexRangeStage = detail::body_kind_v<X> == 0 && detail::impl_traits<X>::nargs == 3;
template <class X>
concept GroupRangeStage = detail::body_kind_v<X> == 1 &&
                          (detail::impl_traits<X>::nargs == 3 || detail::has_local_mem_v<X>);
template <class X>
concept KernelStage = (IndexRangeStage<X> || GroupRangeStage<X>) && std::is_trivially_copyable_v<typename X::Inputs> &&
                      std::is_trivially_copyable_v<typename X::Outputs>;

// -----------------------------------------------------------------------------
// StageArgs and the argument pack of a launch
// -----------------------------------------------------------------------------
namespace detail {

/// Everything a stage reads of its component (reference kernel.hpp:875-886): its units (work-groups, or indices for an
/// index-range stage) as a rank-3 range with leading 1s, its Inputs and Outputs views, its LocalMem shapes.
template <class X>
struct StageArgs {
    std::uint64_t        units[3] = {0, 0, 0};
    typename X::Inputs   in;
    typename X::Outputs  out;
    local_mem_t<X>       local;
};

template <class... X>
struct ArgPack;
template <>
struct ArgPack<> {};
template <class H, class... T>
struct ArgPack<H, T...> {
    StageArgs<H>  head;
    ArgPack<T...> tail;
};

template <std::size_t I, class... X>
struct pack_get;
template <class H, class... T>
struct pack_get<0, H, T...> {
    static CS_HD StageArgs<H>&       get(ArgPack<H, T...>& p) { return p.head; }
    static CS_HD const StageArgs<H>& get(const ArgPack<H, T...>& p) { return p.head; }
};
template <std::size_t I, class H, class... T>
struct pack_get<I, H, T...> {
    static CS_HD auto&       get(ArgPack<H, T...>& p) { return pack_get<I - 1, T...>::get(p.tail); }
    static CS_HD const auto& get(const ArgPack<H, T...>& p) { return pack_get<I - 1, T...>::get(p.tail); }
};

template <class... X>
struct GroupArgs {
    Geometry<sizeof...(X)> geo;
    ArgPack<X...>          args;
};

struct ItemFixup {
    template <int R, int N>
    __device__ static GroupItem<R> make(const StageGeometry& g, const Geometry<N>& geo, const std::uint64_t* coord,
                                        const std::uint64_t* own, std::byte* lds) {
        const std::uint64_t* gr = own;
        if (g.range_mode == static_cast<std::uint32_t>(RangeMode::whole)) gr = geo.seg[g.segment].range;
        return GroupItem<R>{tail_id<R>(coord[0], coord[1], coord[2]), tail_range<R>(gr[0], gr[1], gr[2]), g.segment, geo.segments,
                            geo.seg, geo.ticket_state, lds + g.local_offset, geo.fault, geo.spin_limit};
    }
};

/// Stage K of a group-range launch: segment guard, own-range guard, then the body (reference kernel.hpp:919-950).
template <class X, int K, int N>
__device__ inline void stage_run(const Geometry<N>& geo, const StageArgs<X>& a, const LogicalGroup& lg, std::byte* lds) {
    const StageGeometry& g = geo.stage[K];
    if (lg.segment != g.segment) return;
    if (a.units[0] * a.units[1] * a.units[2] == 0) return;
    if (g.range_mode == static_cast<std::uint32_t>(RangeMode::own))
        for (int d = 0; d < 3; ++d)
            if (lg.coord[d] >= a.units[d]) return;
    constexpr int R = kernel_rank_v<X>;
    if constexpr (GroupRangeStage<X>) {
        const GroupItem<R> gi = ItemFixup::make<R>(g, geo, lg.coord, a.units, lds);
        if constexpr (has_local_mem_v<X>)
            X::compute_forward(gi, a.in, a.out, a.local);
        else
            X::compute_forward(gi, a.in, a.out);
    } else {
        X::compute_forward(tail_id<R>(lg.coord[0], lg.coord[1], lg.coord[2]), a.in, a.out);
    }
}

template <int N, class... X, std::size_t... K>
__device__ inline void run_stages(const GroupArgs<X...>& ga, const LogicalGroup& lg, std::byte* lds, std::index_sequence<K...>) {
    (stage_run<X, static_cast<int>(K), N>(ga.geo, pack_get<K, X...>::get(ga.args), lg, lds), ...);
}

/// THE fused kernel of a group-range group: one instantiation per distinct stage sequence (reference kernel.hpp:1015-1067).
template <int MaxThreads, int MinWaves, class... X>
__global__ __launch_bounds__(MaxThreads) __attribute__((amdgpu_waves_per_eu(MinWaves))) void group_kernel(const GroupArgs<X...> ga) {
    extern __shared__ __attribute__((aligned(16))) std::byte cs_group_lds[];
    constexpr int        N   = sizeof...(X);
    const Geometry<N>&   geo = ga.geo;
    std::uint32_t* const ts  = geo.ticket_state;
    LogicalGroup         lg;
    std::uint32_t        ticket = 0;
    if (ts != nullptr) {
        __shared__ std::uint32_t s_ticket;
        if (local_linear_id() == 0) s_ticket = __hip_atomic_fetch_add(&ts[0], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __syncthreads();
        ticket = s_ticket;
        lg     = locate_linear(geo.seg, geo.segments, ticket, geo.total_groups);
        if (!lg.valid) {
            // a ticket past the last group: the state was not zero when the launch began (two launches sharing it)
            if (local_linear_id() == 0) bounded::report(geo.fault, CS_HIP_FAULT_STATE);
            return;
        }
        if (geo.seg[lg.segment].boundary == static_cast<std::uint64_t>(Boundary::global_barrier))
            ticket_wait_segments(ts, geo.seg, lg.segment, geo.fault, geo.spin_limit);
    } else {
        lg = locate_physical(geo.split, geo.seg, geo.seg_lo, geo.seg_hi, geo.total_groups);
        if (!lg.valid) return;
    }

    run_stages<N>(ga, lg, cs_group_lds, std::make_index_sequence<N>{});

    if (ts != nullptr) {
        // completion: every wave waits for its stores to reach the L2, barrier, the group's stores published (a shared
        // write-back of the XCD's L2, publish_stores), the group's flag and its segment's count, then the finished count;
        // only the last group acquires, and zeroes the state (an acquire per group would invalidate its XCD's L2 every
        // time). The finished count must come after the flag and the count, or the last group could zero a word that one
        // of them then sets: both are read-modify-writes whose results the empty asm consumes, so the finished count is
        // issued only once both have been performed (a release fence would order them too, at the cost of another L2
        // write-back per work-group).
        asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
        __syncthreads();
        __shared__ std::uint32_t s_last;
        if (local_linear_id() == 0) {
            if (!publish_stores(&ts[write_back_word(geo.segments, 0, 0)], geo.spin_limit)) bounded::report(geo.fault, CS_HIP_FAULT_TICKET_WAIT);
            const std::uint32_t was = __hip_atomic_exchange(&ts[flag_word(geo.segments, ticket)], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            const std::uint32_t cnt = __hip_atomic_fetch_add(&ts[segment_word(lg.segment)], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            asm volatile("" ::"v"(was), "v"(cnt) : "memory");
            const std::uint32_t f    = __hip_atomic_fetch_add(&ts[finished_word()], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            const bool          last = static_cast<std::uint64_t>(f) + 1 == geo.total_groups;
            if (last) __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
            s_last = last ? 1u : 0u;
        }
        __syncthreads();
        if (s_last) {
            const std::uint64_t words = ticket_words(geo.segments, geo.total_groups);
            for (std::uint64_t w = local_linear_id(); w < words; w += local_linear_range())
                __hip_atomic_store(&ts[w], 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
    }
}

/// The fused kernel of an index-range group: a flat 1-D launch of indices (reference kernel.hpp:1073-1087).
template <class X, int K, int N>
__device__ inline void index_stage_run(const Geometry<N>& geo, const StageArgs<X>& a, const LogicalGroup& lg) {
    const StageGeometry& g = geo.stage[K];
    if (lg.segment != g.segment) return;
    if (g.range_mode == static_cast<std::uint32_t>(RangeMode::own))
        for (int d = 0; d < 3; ++d)
            if (lg.coord[d] >= a.units[d]) return;
    constexpr int R = kernel_rank_v<X>;
    X::compute_forward(tail_id<R>(lg.coord[0], lg.coord[1], lg.coord[2]), a.in, a.out);
}

template <int N, class... X, std::size_t... K>
__device__ inline void run_index_stages(const GroupArgs<X...>& ga, const LogicalGroup& lg, std::index_sequence<K...>) {
    (index_stage_run<X, static_cast<int>(K), N>(ga.geo, pack_get<K, X...>::get(ga.args), lg), ...);
}

template <class... X>
__global__ __launch_bounds__(256) void index_kernel(const GroupArgs<X...> ga) {
    constexpr int       N   = sizeof...(X);
    const std::uint64_t lin = ga.geo.seg[ga.geo.seg_lo].first + static_cast<std::uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    const LogicalGroup  lg  = locate_linear(ga.geo.seg, ga.geo.seg_hi, lin, ga.geo.total_groups);
    if (!lg.valid) return;
    run_index_stages<N>(ga, lg, std::make_index_sequence<N>{});
}

}  // namespace detail
}  // namespace naga::cospan
