#pragma once

// =============================================================================
// naga/cospan/group_launch.hpp — one launch of a fused group (host side)
// =============================================================================
//
// The host half of the reference's GroupLauncher::launch (cospanmdao/include/naga/cospan/component.hpp:779-970): from the
// stages' launch facts (units, work-group shape, LDS bytes) and the segment layout it builds the geometry (one entry per
// stage, the segment table), the physical grid of the range_split strategy (component.hpp:936-954), the LDS block
// (partitioned: a slice per stage; shared: one block), the ticket state when the group waits, and submits the group's
// kernel through the C boundary (cs_hip_launch): once, or once per barrier-free run of segments when the group's global
// barriers are launch boundaries. Checks with the reference's messages: one work-group shape per group
// (component.hpp:815-829), the LDS limit (:880-890).

#include <algorithm>
#include <cstddef>
#include <cstring>
#include <functional>
#include <new>
#include <type_traits>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

#include "cospan_hip.h"
#include "naga/cospan/kernel.hpp"

namespace naga::cospan::detail {

/// A C-boundary status as an exception (the reference's exception types: std::invalid_argument for a bad argument,
/// std::logic_error for a call out of order, std::runtime_error for the runtime).
inline void hip_check(int status, const std::string& what) {
    if (status == CS_HIP_OK) return;
    char msg[512];
    cs_hip_last_error(msg, sizeof msg);
    const std::string text = "naga::cospan: " + what + ": " + msg;
    if (status == CS_HIP_ERR_INVALID) throw std::invalid_argument(text);
    if (status == CS_HIP_ERR_STATE) throw std::logic_error(text);
    throw std::runtime_error(text);
}

/// Reads and clears the device's fault word after completed work; throws std::runtime_error when a bounded device wait of
/// that work gave up (its results are then not valid).
inline void check_device_faults(const std::string& who) {
    std::uint32_t bits = 0;
    hip_check(cs_hip_take_fault(&bits), who);
    if (bits == 0) return;
    std::string what;
    auto        name = [&](std::uint32_t bit, const char* text) {
        if (bits & bit) what += (what.empty() ? "" : ", ") + std::string(text);
    };
    name(CS_HIP_FAULT_TICKET_WAIT, "a fused group's wait for an earlier group");
    name(CS_HIP_FAULT_SCAN_LOOKBACK, "the device scan's look-back");
    name(CS_HIP_FAULT_STATE, "ticket state or scan workspace not zeroed at launch");
    name(CS_HIP_FAULT_UNIT_WAIT, "a work-queue kernel's wait for an earlier unit");
    throw std::runtime_error("naga::cospan: " + who + ": a bounded device wait gave up (fault bits " + std::to_string(bits) + ": " +
                             what + "); the results of that work are not valid");
}

/// What a launcher needs of one stage.
struct StagePlan {
    std::uint64_t units[3] = {1, 1, 1};  ///< work-groups (group-range) or indices (index-range), leading 1s
    std::uint32_t wg[3]    = {1, 1, 1};  ///< work-group shape, physical order (d0, d1, d2), leading 1s
    std::size_t   local_bytes = 0;
    std::size_t   local_align = 1;
    bool          group_range = true;
    std::string   who;  ///< "'path'" for messages
};

/// What a launcher needs of one segment: its stages [first, first + count), options, the boundary before it.
struct SegmentPlan {
    int       first = 0;
    int       count = 0;
    RangeMode range = RangeMode::own;
    LocalMode local = LocalMode::partitioned;
    Boundary  boundary = Boundary::none;
};

/// The ticket state of a waiting group: device words, zeroed once at allocation; every launch leaves them zeroed.
class TicketState {
public:
    TicketState() = default;
    TicketState(const TicketState&)            = delete;
    TicketState& operator=(const TicketState&) = delete;
    ~TicketState() {
        if (ptr_) cs_hip_free(ptr_);
    }
    std::uint32_t* ensure(std::size_t words, cs_hip_stream s) {
        if (words > words_) {
            if (ptr_) {
                hip_check(cs_hip_stream_synchronize(s), "ticket state");
                hip_check(cs_hip_free(ptr_), "ticket state");
                ptr_ = nullptr;
            }
            const std::size_t w = words < 1024 ? 1024 : words;
            void*             p = nullptr;
            hip_check(cs_hip_malloc(&p, w * 4), "ticket state");
            hip_check(cs_hip_fill_u32_async(p, 0u, w, s), "ticket state");
            ptr_   = static_cast<std::uint32_t*>(p);
            words_ = w;
        }
        return ptr_;
    }

private:
    std::uint32_t* ptr_   = nullptr;
    std::size_t    words_ = 0;
};

/// One prepared launch: the kernel's host stub, its shape and a copy of its argument bytes. Submitting it is one
/// cs_hip_launch; two runs whose records are byte-identical launch identical work (the test for graph replay).
struct LaunchRecord {
    const void*            kernel   = nullptr;
    std::uint32_t          grid[3]  = {0, 0, 0};
    std::uint32_t          block[3] = {1, 1, 1};
    std::size_t            lds      = 0;
    std::vector<std::byte> args;  ///< the one by-value kernel argument (GroupArgs<X...>), or a custom op's ports
    std::string            name;
    /// A custom-op launch (CustomComponent): native work enqueued on the stream instead of a group kernel.
    std::function<void(cs_hip_stream)> custom;
    /// Launches that follow this one on the same stream (a fused group whose global barriers are launch boundaries).
    std::vector<LaunchRecord> then;

    bool empty() const { return kernel == nullptr && !custom && then.empty(); }
    void submit(cs_hip_stream s) const {
        if (custom) {
            custom(s);
            return;
        }
        if (kernel) {
            void* a[] = {const_cast<std::byte*>(args.data())};
            hip_check(cs_hip_launch(kernel, grid, block, lds, a, s), "launch of '" + name + "'");
        }
        for (const LaunchRecord& t : then) t.submit(s);
    }
    /// Appends the identity of this launch to `sig`.
    void append_signature(std::vector<std::byte>& sig) const {
        auto put = [&](const void* p, std::size_t n) {
            const auto* b = static_cast<const std::byte*>(p);
            sig.insert(sig.end(), b, b + n);
        };
        put(&kernel, sizeof kernel);
        put(grid, sizeof grid);
        put(block, sizeof block);
        put(&lds, sizeof lds);
        put(args.data(), args.size());
        for (const LaunchRecord& t : then) t.append_signature(sig);
    }
};

/// Member-wise copies of a stage's arguments (the padding of the destination is left as it is).
template <class X>
inline void assign_stage(StageArgs<X>& d, const StageArgs<X>& s) {
    for (int i = 0; i < 3; ++i) d.units[i] = s.units[i];
    d.in  = s.in;
    d.out = s.out;
    if constexpr (!std::is_empty_v<local_mem_t<X>>) d.local = s.local;
}
template <class H, class... T>
inline void assign_pack(ArgPack<H, T...>& d, const ArgPack<H, T...>& s) {
    assign_stage(d.head, s.head);
    if constexpr (sizeof...(T) > 0) assign_pack(d.tail, s.tail);
}

/// Builds the geometry of a group and launches its kernel.
template <class... X>
class GroupLauncher {
public:
    static constexpr int N = sizeof...(X);
    static_assert(N >= 1 && N <= kMaxStages, "cospan: a fused group has 1..16 stages");
    static constexpr bool kGroupRange = (GroupRangeStage<X> && ...);
    static constexpr bool kIndexRange = (IndexRangeStage<X> && ...);
    static_assert(kGroupRange || kIndexRange,
                  "cospan: the stages of one fused group are all group-range or all index-range components");
    static constexpr int kMaxThreads = std::min({max_wg_of<X>::value...});
    static constexpr int kMinWaves   = std::max({min_waves_of<X>::value...});

    /// The host stub of this group's kernel (what the process-wide kernel cache of the reference keys by stage sequence:
    /// here the C++ type system does it, one instantiation per sequence).
    static const void* kernel() {
        if constexpr (kGroupRange)
            return reinterpret_cast<const void*>(&group_kernel<kMaxThreads, kMinWaves, X...>);
        else
            return reinterpret_cast<const void*>(&index_kernel<X...>);
    }

    /// Launches the group on `s`. `split`: the range_split strategy; `waits`: group_waits_enabled.
    void launch(Split split, bool waits, const std::vector<SegmentPlan>& segs, const StagePlan (&stages)[N], const ArgPack<X...>& args,
                cs_hip_stream s, const std::string& group_name) {
        prepare(split, waits, segs, stages, args, s, group_name).submit(s);
    }

    /// The launch of the group as a record (nothing is submitted; the ticket state is allocated on `s` when first needed).
    /// `waits`: group waits (ticket mode for the whole group). Without them, every global barrier is a launch boundary: each
    /// barrier-free run of segments is one launch of the same kernel, in order on one stream, so no group ever waits and
    /// no group pays an agent-scope release to tell others it is done (on CDNA that release writes the L2 back).
    LaunchRecord prepare(Split split, bool waits, const std::vector<SegmentPlan>& segs, const StagePlan (&stages)[N],
                         const ArgPack<X...>& args, cs_hip_stream s, const std::string& group_name) {
        const std::size_t nseg = segs.size();
        if (nseg == 0 || nseg > static_cast<std::size_t>(N)) throw std::logic_error("naga::cospan: '" + group_name + "': bad segment layout");
        bool barrier = false;
        for (const SegmentPlan& sp : segs) barrier = barrier || sp.boundary == Boundary::global_barrier;
        if (waits || !barrier) return prepare_run(split, waits, segs, 0, nseg, stages, args, s, group_name);
        LaunchRecord rec;
        rec.name       = group_name;
        bool        first = true;
        std::size_t a     = 0;
        for (std::size_t b = 1; b <= nseg; ++b) {
            if (b < nseg && segs[b].boundary != Boundary::global_barrier) continue;
            LaunchRecord part = prepare_run(split, false, segs, a, b, stages, args, s, group_name);
            if (first)
                rec = std::move(part);
            else
                rec.then.push_back(std::move(part));
            first = false;
            a     = b;
        }
        return rec;
    }

private:
    /// One launch running segments [a, b) of the group. The segment table always holds every segment with the indices,
    /// group counts and first tickets the stages see (get_segment, get_segment_count, get_segment_group_count); the launch
    /// covers the groups of [a, b) only.
    LaunchRecord prepare_run(Split split, bool waits, const std::vector<SegmentPlan>& segs, std::size_t a, std::size_t b,
                             const StagePlan (&stages)[N], const ArgPack<X...>& args, cs_hip_stream s, const std::string& group_name) {
        LaunchRecord    rec;
        rec.name = group_name;
        GroupArgs<X...> ga{};
        ga.args                 = args;
        Geometry<N>&  geo       = ga.geo;
        const auto    nseg      = static_cast<std::uint32_t>(segs.size());
        geo.segments = nseg;
        geo.seg_lo   = static_cast<std::uint32_t>(a);
        geo.seg_hi   = static_cast<std::uint32_t>(b);
        geo.split    = static_cast<std::uint32_t>(split);

        // the one work-group shape of the group (group-range stages)
        std::uint32_t wg[3] = {1, 1, 1};
        if constexpr (kGroupRange) {
            bool first = true;
            for (int k = 0; k < N; ++k) {
                if (first) {
                    for (int d = 0; d < 3; ++d) wg[d] = stages[k].wg[d];
                    first = false;
                } else if (stages[k].wg[0] != wg[0] || stages[k].wg[1] != wg[1] || stages[k].wg[2] != wg[2]) {
                    throw std::invalid_argument("naga::cospan: fused group '" + group_name + "': stage " + stages[k].who +
                                                " has a different work-group shape; all stages of a group share one");
                }
            }
            if (static_cast<std::uint64_t>(wg[0]) * wg[1] * wg[2] > static_cast<std::uint64_t>(kMaxThreads))
                throw std::invalid_argument("naga::cospan: fused group '" + group_name + "': a work-group of " +
                                            std::to_string(wg[0] * wg[1] * wg[2]) + " work-items exceeds the stages' bound of " +
                                            std::to_string(kMaxThreads));
        }

        // segment table (all segments), LDS layout; the launch's extent from segments [a, b)
        std::uint64_t first_ticket = 0, offset = 0, lds = 0, run_groups = 0, run_offset = 0;
        std::uint64_t maxr[3]      = {0, 0, 0};
        for (std::uint32_t si = 0; si < nseg; ++si) {
            const SegmentPlan& sp = segs[si];
            SegmentInfo&       info = geo.seg[si];
            std::uint64_t      seg_lds = 0;
            for (int k = sp.first; k < sp.first + sp.count; ++k) {
                for (int d = 0; d < 3; ++d)
                    if (stages[k].units[d] > info.range[d]) info.range[d] = stages[k].units[d];
                StageGeometry& g = geo.stage[k];
                g.segment        = si;
                g.range_mode     = static_cast<std::uint32_t>(sp.range);
                if (sp.local == LocalMode::shared) {
                    g.local_offset = 0;
                    if (stages[k].local_bytes > seg_lds) seg_lds = stages[k].local_bytes;
                } else {
                    seg_lds        = align_up(seg_lds, stages[k].local_align);
                    g.local_offset = static_cast<std::uint32_t>(seg_lds);
                    seg_lds += stages[k].local_bytes;
                }
                g.local_bytes = static_cast<std::uint32_t>(stages[k].local_bytes);
            }
            info.count    = info.range[0] * info.range[1] * info.range[2];
            info.first    = first_ticket;
            info.offset   = offset;
            info.boundary = static_cast<std::uint64_t>(sp.boundary);
            first_ticket += info.count;
            offset += info.range[2];
            if (si >= a && si < b) {
                run_groups += info.count;
                run_offset += info.range[2];
                for (int d = 0; d < 3; ++d)
                    if (info.range[d] > maxr[d]) maxr[d] = info.range[d];
                if (seg_lds > lds) lds = seg_lds;
            }
        }
        // the bound of a linear group index in this launch (tickets and flattened ids count from segment 0)
        geo.total_groups = geo.seg[b - 1].first + geo.seg[b - 1].count;
        if (lds > 160u * 1024u)
            throw std::invalid_argument("naga::cospan: fused group '" + group_name + "': " + std::to_string(lds) +
                                        " bytes of local memory exceed the 163840 of a CDNA4 work-group");
        if (run_groups == 0) return rec;  // nothing to run: an empty record

        // physical grid (x is physical dimension 2); a hardware grid allows 65535 groups in y and z, beyond that the
        // groups are flattened along x (the logical coordinates are the same either way)
        if (split != Split::flatten && (maxr[0] > 65535 || maxr[1] > 65535 || b - a > 65535)) {
            split     = Split::flatten;
            geo.split = static_cast<std::uint32_t>(split);
        }
        std::uint32_t grid[3] = {1, 1, 1};
        if constexpr (kIndexRange) {
            grid[0] = static_cast<std::uint32_t>((run_groups + 255) / 256);
            wg[2]   = 256;
            if (waits) throw std::invalid_argument("naga::cospan: fused group '" + group_name + "': index-range stages cannot wait");
            geo.split = static_cast<std::uint32_t>(Split::flatten);
        } else if (waits || split == Split::flatten) {
            if (run_groups > 0x7fffffffull)
                throw std::invalid_argument("naga::cospan: fused group '" + group_name + "': too many work-groups for one launch");
            grid[0] = static_cast<std::uint32_t>(run_groups);
            if (waits) geo.split = static_cast<std::uint32_t>(Split::flatten);
        } else if (split == Split::pad) {
            grid[0] = static_cast<std::uint32_t>(run_offset);
            grid[1] = static_cast<std::uint32_t>(maxr[1]);
            grid[2] = static_cast<std::uint32_t>(maxr[0]);
        } else {
            if (maxr[0] > 1) throw std::invalid_argument("naga::cospan: fused group '" + group_name + "': range_split::padded needs stages of rank 2 or lower");
            grid[0] = static_cast<std::uint32_t>(maxr[2]);
            grid[1] = static_cast<std::uint32_t>(maxr[1]);
            grid[2] = static_cast<std::uint32_t>(b - a);
        }

        if (waits) geo.ticket_state = tickets_.ensure(ticket_words(nseg, geo.total_groups), s);
        // where a wait that gives up reports, and how often a wait polls (both are argument bytes: a changed bound is a
        // different launch, so a captured graph is not replayed with the old one)
        hip_check(cs_hip_fault_word(&geo.fault), "fault word of '" + group_name + "'");
        geo.spin_limit = cs_hip_spin_limit();

        static_assert(sizeof(GroupArgs<X...>) <= 4096, "cospan: a fused group's arguments exceed the 4 KiB kernel-argument limit");
        static_assert(std::is_trivially_copyable_v<GroupArgs<X...>>, "cospan: a fused group's arguments are copied as bytes");
        static_assert(std::has_unique_object_representations_v<Geometry<N>>,
                      "cospan: the geometry has padding bytes, which would differ between identical launches");
        rec.kernel = kernel();
        for (int d = 0; d < 3; ++d) rec.grid[d] = grid[d];
        rec.block[0] = wg[2];
        rec.block[1] = wg[1];
        rec.block[2] = wg[0];
        rec.lds      = static_cast<std::size_t>(lds);
        // the argument bytes: a zeroed buffer, then member-wise assignment, so padding is zero and identical launches have
        // identical bytes (the graph-replay signature compares them)
        rec.args.assign(sizeof ga, std::byte{0});
        auto* out = ::new (static_cast<void*>(rec.args.data())) GroupArgs<X...>;
        out->geo  = ga.geo;
        assign_pack(out->args, ga.args);
        return rec;
    }

    TicketState tickets_;
};

}  // namespace naga::cospan::detail
