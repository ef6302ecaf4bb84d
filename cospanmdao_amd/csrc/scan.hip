// cospanmdao_amd/csrc/scan.hip — the device exclusive scan and the ragged column build of libcospan_hip.
//
// Replaces naga::cospan::detail::exclusive_scan_u32 (cospanmdao/include/naga/cospan/device_scan.hpp:46-110: reduce-then-
// scan over work-groups, three kernels plus recursion) and build_column (shape_hook.hpp:215-348: a zeroing task, an
// extents kernel, the scan, a wrap-check kernel and a total task) with ONE launch each:
//
//   * a decoupled look-back scan: each work-group takes a ticket (tiles in start order, so a group only ever waits for
//     groups that have started), stages its 2048-element tile through LDS (coalesced striped loads, blocked per-thread
//     runs of 8), scans it (thread runs, DPP wave scans, wave totals through LDS), publishes its aggregate, and wave 0
//     looks back over 64 predecessors at a time (one status word per lane, ballot for the nearest inclusive prefix,
//     DPP sum of the window) before publishing its inclusive prefix;
//   * build_column's narrowing (max(v, 0), the negative / past-32-bits flags) is the scan's load prologue, its wrap
//     check (offs[i + 1] < offs[i], exact because every extent is below 2^32) the epilogue's per-element test;
//   * the last work-group to finish writes [total, flag] (the reference's host buffer contents) and leaves the
//     workspace zeroed for the next call: no memset, so a scan can sit inside a captured graph.

#include <hip/hip_runtime.h>

#include <cstdint>
#include <map>
#include <mutex>
#include <utility>
#include <vector>

#include "cospan_hip.h"
#include "internal.hpp"
#include "naga/cospan/device.hpp"

namespace {

namespace cs = naga::cospan;

constexpr int      kThreads = 256;
constexpr int      kItems   = 8;
constexpr int      kTile    = kThreads * kItems;
constexpr uint64_t kStateAgg  = 1ull << 62;
constexpr uint64_t kStateIncl = 2ull << 62;

/// The workspace: a header, then one status word per tile (state in bits 63..62, value in bits 31..0).
struct ScanHeader {
    uint32_t ticket;
    uint32_t done;
    uint32_t flag;
    uint32_t pad;
};

/// Where a wait that gives up reports, and how often a wait polls (cs_hip_spin_limit at launch).
struct WaitBound {
    uint32_t* fault;
    uint32_t  limit;
};

__device__ inline uint32_t atomic_load_u32(uint32_t* p) { return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT); }
__device__ inline void     atomic_store_u32(uint32_t* p, uint32_t v) { __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT); }
__device__ inline uint64_t atomic_load_u64(uint64_t* p) { return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT); }
__device__ inline void     atomic_store_u64(uint64_t* p, uint64_t v) { __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT); }

/// A source value as 32 bits with the flag bits of shape_hook.hpp:180-201 (negative: bit 1; past 2^32 - 1: bit 2; the
/// value is then 0).
template <class Src>
__device__ inline uint32_t narrow(Src v, uint32_t& fl) {
    if constexpr (std::is_signed_v<Src>) {
        if (v < Src{0}) {
            fl |= CS_HIP_FLAG_NEGATIVE;
            return 0u;
        }
    }
    if constexpr (sizeof(Src) > 4) {
        if (static_cast<uint64_t>(v) > 0xffffffffull) {
            fl |= CS_HIP_FLAG_EXCEEDS32;
            return 0u;
        }
    }
    return static_cast<uint32_t>(v);
}

/// Completion: the group's flag bits into the workspace, then the done counter. The last group to finish writes the
/// result words (when asked) and zeroes the workspace. Returns nothing; every work-item calls it.
__device__ inline void finish(ScanHeader* ws, uint64_t* status, uint32_t tiles, uint32_t group_flag, uint32_t* result,
                              uint32_t result_total, bool total_from_status) {
    __shared__ uint32_t s_last;
    const int           tid = threadIdx.x;
    if (tid == 0) {
        if (group_flag) __hip_atomic_fetch_or(&ws->flag, group_flag, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        const uint32_t d = __hip_atomic_fetch_add(&ws->done, 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
        s_last           = d == tiles - 1 ? 1u : 0u;
    }
    __syncthreads();
    if (!s_last) return;
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    if (tid == 0 && result) {
        result[0] = total_from_status ? static_cast<uint32_t>(atomic_load_u64(&status[tiles - 1])) : result_total;
        result[1] = atomic_load_u32(&ws->flag);
    }
    __syncthreads();  // the total is read before any work-item of this group zeroes the status words
    if (status)
        for (uint32_t t = static_cast<uint32_t>(tid); t < tiles; t += kThreads) atomic_store_u64(&status[t], 0);
    if (tid == 0) {
        atomic_store_u32(&ws->ticket, 0);
        atomic_store_u32(&ws->flag, 0);
        atomic_store_u32(&ws->done, 0);
    }
}

/// The scan. Column = true: narrowing prologue (extents out), wrap check, [total, flag] result.
template <bool Column, class Src>
__global__ __launch_bounds__(kThreads) void scan_kernel(const Src* __restrict__ in, uint64_t kstride, uint32_t* __restrict__ extents,
                                                        uint32_t* __restrict__ out, uint64_t n, uint32_t tiles, ScanHeader* ws,
                                                        uint32_t* result, WaitBound wb) {
    __shared__ uint32_t s_items[kTile];
    __shared__ uint32_t s_tile, s_prefix, s_flag;
    uint64_t* const     status = reinterpret_cast<uint64_t*>(ws + 1);
    const int           tid    = threadIdx.x;

    if (tid == 0) {
        s_tile = __hip_atomic_fetch_add(&ws->ticket, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        s_flag = 0;
    }
    __syncthreads();
    const uint32_t tile = s_tile;
    if (tile >= tiles) {
        // only with a workspace that was not zeroed (two scans sharing it): never write outside it, and say so
        if (tid == 0) cs::bounded::report(wb.fault, CS_HIP_FAULT_STATE);
        return;
    }
    const uint64_t base = static_cast<uint64_t>(tile) * kTile;

    // striped, coalesced loads into LDS (the narrowing prologue of a column build on the way)
    uint32_t fl = 0;
#pragma unroll
    for (int it = 0; it < kItems; ++it) {
        const uint64_t idx = base + static_cast<uint64_t>(it) * kThreads + tid;
        uint32_t       v   = 0;
        if (idx < n) {
            if constexpr (Column) {
                v            = narrow(in[idx * kstride], fl);
                extents[idx] = v;
            } else {
                v = static_cast<uint32_t>(in[idx]);
            }
        }
        s_items[it * kThreads + tid] = v;
    }
    __syncthreads();

    // blocked runs of kItems per work-item
    uint32_t x[kItems];
    uint32_t sum = 0;
#pragma unroll
    for (int k = 0; k < kItems; ++k) {
        x[k] = s_items[tid * kItems + k];
        sum += x[k];
    }
    uint32_t       agg     = 0;
    const uint32_t tprefix = cs::block::exclusive_scan(sum, cs::plus<uint32_t>{}, 0u, &agg);  // ends with a barrier

    // publish, then look back (wave 0)
    if (tid < cs::wave::kSize) {
        uint32_t prefix = 0;
        if (tile == 0) {
            if (tid == 0) atomic_store_u64(&status[0], kStateIncl | agg);
        } else {
            if (tid == 0) atomic_store_u64(&status[tile], kStateAgg | agg);
            int64_t end = tile;
            while (true) {
                const int64_t p  = end - 1 - tid;
                uint64_t      st = kStateIncl;  // before tile 0: an inclusive prefix of 0
                if (p >= 0) {
                    // every predecessor has started (tickets in start order) and publishes without waiting; the bound
                    // only guarantees that a corrupted workspace can neither hang the device nor pass unnoticed (a
                    // lane that gives up counts the tile as 0 and the look-back walks on)
                    st = 0;
                    if (!cs::bounded::poll([&] { return ((st = atomic_load_u64(&status[p])) >> 62) != 0; }, wb.limit)) {
                        st = 0;
                        cs::bounded::report(wb.fault, CS_HIP_FAULT_SCAN_LOOKBACK);
                        atomicOr(&s_flag, CS_HIP_FLAG_WAIT_TIMEOUT);
                    }
                }
                const uint64_t incl = cs::wave::ballot((st >> 62) == 2);
                uint32_t       v    = static_cast<uint32_t>(st);
                if (incl != 0 && tid > __builtin_ctzll(incl)) v = 0;
                const uint32_t window = cs::wave::readlane(cs::wave::inclusive_scan(v, cs::plus<uint32_t>{}, 0u), 63);
                prefix += window;
                if (incl != 0) break;
                end -= cs::wave::kSize;
            }
            if (tid == 0) atomic_store_u64(&status[tile], kStateIncl | (prefix + agg));
        }
        if (tid == 0) s_prefix = prefix;
    }
    __syncthreads();

    uint32_t run = s_prefix + tprefix;
#pragma unroll
    for (int k = 0; k < kItems; ++k) {
        const uint32_t e = run;
        run += x[k];
        if constexpr (Column) {
            if (run < e) fl |= CS_HIP_FLAG_EXCEEDS32;  // the running sum wrapped (exact: each extent is below 2^32)
        }
        s_items[tid * kItems + k] = e;
    }
    __syncthreads();
#pragma unroll
    for (int it = 0; it < kItems; ++it) {
        const uint64_t idx = base + static_cast<uint64_t>(it) * kThreads + tid;
        if (idx < n) out[idx] = s_items[it * kThreads + tid];
    }
    if (tile == tiles - 1 && tid == 0) out[n] = s_prefix + agg;
    if constexpr (Column) {
        if (fl) atomicOr(&s_flag, fl);
    }
    __syncthreads();
    finish(ws, status, tiles, Column ? s_flag : 0u, Column ? result : nullptr, 0u, true);
}

/// build_column, offsets mode (shape_hook.hpp:303-330): offs[i] = the value, extents[i] = offs[i + 1] - offs[i].
template <class Src>
__global__ __launch_bounds__(kThreads) void offsets_kernel(const Src* __restrict__ in, uint64_t count_src,
                                                           uint32_t* __restrict__ extents, uint32_t* __restrict__ offs,
                                                           uint32_t groups, ScanHeader* ws, uint32_t* result) {
    __shared__ uint32_t s_flag;
    const int           tid = threadIdx.x;
    if (tid == 0) s_flag = 0;
    __syncthreads();
    const uint64_t i  = static_cast<uint64_t>(blockIdx.x) * kThreads + tid;
    uint32_t       fl = 0;
    if (i < count_src) {
        uint32_t       fb = 0;
        const uint32_t x  = narrow(in[i], fb);
        if (i == 0 && x != 0) fb |= CS_HIP_FLAG_NEGATIVE;
        offs[i] = x;
        fl |= fb;
        if (i + 1 < count_src) {
            uint32_t       ignored = 0;
            const uint32_t nx      = narrow(in[i + 1], ignored);
            if (nx < x) {
                extents[i] = 0;
                fl |= CS_HIP_FLAG_NEGATIVE;
            } else {
                extents[i] = nx - x;
            }
        }
    }
    if (fl) atomicOr(&s_flag, fl);
    __syncthreads();
    uint32_t total = 0;
    {
        uint32_t ignored = 0;
        total            = narrow(in[count_src - 1], ignored);
    }
    finish(ws, nullptr, groups, s_flag, result, total, false);
}

/// The library-owned workspaces: one per (device, stream), so scans on different streams never share one; scans on one
/// stream run in order. An outgrown workspace is retired, not freed: a captured graph may still name it.
struct OwnedWorkspace {
    void*  ptr   = nullptr;
    size_t bytes = 0;
};
std::mutex                                             g_ws_mu;
std::map<std::pair<int, hipStream_t>, OwnedWorkspace> g_ws;
std::vector<void*>                                     g_ws_retired;

/// Runs `launch(workspace)` with the library-owned workspace of the current device and stream `s`, grown (outside any
/// capture) to `need` bytes. The lock is held through the launch, so no other thread can grow the workspace between
/// the lookup and the enqueue.
template <class Launch>
int with_owned_workspace(size_t need, cs_hip_stream s, Launch&& launch) {
    int dev = 0;
    if (int r = cs_hip::check(hipGetDevice(&dev), "hipGetDevice")) return r;
    std::lock_guard<std::mutex> lock{g_ws_mu};
    OwnedWorkspace&             w = g_ws[{dev, cs_hip::raw(s)}];
    if (w.bytes < need) {
        hipStreamCaptureStatus st = hipStreamCaptureStatusNone;
        if (int r = cs_hip::check(hipStreamIsCapturing(cs_hip::raw(s), &st), "hipStreamIsCapturing")) return r;
        if (st != hipStreamCaptureStatusNone)
            return cs_hip::fail(CS_HIP_ERR_STATE,
                                "scan: the library workspace must grow to %zu bytes, which cannot happen during graph capture; "
                                "run the scan once before capturing or pass a workspace",
                                need);
        const size_t bytes = (need < 4096 ? 4096 : need * 2 + 3) / 4 * 4;
        void*        p     = nullptr;
        if (int r = cs_hip::check(hipMalloc(&p, bytes), "hipMalloc")) return r;
        // zeroed on the scan's own stream, ahead of the scan (a kernel, not a memset)
        if (int r = cs_hip_fill_u32_async(p, 0u, bytes / 4, s)) return r;
        if (w.ptr) g_ws_retired.push_back(w.ptr);
        w.ptr   = p;
        w.bytes = bytes;
    }
    return launch(w.ptr);
}

WaitBound wait_bound() {
    WaitBound b{nullptr, cs_hip_spin_limit()};
    if (cs_hip::fault_word(&b.fault) != CS_HIP_OK) b.fault = nullptr;
    return b;
}

uint32_t tiles_for(size_t n) { return static_cast<uint32_t>(n == 0 ? 1 : (n + kTile - 1) / kTile); }

// ---- the scan that allocates nothing (cs_hip_exclusive_scan_u32_static) -------------------------------------------
// Its workspace is `scratch`, a file-scope device array (each device holds its own copy until the process ends), sized
// for CS_HIP_SCAN_STATIC_MAX_N elements: the header and one status word per tile. Host code never passes the array's name
// to a kernel (on the host it is only a stand-in): the first call on a device looks up the array's device address there
// with hipGetSymbolAddress and keeps it; kernels and the one-time clearing get that address.
constexpr uint32_t kStaticTiles = static_cast<uint32_t>((CS_HIP_SCAN_STATIC_MAX_N + kTile - 1) / kTile);
constexpr size_t   kStaticBytes = sizeof(ScanHeader) + sizeof(uint64_t) * kStaticTiles;
static_assert(kStaticBytes % sizeof(uint64_t) == 0, "the scratch array holds whole status words");
static __device__ __align__(256) uint64_t scratch[kStaticBytes / sizeof(uint64_t)];

constexpr int kMaxDevices = 64;
std::mutex    g_static_mu;
void*         g_static_ptr[kMaxDevices] = {};  // per device: the device address of `scratch` there (no host allocation)

/// The address of the current device's `scratch`, looked up (and the array cleared, on stream `s`, by a fill kernel) on
/// the first call on that device. Every scan leaves the array zeroed again, so later calls need no clearing.
int static_workspace(cs_hip_stream s, void** out) {
    int dev = 0;
    if (int r = cs_hip::check(hipGetDevice(&dev), "hipGetDevice")) return r;
    if (dev < 0 || dev >= kMaxDevices)
        return cs_hip::fail(CS_HIP_ERR_INVALID, "cs_hip_exclusive_scan_u32_static: device %d is past the %d this library tracks", dev,
                            kMaxDevices);
    std::lock_guard<std::mutex> lock{g_static_mu};
    if (g_static_ptr[dev] == nullptr) {
        void* p = nullptr;
        if (int r = cs_hip::check(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)), "hipGetSymbolAddress(scratch)")) return r;
        if (p == nullptr) return cs_hip::fail(CS_HIP_ERR_RUNTIME, "hipGetSymbolAddress(scratch) gave a null address");
        if (int r = cs_hip_fill_u32_async(p, 0u, kStaticBytes / 4, s)) return r;
        g_static_ptr[dev] = p;
    }
    *out = g_static_ptr[dev];
    return CS_HIP_OK;
}

template <bool Column, class Src>
int launch_scan(const Src* in, size_t kstride, uint32_t* extents, uint32_t* out, size_t n, uint32_t* result, void* workspace,
                cs_hip_stream s) {
    const uint32_t  tiles = tiles_for(n);
    const WaitBound wb    = wait_bound();
    auto            go    = [&](void* ws) {
        hipLaunchKernelGGL((scan_kernel<Column, Src>), dim3(tiles), dim3(kThreads), 0, cs_hip::raw(s), in, static_cast<uint64_t>(kstride),
                           extents, out, static_cast<uint64_t>(n), tiles, static_cast<ScanHeader*>(ws), result, wb);
        return cs_hip::check(hipGetLastError(), "scan launch");
    };
    return workspace ? go(workspace) : with_owned_workspace(cs_hip_scan_workspace_bytes(n), s, go);
}

template <class Src>
int launch_offsets(const Src* in, size_t count_src, uint32_t* extents, uint32_t* offs, uint32_t* result, void* workspace,
                   cs_hip_stream s) {
    const uint32_t groups = static_cast<uint32_t>((count_src + kThreads - 1) / kThreads);
    auto           go     = [&](void* ws) {
        hipLaunchKernelGGL((offsets_kernel<Src>), dim3(groups), dim3(kThreads), 0, cs_hip::raw(s), in, static_cast<uint64_t>(count_src),
                           extents, offs, groups, static_cast<ScanHeader*>(ws), result);
        return cs_hip::check(hipGetLastError(), "offsets launch");
    };
    return workspace ? go(workspace) : with_owned_workspace(cs_hip_scan_workspace_bytes(count_src), s, go);
}

}  // namespace

extern "C" {

size_t cs_hip_scan_workspace_bytes(size_t n) { return sizeof(ScanHeader) + sizeof(uint64_t) * static_cast<size_t>(tiles_for(n)); }

int cs_hip_scan_workspace_init(void* workspace, size_t bytes, cs_hip_stream s) {
    if (!workspace || bytes < sizeof(ScanHeader) || bytes % 4 != 0)
        return cs_hip::fail(CS_HIP_ERR_INVALID, "cs_hip_scan_workspace_init: a workspace of at least %zu bytes, a multiple of 4",
                            sizeof(ScanHeader));
    return cs_hip_fill_u32_async(workspace, 0u, bytes / 4, s);
}

int cs_hip_exclusive_scan_u32(const uint32_t* in, uint32_t* out, size_t n, void* workspace, cs_hip_stream s) {
    if (!out || (n != 0 && !in)) return cs_hip::fail(CS_HIP_ERR_INVALID, "cs_hip_exclusive_scan_u32: null pointer");
    if (n > (size_t{1} << 40)) return cs_hip::fail(CS_HIP_ERR_INVALID, "cs_hip_exclusive_scan_u32: n = %zu is too large", n);
    return launch_scan<false, uint32_t>(in, 1, nullptr, out, n, nullptr, workspace, s);
}

int cs_hip_exclusive_scan_u32_static(const uint32_t* in, uint32_t* out, size_t n, cs_hip_stream s) {
    if (!out || (n != 0 && !in)) return cs_hip::fail(CS_HIP_ERR_INVALID, "cs_hip_exclusive_scan_u32_static: null pointer");
    if (n > CS_HIP_SCAN_STATIC_MAX_N)
        return cs_hip::fail(CS_HIP_ERR_INVALID,
                            "cs_hip_exclusive_scan_u32_static: n = %zu; this scan's fixed workspace holds at most %llu elements "
                            "(use cs_hip_exclusive_scan_u32 for more)",
                            n, static_cast<unsigned long long>(CS_HIP_SCAN_STATIC_MAX_N));
    void* ws = nullptr;
    if (int r = static_workspace(s, &ws)) return r;
    return launch_scan<false, uint32_t>(in, 1, nullptr, out, n, nullptr, ws, s);
}

int cs_hip_build_column(int src_dtype, const void* src, size_t count_src, size_t k, int offsets_mode, uint32_t* extents,
                        uint32_t* offs, uint32_t* result, void* workspace, cs_hip_stream s) {
    if (!offs || !result || (count_src != 0 && (!src || !extents)))
        return cs_hip::fail(CS_HIP_ERR_INVALID, "cs_hip_build_column: null pointer");
    if (src_dtype < 0 || src_dtype > 3) return cs_hip::fail(CS_HIP_ERR_INVALID, "cs_hip_build_column: source dtype %d", src_dtype);
    if (offsets_mode) {
        if (count_src == 0)
            return cs_hip::fail(CS_HIP_ERR_INVALID, "cs_hip_build_column: an offsets source holds no offsets (it needs N + 1 >= 1 entries)");
        switch (src_dtype) {
            case 0: return launch_offsets(static_cast<const int32_t*>(src), count_src, extents, offs, result, workspace, s);
            case 1: return launch_offsets(static_cast<const uint32_t*>(src), count_src, extents, offs, result, workspace, s);
            case 2: return launch_offsets(static_cast<const int64_t*>(src), count_src, extents, offs, result, workspace, s);
            default: return launch_offsets(static_cast<const uint64_t*>(src), count_src, extents, offs, result, workspace, s);
        }
    }
    if (k == 0) return cs_hip::fail(CS_HIP_ERR_INVALID, "cs_hip_build_column: a counts source has no column 0 (k = 0)");
    switch (src_dtype) {
        case 0: return launch_scan<true>(static_cast<const int32_t*>(src), k, extents, offs, count_src, result, workspace, s);
        case 1: return launch_scan<true>(static_cast<const uint32_t*>(src), k, extents, offs, count_src, result, workspace, s);
        case 2: return launch_scan<true>(static_cast<const int64_t*>(src), k, extents, offs, count_src, result, workspace, s);
        default: return launch_scan<true>(static_cast<const uint64_t*>(src), k, extents, offs, count_src, result, workspace, s);
    }
}

}  // extern "C"
