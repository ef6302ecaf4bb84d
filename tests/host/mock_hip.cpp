// tests/host/mock_hip.cpp — the CPU implementation of include/cospan_hip.h behind host-logic tests (TEST INFRASTRUCTURE).
// See mock_hip.hpp for the model of streams, events, graphs and device memory it implements.
#include "mock_hip.hpp"

#include <algorithm>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <string>
#include <vector>

#include "cospan_hip.h"
#include "oracle.h"

namespace {

struct Event {
    std::uint64_t recorded  = 0;  // record calls so far
    std::uint64_t completed = 0;  // the latest record op that ran
    bool          in_capture = false;
};

struct Op {
    enum Kind { work, wait, record } kind = work;
    std::function<void()> fn;
    Event*                ev  = nullptr;
    std::uint64_t         gen = 0;
};

struct Stream {
    std::deque<Op> q;
    int            id = 0;
};

struct Graph {
    std::vector<std::function<void()>> ops;
};

struct Kernel {
    std::size_t    bytes = 0;
    mock::KernelFn fn;
};

struct State {
    std::map<const void*, Kernel>      kernels;
    mock::KernelFn                     fallback;
    std::map<std::uintptr_t, std::size_t> allocs;  // start -> bytes
    std::vector<std::unique_ptr<Stream>> streams;  // [0] is the null stream
    std::vector<std::unique_ptr<Event>>  events;
    bool                                 capturing = false;
    Graph*                               graph     = nullptr;
    mock::Stats                          stats;
    std::string                          failure;
    State() {
        streams.push_back(std::make_unique<Stream>());
        streams.back()->id = 0;
    }
};

State& S() {
    static State s;
    return s;
}

thread_local std::string t_err;

int fail(int code, const char* fmt, ...) {
    char    buf[512];
    va_list ap;
    va_start(ap, fmt);
    std::vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    t_err = buf;
    return code;
}

void note_failure(const std::string& what) {
    if (S().failure.empty()) S().failure = what;
}

Stream* stream_of(cs_hip_stream s) { return s ? reinterpret_cast<Stream*>(s) : S().streams[0].get(); }

/// Work enqueued on `s` (recorded into the graph while capturing).
void enqueue(cs_hip_stream s, std::function<void()> fn) {
    State& st = S();
    if (st.capturing) {
        st.graph->ops.push_back(std::move(fn));
        return;
    }
    Op op;
    op.fn = std::move(fn);
    stream_of(s)->q.push_back(std::move(op));
}

bool runnable(const Op& op) { return op.kind != Op::wait || op.ev->completed >= op.gen; }

/// Runs queued work until every queue is empty, one operation at a time, always from the highest-numbered stream that has
/// runnable work (an operation that unblocks a higher stream hands over to it at once).
void drain() {
    State& st = S();
    for (;;) {
        bool progress = false, pending = false;
        for (auto it = st.streams.rbegin(); it != st.streams.rend(); ++it) {
            Stream& q = **it;
            if (!q.q.empty() && runnable(q.q.front())) {
                Op op = std::move(q.q.front());
                q.q.pop_front();
                if (op.kind == Op::work) op.fn();
                if (op.kind == Op::record) op.ev->completed = std::max(op.ev->completed, op.gen);
                progress = true;
                pending  = true;
                break;
            }
            pending = pending || !q.q.empty();
        }
        if (!pending) return;
        if (!progress) {
            note_failure("deadlock: a stream waits on an event whose record never runs");
            for (auto& q : st.streams) q->q.clear();
            return;
        }
    }
}

bool device_range(const void* p, std::size_t bytes) {
    const auto   a  = reinterpret_cast<std::uintptr_t>(p);
    const State& st = S();
    auto         it = st.allocs.upper_bound(a);
    if (it == st.allocs.begin()) return false;
    --it;
    return a >= it->first && a + bytes <= it->first + it->second;
}

// the device hash: the oracle's generator (examples/gemver-chain/common/chain_data.hpp:34-49), bit-identical to the fill kernel
float hash_unit(std::uint32_t seed, std::uint64_t index) { return or_hash_unit(seed, index); }

template <class Src>
std::uint32_t narrow(Src v, std::uint32_t& fl) {
    if constexpr (std::is_signed_v<Src>) {
        if (v < Src{0}) {
            fl |= CS_HIP_FLAG_NEGATIVE;
            return 0;
        }
    }
    if constexpr (sizeof(Src) > 4) {
        if (static_cast<std::uint64_t>(v) > 0xffffffffull) {
            fl |= CS_HIP_FLAG_EXCEEDS32;
            return 0;
        }
    }
    return static_cast<std::uint32_t>(v);
}

bool all_zero(const void* p, std::size_t bytes) {
    const auto* b = static_cast<const unsigned char*>(p);
    for (std::size_t i = 0; i < bytes; ++i)
        if (b[i]) return false;
    return true;
}

template <class Src>
void column_counts(const Src* in, std::size_t k, std::uint32_t* ext, std::uint32_t* offs, std::size_t n, std::uint32_t* result) {
    std::uint32_t fl = 0, run = 0;
    for (std::size_t i = 0; i < n; ++i) {
        const std::uint32_t v = narrow(in[i * k], fl);
        if (ext) ext[i] = v;
        offs[i] = run;
        const std::uint32_t next = run + v;
        if (next < run) fl |= CS_HIP_FLAG_EXCEEDS32;
        run = next;
    }
    offs[n] = run;
    if (result) {
        result[0] = run;
        result[1] = fl;
    }
}

}  // namespace

namespace mock {
void register_kernel(const void* stub, std::size_t arg_bytes, KernelFn fn) { S().kernels[stub] = Kernel{arg_bytes, std::move(fn)}; }
void set_fallback_kernel(KernelFn fn) { S().fallback = std::move(fn); }
bool is_device(const void* p) { return device_range(p, 1); }
Stats stats() { return S().stats; }
std::string failure() { return S().failure; }
}  // namespace mock

extern "C" {

int cs_hip_last_error(char* buf, size_t len) {
    if (buf && len) {
        std::snprintf(buf, len, "%s", t_err.c_str());
    }
    return static_cast<int>(t_err.size());
}

int cs_hip_init(int) { return CS_HIP_OK; }
int cs_hip_device_count(int* c) {
    *c = 1;
    return CS_HIP_OK;
}
int cs_hip_device_info(int, char* name, size_t len, int* cus, int* lds, int* wg) {
    if (name && len) std::snprintf(name, len, "mock");
    if (cus) *cus = 8;
    if (lds) *lds = 163840;
    if (wg) *wg = 1024;
    return CS_HIP_OK;
}
int cs_hip_device_synchronize(void) {
    if (S().capturing) note_failure("device synchronize during capture");
    drain();
    return CS_HIP_OK;
}

int cs_hip_stream_create(cs_hip_stream* s) {
    auto st = std::make_unique<Stream>();
    st->id  = static_cast<int>(S().streams.size());
    *s      = reinterpret_cast<cs_hip_stream>(st.get());
    S().streams.push_back(std::move(st));
    return CS_HIP_OK;
}
int cs_hip_stream_destroy(cs_hip_stream s) {
    drain();
    auto& v = S().streams;
    v.erase(std::remove_if(v.begin(), v.end(), [&](const std::unique_ptr<Stream>& p) { return reinterpret_cast<cs_hip_stream>(p.get()) == s; }),
            v.end());
    return CS_HIP_OK;
}
int cs_hip_stream_synchronize(cs_hip_stream) {
    if (S().capturing) note_failure("stream synchronize during capture");
    drain();
    return CS_HIP_OK;
}

int cs_hip_event_create(cs_hip_event* e, int) {
    auto ev = std::make_unique<Event>();
    *e      = reinterpret_cast<cs_hip_event>(ev.get());
    S().events.push_back(std::move(ev));
    return CS_HIP_OK;
}
int cs_hip_event_destroy(cs_hip_event) { return CS_HIP_OK; }  // kept alive: queued ops may still point at it
int cs_hip_event_record(cs_hip_event e, cs_hip_stream s) {
    Event* ev = reinterpret_cast<Event*>(e);
    State& st = S();
    if (st.capturing) {
        ev->in_capture = true;  // fork/join inside the graph: order is the recorded order
        return CS_HIP_OK;
    }
    ev->in_capture = false;
    Op op;
    op.kind = Op::record;
    op.ev   = ev;
    op.gen  = ++ev->recorded;
    stream_of(s)->q.push_back(std::move(op));
    return CS_HIP_OK;
}
int cs_hip_stream_wait_event(cs_hip_stream s, cs_hip_event e) {
    Event* ev = reinterpret_cast<Event*>(e);
    State& st = S();
    if (st.capturing) return CS_HIP_OK;
    if (ev->in_capture) note_failure("a stream outside a capture waits on an event recorded inside one");
    if (ev->recorded == 0) return CS_HIP_OK;
    Op op;
    op.kind = Op::wait;
    op.ev   = ev;
    op.gen  = ev->recorded;
    stream_of(s)->q.push_back(std::move(op));
    return CS_HIP_OK;
}
int cs_hip_event_query(cs_hip_event e, int* done) {
    drain();
    Event* ev = reinterpret_cast<Event*>(e);
    *done     = ev->completed >= ev->recorded ? 1 : 0;
    return CS_HIP_OK;
}
int cs_hip_event_synchronize(cs_hip_event) {
    drain();
    return CS_HIP_OK;
}
int cs_hip_event_elapsed_ms(float* ms, cs_hip_event, cs_hip_event) {
    drain();
    *ms = 0.001f;
    return CS_HIP_OK;
}

int cs_hip_malloc(void** p, size_t bytes) {
    if (S().capturing) note_failure("allocation during capture");
    void* m = std::aligned_alloc(256, (bytes + 255) / 256 * 256);
    if (!m) return fail(CS_HIP_ERR_RUNTIME, "mock: out of memory (%zu bytes)", bytes);
    std::memset(m, 0xCD, (bytes + 255) / 256 * 256);  // fresh device memory holds anything
    S().allocs[reinterpret_cast<std::uintptr_t>(m)] = bytes;
    *p = m;
    return CS_HIP_OK;
}
int cs_hip_free(void* p) {
    if (!p) return CS_HIP_OK;
    drain();  // hipFree waits for the device
    S().allocs.erase(reinterpret_cast<std::uintptr_t>(p));
    std::free(p);
    return CS_HIP_OK;
}
int cs_hip_host_malloc(void** p, size_t bytes) {
    *p = std::malloc(bytes ? bytes : 1);
    return CS_HIP_OK;
}
int cs_hip_host_free(void* p) {
    std::free(p);
    return CS_HIP_OK;
}
int cs_hip_memcpy_async(void* dst, const void* src, size_t bytes, cs_hip_stream s) {
    if (bytes == 0) return CS_HIP_OK;
    S().stats.copies++;
    if (!device_range(src, bytes)) {
        // a pageable host source is staged when the call is made
        auto copy = std::make_shared<std::vector<unsigned char>>(static_cast<const unsigned char*>(src),
                                                                 static_cast<const unsigned char*>(src) + bytes);
        enqueue(s, [dst, copy] { std::memcpy(dst, copy->data(), copy->size()); });
    } else {
        enqueue(s, [dst, src, bytes] { std::memcpy(dst, src, bytes); });
    }
    return CS_HIP_OK;
}
int cs_hip_fill_u32_async(void* dst, uint32_t value, size_t count, cs_hip_stream s) {
    if (count == 0) return CS_HIP_OK;
    if (!device_range(dst, count * 4)) return fail(CS_HIP_ERR_INVALID, "mock: fill_u32 outside device memory");
    S().stats.fills++;
    enqueue(s, [dst, value, count] {
        auto* p = static_cast<std::uint32_t*>(dst);
        for (size_t i = 0; i < count; ++i) p[i] = value;
    });
    return CS_HIP_OK;
}
int cs_hip_fill_u8_async(void* dst, uint8_t value, size_t count, cs_hip_stream s) {
    if (count == 0) return CS_HIP_OK;
    if (!device_range(dst, count)) return fail(CS_HIP_ERR_INVALID, "mock: fill_u8 outside device memory");
    S().stats.fills++;
    enqueue(s, [dst, value, count] { std::memset(dst, value, count); });
    return CS_HIP_OK;
}

int cs_hip_launch(const void* kernel, const uint32_t grid[3], const uint32_t block[3], size_t, void** args, cs_hip_stream s) {
    auto it = S().kernels.find(kernel);
    if (it == S().kernels.end() && S().fallback) {
        S().stats.launches++;
        mock::KernelFn fn = S().fallback;
        const std::uint32_t g0 = grid[0], g1 = grid[1], g2 = grid[2], b0 = block[0], b1 = block[1], b2 = block[2];
        enqueue(s, [fn, g0, g1, g2, b0, b1, b2] {
            const std::uint32_t gg[3] = {g0, g1, g2}, bb[3] = {b0, b1, b2};
            fn(nullptr, gg, bb);
        });
        return CS_HIP_OK;
    }
    if (it == S().kernels.end()) {
        note_failure("launch of a kernel with no CPU emulation");
        return fail(CS_HIP_ERR_RUNTIME, "mock: no emulation for this kernel");
    }
    S().stats.launches++;
    auto bytes = std::make_shared<std::vector<unsigned char>>(static_cast<const unsigned char*>(args[0]),
                                                              static_cast<const unsigned char*>(args[0]) + it->second.bytes);
    std::uint32_t g[3] = {grid[0], grid[1], grid[2]}, b[3] = {block[0], block[1], block[2]};
    mock::KernelFn fn = it->second.fn;
    enqueue(s, [fn, bytes, g0 = g[0], g1 = g[1], g2 = g[2], b0 = b[0], b1 = b[1], b2 = b[2]] {
        const std::uint32_t gg[3] = {g0, g1, g2}, bb[3] = {b0, b1, b2};
        fn(bytes->data(), gg, bb);
    });
    return CS_HIP_OK;
}
int cs_hip_occupancy(const void*, int, size_t, int* per_cu, int* cus) {
    if (per_cu) *per_cu = 8;
    if (cus) *cus = 8;
    return CS_HIP_OK;
}

int cs_hip_graph_begin(cs_hip_stream) {
    State& st = S();
    if (st.capturing) return fail(CS_HIP_ERR_STATE, "mock: capture already begun");
    drain();
    st.capturing = true;
    st.graph     = new Graph;
    return CS_HIP_OK;
}
int cs_hip_graph_end(cs_hip_stream, cs_hip_graph* g) {
    State& st = S();
    if (!st.capturing) return fail(CS_HIP_ERR_STATE, "mock: no capture");
    st.capturing = false;
    *g           = reinterpret_cast<cs_hip_graph>(st.graph);
    st.graph     = nullptr;
    st.stats.captures++;
    return CS_HIP_OK;
}
int cs_hip_graph_launch(cs_hip_graph g, cs_hip_stream s) {
    S().stats.graph_launches++;
    for (auto& op : reinterpret_cast<Graph*>(g)->ops) enqueue(s, op);
    return CS_HIP_OK;
}
int cs_hip_graph_destroy(cs_hip_graph g) {
    delete reinterpret_cast<Graph*>(g);
    return CS_HIP_OK;
}
const char* cs_hip_build_id(void) { return "mock"; }

// the fault word and spin limit: a host word that emulated kernels OR their bits into, taken after the queues drain
namespace {
std::uint32_t g_fault_word = 0;
std::uint32_t g_spin_limit = CS_HIP_DEFAULT_SPIN_LIMIT;
}  // namespace
int cs_hip_set_spin_limit(uint32_t limit) {
    g_spin_limit = limit;
    return CS_HIP_OK;
}
uint32_t cs_hip_spin_limit(void) { return g_spin_limit; }
int      cs_hip_fault_word(uint32_t** p) {
    if (!p) return fail(CS_HIP_ERR_INVALID, "cs_hip_fault_word: null");
    *p = &g_fault_word;
    return CS_HIP_OK;
}
int cs_hip_take_fault(uint32_t* bits) {
    if (!bits) return fail(CS_HIP_ERR_INVALID, "cs_hip_take_fault: null");
    drain();
    *bits        = g_fault_word;
    g_fault_word = 0;
    return CS_HIP_OK;
}

int cs_hip_stream_capturing(cs_hip_stream, int* c) {
    *c = S().capturing ? 1 : 0;
    return CS_HIP_OK;
}

size_t cs_hip_scan_workspace_bytes(size_t n) { return 16 + 8 * (n == 0 ? 1 : (n + 2047) / 2048); }
int    cs_hip_scan_workspace_init(void* ws, size_t bytes, cs_hip_stream s) { return cs_hip_fill_u8_async(ws, 0, bytes, s); }
int    cs_hip_exclusive_scan_u32(const uint32_t* in, uint32_t* out, size_t n, void* ws, cs_hip_stream s) {
    S().stats.scans++;
    const size_t wsb = cs_hip_scan_workspace_bytes(n);
    enqueue(s, [in, out, n, ws, wsb] {
        if (ws && !all_zero(ws, wsb)) note_failure("exclusive_scan_u32: the workspace is not zero when the scan starts");
        column_counts(in, 1, nullptr, out, n, nullptr);
    });
    return CS_HIP_OK;
}
int cs_hip_exclusive_scan_u32_static(const uint32_t* in, uint32_t* out, size_t n, cs_hip_stream s) {
    if (n > CS_HIP_SCAN_STATIC_MAX_N) return fail(CS_HIP_ERR_INVALID, "cs_hip_exclusive_scan_u32_static: n past the fixed workspace");
    return cs_hip_exclusive_scan_u32(in, out, n, nullptr, s);
}
int cs_hip_build_column(int dt, const void* src, size_t count, size_t k, int offsets_mode, uint32_t* ext, uint32_t* offs, uint32_t* result,
                        void* ws, cs_hip_stream s) {
    if (!offs || !result || (count && (!src || !ext))) return fail(CS_HIP_ERR_INVALID, "cs_hip_build_column: null pointer");
    S().stats.scans++;
    const size_t wsb = cs_hip_scan_workspace_bytes(count);
    if (offsets_mode) {
        if (count == 0) return fail(CS_HIP_ERR_INVALID, "cs_hip_build_column: an offsets source holds no offsets");
        enqueue(s, [=] {
            if (ws && !all_zero(ws, wsb)) note_failure("build_column: the workspace is not zero when the scan starts");
            uint32_t total = 0;
            result[1]      = or_build_column(dt, src, count, 1, 1, ext, offs, &total);
            result[0]      = total;
        });
        return CS_HIP_OK;
    }
    enqueue(s, [=] {
        if (ws && !all_zero(ws, wsb)) note_failure("build_column: the workspace is not zero when the scan starts");
        switch (dt) {
            case 0: column_counts(static_cast<const std::int32_t*>(src), k, ext, offs, count, result); break;
            case 1: column_counts(static_cast<const std::uint32_t*>(src), k, ext, offs, count, result); break;
            case 2: column_counts(static_cast<const std::int64_t*>(src), k, ext, offs, count, result); break;
            default: column_counts(static_cast<const std::uint64_t*>(src), k, ext, offs, count, result); break;
        }
    });
    return CS_HIP_OK;
}

int cs_hip_fill_hash_f32(uint32_t seed, float* out, size_t rows, size_t cols, size_t ld, cs_hip_stream s) {
    if (rows == 0 || cols == 0) return CS_HIP_OK;
    if (!device_range(out, ((rows - 1) * ld + cols) * 4)) return fail(CS_HIP_ERR_INVALID, "mock: fill_hash outside device memory");
    S().stats.fills++;
    enqueue(s, [=] {
        for (size_t i = 0; i < rows; ++i)
            for (size_t j = 0; j < cols; ++j) out[i * ld + j] = hash_unit(seed, static_cast<std::uint64_t>(i) * cols + j);
    });
    return CS_HIP_OK;
}
int cs_hip_quantize_bf16(const float*, uint16_t*, size_t, size_t, size_t, size_t, cs_hip_stream) {
    return fail(CS_HIP_ERR_RUNTIME, "mock: quantize_bf16 is not emulated");
}
int cs_hip_quantize_e4m3_rows(const float*, uint8_t*, float*, size_t, size_t, size_t, size_t, cs_hip_stream) {
    return fail(CS_HIP_ERR_RUNTIME, "mock: quantize_e4m3_rows is not emulated");
}

}  // extern "C"
