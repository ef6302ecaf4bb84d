// cospanmdao_amd/csrc/runtime.hip — libcospan_hip: streams, events, memory, launches and graphs behind a C boundary.
//
// Every entry point returns a status (include/cospan_hip.h) and records a thread-local message on failure; nothing
// throws across the boundary.

#include <hip/hip_runtime.h>

#include <atomic>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "cospan_hip.h"
#include "internal.hpp"

namespace cs_hip {

thread_local std::string g_error;

int fail(int code, const char* fmt, ...) {
    char    buf[512];
    va_list ap;
    va_start(ap, fmt);
    std::vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_error = buf;
    return code;
}

int check(hipError_t e, const char* what) {
    if (e == hipSuccess) return CS_HIP_OK;
    return fail(CS_HIP_ERR_RUNTIME, "%s: %s", what, hipGetErrorString(e));
}

}  // namespace cs_hip

using cs_hip::check;
using cs_hip::fail;

struct cs_hip_stream_s {
    hipStream_t s;
};
struct cs_hip_event_s {
    hipEvent_t e;
};
struct cs_hip_graph_s {
    hipGraph_t     g;
    hipGraphExec_t exec;
};

namespace {

/// The fault word of each device (every device holds its own copy of this variable at its own address). Kernels OR
/// CS_HIP_FAULT_* bits into it when a bounded wait gives up; the host reads and clears it after the work completed.
__device__ uint32_t g_fault_word;

std::atomic<uint32_t>  g_spin_limit{CS_HIP_DEFAULT_SPIN_LIMIT};
std::mutex             g_fault_mu;
std::vector<uint32_t*> g_fault_ptr;  // per device: the address of g_fault_word there

}  // namespace

namespace cs_hip {
hipStream_t raw(cs_hip_stream s) { return s ? s->s : nullptr; }

int fault_word(uint32_t** out) {
    int dev = 0;
    if (int r = check(hipGetDevice(&dev), "hipGetDevice")) return r;
    std::lock_guard<std::mutex> lock{g_fault_mu};
    if (g_fault_ptr.size() <= static_cast<size_t>(dev)) g_fault_ptr.resize(static_cast<size_t>(dev) + 1, nullptr);
    if (!g_fault_ptr[dev]) {
        void* p = nullptr;
        if (int r = check(hipGetSymbolAddress(&p, HIP_SYMBOL(g_fault_word)), "hipGetSymbolAddress(fault word)")) return r;
        g_fault_ptr[dev] = static_cast<uint32_t*>(p);
    }
    *out = g_fault_ptr[dev];
    return CS_HIP_OK;
}
}  // namespace cs_hip

extern "C" {

int cs_hip_last_error(char* buf, size_t len) {
    const std::string& m = cs_hip::g_error;
    if (buf && len) {
        const size_t k = m.size() < len - 1 ? m.size() : len - 1;
        std::memcpy(buf, m.data(), k);
        buf[k] = '\0';
    }
    return static_cast<int>(m.size());
}

int cs_hip_init(int device) {
    int n = 0;
    if (int r = check(hipGetDeviceCount(&n), "hipGetDeviceCount")) return r;
    if (device < 0 || device >= n) return fail(CS_HIP_ERR_INVALID, "cs_hip_init: device %d of %d", device, n);
    return check(hipSetDevice(device), "hipSetDevice");
}

int cs_hip_device_count(int* count) {
    if (!count) return fail(CS_HIP_ERR_INVALID, "cs_hip_device_count: null");
    *count = 0;
    const hipError_t e = hipGetDeviceCount(count);
    if (e == hipErrorNoDevice) {
        *count = 0;
        return CS_HIP_OK;
    }
    return check(e, "hipGetDeviceCount");
}

int cs_hip_device_info(int device, char* name, size_t name_len, int* cus, int* lds_bytes, int* max_wg) {
    hipDeviceProp_t p;
    if (int r = check(hipGetDeviceProperties(&p, device), "hipGetDeviceProperties")) return r;
    if (name && name_len) {
        std::snprintf(name, name_len, "%s", p.gcnArchName);
    }
    if (cus) *cus = p.multiProcessorCount;
    if (lds_bytes) *lds_bytes = static_cast<int>(p.sharedMemPerBlock);
    if (max_wg) *max_wg = p.maxThreadsPerBlock;
    return CS_HIP_OK;
}

int cs_hip_device_synchronize(void) { return check(hipDeviceSynchronize(), "hipDeviceSynchronize"); }

int cs_hip_stream_create(cs_hip_stream* s) {
    if (!s) return fail(CS_HIP_ERR_INVALID, "cs_hip_stream_create: null");
    hipStream_t h;
    // default priority, non-blocking with respect to the null stream
    if (int r = check(hipStreamCreateWithFlags(&h, hipStreamNonBlocking), "hipStreamCreateWithFlags")) return r;
    *s = new cs_hip_stream_s{h};
    return CS_HIP_OK;
}

int cs_hip_stream_destroy(cs_hip_stream s) {
    if (!s) return CS_HIP_OK;
    const int r = check(hipStreamDestroy(s->s), "hipStreamDestroy");
    delete s;
    return r;
}

int cs_hip_stream_synchronize(cs_hip_stream s) { return check(hipStreamSynchronize(cs_hip::raw(s)), "hipStreamSynchronize"); }

int cs_hip_event_create(cs_hip_event* e, int timing) {
    if (!e) return fail(CS_HIP_ERR_INVALID, "cs_hip_event_create: null");
    hipEvent_t h;
    if (int r = check(hipEventCreateWithFlags(&h, timing ? hipEventDefault : hipEventDisableTiming), "hipEventCreateWithFlags"))
        return r;
    *e = new cs_hip_event_s{h};
    return CS_HIP_OK;
}

int cs_hip_event_destroy(cs_hip_event e) {
    if (!e) return CS_HIP_OK;
    const int r = check(hipEventDestroy(e->e), "hipEventDestroy");
    delete e;
    return r;
}

int cs_hip_event_record(cs_hip_event e, cs_hip_stream s) {
    if (!e) return fail(CS_HIP_ERR_INVALID, "cs_hip_event_record: null event");
    return check(hipEventRecord(e->e, cs_hip::raw(s)), "hipEventRecord");
}

int cs_hip_stream_wait_event(cs_hip_stream s, cs_hip_event e) {
    if (!e) return fail(CS_HIP_ERR_INVALID, "cs_hip_stream_wait_event: null event");
    return check(hipStreamWaitEvent(cs_hip::raw(s), e->e, 0), "hipStreamWaitEvent");
}

int cs_hip_event_query(cs_hip_event e, int* done) {
    if (!e || !done) return fail(CS_HIP_ERR_INVALID, "cs_hip_event_query: null");
    const hipError_t r = hipEventQuery(e->e);
    if (r == hipSuccess) {
        *done = 1;
        return CS_HIP_OK;
    }
    if (r == hipErrorNotReady) {
        *done = 0;
        return CS_HIP_OK;
    }
    return check(r, "hipEventQuery");
}

int cs_hip_event_synchronize(cs_hip_event e) {
    if (!e) return fail(CS_HIP_ERR_INVALID, "cs_hip_event_synchronize: null");
    return check(hipEventSynchronize(e->e), "hipEventSynchronize");
}

int cs_hip_event_elapsed_ms(float* ms, cs_hip_event start, cs_hip_event stop) {
    if (!ms || !start || !stop) return fail(CS_HIP_ERR_INVALID, "cs_hip_event_elapsed_ms: null");
    return check(hipEventElapsedTime(ms, start->e, stop->e), "hipEventElapsedTime");
}

int cs_hip_malloc(void** p, size_t bytes) {
    if (!p) return fail(CS_HIP_ERR_INVALID, "cs_hip_malloc: null");
    *p = nullptr;
    if (bytes == 0) return CS_HIP_OK;
    return check(hipMalloc(p, bytes), "hipMalloc");
}

int cs_hip_free(void* p) { return p ? check(hipFree(p), "hipFree") : CS_HIP_OK; }

int cs_hip_host_malloc(void** p, size_t bytes) {
    if (!p) return fail(CS_HIP_ERR_INVALID, "cs_hip_host_malloc: null");
    *p = nullptr;
    if (bytes == 0) return CS_HIP_OK;
    return check(hipHostMalloc(p, bytes, hipHostMallocDefault), "hipHostMalloc");
}

int cs_hip_host_free(void* p) { return p ? check(hipHostFree(p), "hipHostFree") : CS_HIP_OK; }

int cs_hip_memcpy_async(void* dst, const void* src, size_t bytes, cs_hip_stream s) {
    if (bytes == 0) return CS_HIP_OK;
    if (!dst || !src) return fail(CS_HIP_ERR_INVALID, "cs_hip_memcpy_async: null pointer");
    return check(hipMemcpyAsync(dst, src, bytes, hipMemcpyDefault, cs_hip::raw(s)), "hipMemcpyAsync");
}

int cs_hip_launch(const void* kernel, const uint32_t grid[3], const uint32_t block[3], size_t lds_bytes, void** args,
                  cs_hip_stream s) {
    if (!kernel || !grid || !block) return fail(CS_HIP_ERR_INVALID, "cs_hip_launch: null");
    if (grid[0] == 0 || grid[1] == 0 || grid[2] == 0) return CS_HIP_OK;  // an empty launch submits nothing
    if (grid[1] > 65535u || grid[2] > 65535u)
        return fail(CS_HIP_ERR_INVALID, "cs_hip_launch: grid (%u, %u, %u) exceeds 65535 in y or z", grid[0], grid[1], grid[2]);
    if (static_cast<uint64_t>(block[0]) * block[1] * block[2] > 1024u)
        return fail(CS_HIP_ERR_INVALID, "cs_hip_launch: block (%u, %u, %u) exceeds 1024 work-items", block[0], block[1], block[2]);
    if (lds_bytes > 160u * 1024u)
        return fail(CS_HIP_ERR_INVALID, "cs_hip_launch: %zu bytes of LDS exceed the 160 KiB of a CDNA4 work-group", lds_bytes);
    return check(hipLaunchKernel(kernel, dim3(grid[0], grid[1], grid[2]), dim3(block[0], block[1], block[2]), args, lds_bytes,
                                 cs_hip::raw(s)),
                 "hipLaunchKernel");
}

int cs_hip_occupancy(const void* kernel, int block_threads, size_t lds_bytes, int* groups_per_cu, int* cus) {
    if (!kernel || !groups_per_cu) return fail(CS_HIP_ERR_INVALID, "cs_hip_occupancy: null");
    if (int r = check(hipOccupancyMaxActiveBlocksPerMultiprocessor(groups_per_cu, kernel, block_threads, lds_bytes),
                      "hipOccupancyMaxActiveBlocksPerMultiprocessor"))
        return r;
    if (cus) {
        int dev = 0;
        if (int r = check(hipGetDevice(&dev), "hipGetDevice")) return r;
        if (int r = check(hipDeviceGetAttribute(cus, hipDeviceAttributeMultiprocessorCount, dev), "hipDeviceGetAttribute")) return r;
    }
    return CS_HIP_OK;
}

int cs_hip_graph_begin(cs_hip_stream s) {
    if (!s) return fail(CS_HIP_ERR_INVALID, "cs_hip_graph_begin: capture needs a created stream");
    return check(hipStreamBeginCapture(s->s, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture");
}

int cs_hip_graph_end(cs_hip_stream s, cs_hip_graph* g) {
    if (!s || !g) return fail(CS_HIP_ERR_INVALID, "cs_hip_graph_end: null");
    hipGraph_t graph = nullptr;
    if (int r = check(hipStreamEndCapture(s->s, &graph), "hipStreamEndCapture")) return r;
    hipGraphExec_t exec = nullptr;
    if (int r = check(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0), "hipGraphInstantiate")) {
        (void)hipGraphDestroy(graph);
        return r;
    }
    *g = new cs_hip_graph_s{graph, exec};
    return CS_HIP_OK;
}

int cs_hip_graph_launch(cs_hip_graph g, cs_hip_stream s) {
    if (!g) return fail(CS_HIP_ERR_INVALID, "cs_hip_graph_launch: null graph");
    return check(hipGraphLaunch(g->exec, cs_hip::raw(s)), "hipGraphLaunch");
}

int cs_hip_graph_destroy(cs_hip_graph g) {
    if (!g) return CS_HIP_OK;
    int r = check(hipGraphExecDestroy(g->exec), "hipGraphExecDestroy");
    if (int r2 = check(hipGraphDestroy(g->g), "hipGraphDestroy")) r = r ? r : r2;
    delete g;
    return r;
}

int cs_hip_set_spin_limit(uint32_t limit) {
    g_spin_limit.store(limit, std::memory_order_relaxed);
    return CS_HIP_OK;
}

uint32_t cs_hip_spin_limit(void) { return g_spin_limit.load(std::memory_order_relaxed); }

int cs_hip_fault_word(uint32_t** device_word) {
    if (!device_word) return fail(CS_HIP_ERR_INVALID, "cs_hip_fault_word: null");
    return cs_hip::fault_word(device_word);
}

int cs_hip_take_fault(uint32_t* bits) {
    if (!bits) return fail(CS_HIP_ERR_INVALID, "cs_hip_take_fault: null");
    uint32_t* p = nullptr;
    if (int r = cs_hip::fault_word(&p)) return r;
    uint32_t v = 0;
    if (int r = check(hipMemcpy(&v, p, sizeof v, hipMemcpyDeviceToHost), "hipMemcpy(fault word)")) return r;
    if (v != 0) {
        const uint32_t zero = 0;
        if (int r = check(hipMemcpy(p, &zero, sizeof zero, hipMemcpyHostToDevice), "hipMemcpy(fault word)")) return r;
    }
    *bits = v;
    return CS_HIP_OK;
}

int cs_hip_stream_capturing(cs_hip_stream s, int* capturing) {
    if (!capturing) return fail(CS_HIP_ERR_INVALID, "cs_hip_stream_capturing: null");
    hipStreamCaptureStatus st = hipStreamCaptureStatusNone;
    if (int r = check(hipStreamIsCapturing(cs_hip::raw(s), &st), "hipStreamIsCapturing")) return r;
    *capturing = st == hipStreamCaptureStatusActive ? 1 : 0;
    return CS_HIP_OK;
}

}  // extern "C"
