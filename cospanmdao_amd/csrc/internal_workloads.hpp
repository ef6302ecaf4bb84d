// cospanmdao_amd/csrc/internal_workloads.hpp — shared by the workload libraries (not installed): device buffers owned by a
// model handle, exception-to-status mapping at the C boundary, and launch timing with device events.
#pragma once

#include <functional>
#include <stdexcept>
#include <string>
#include <vector>

#include "cospan_hip.h"
#include "naga/cospan/group_launch.hpp"

namespace cs_hip {
int fail(int code, const char* fmt, ...) __attribute__((format(printf, 2, 3)));
}

namespace wl {

/// Runs `f`; an exception becomes a status with its message (the C boundary never throws).
inline int guard(const std::function<void()>& f) {
    try {
        f();
        return CS_HIP_OK;
    } catch (const std::invalid_argument& e) {
        return cs_hip::fail(CS_HIP_ERR_INVALID, "%s", e.what());
    } catch (const std::logic_error& e) {
        return cs_hip::fail(CS_HIP_ERR_STATE, "%s", e.what());
    } catch (const std::exception& e) {
        return cs_hip::fail(CS_HIP_ERR_RUNTIME, "%s", e.what());
    }
}

/// Device allocations freed with the owner.
class DeviceBuffers {
public:
    DeviceBuffers() = default;
    DeviceBuffers(const DeviceBuffers&)            = delete;
    DeviceBuffers& operator=(const DeviceBuffers&) = delete;
    ~DeviceBuffers() {
        for (void* p : ptrs_) cs_hip_free(p);
    }
    void* alloc(std::size_t bytes) {
        void* p = nullptr;
        naga::cospan::detail::hip_check(cs_hip_malloc(&p, bytes < 16 ? 16 : bytes), "allocation");
        ptrs_.push_back(p);
        return p;
    }
    float* alloc_f32(std::size_t n) { return static_cast<float*>(alloc(n * 4)); }

private:
    std::vector<void*> ptrs_;
};

/// Two timing events.
class EventPair {
public:
    EventPair() {
        naga::cospan::detail::hip_check(cs_hip_event_create(&a_, 1), "event");
        naga::cospan::detail::hip_check(cs_hip_event_create(&b_, 1), "event");
    }
    ~EventPair() {
        cs_hip_event_destroy(a_);
        cs_hip_event_destroy(b_);
    }
    void  record_start(cs_hip_stream s) { naga::cospan::detail::hip_check(cs_hip_event_record(a_, s), "event"); }
    void  record_stop(cs_hip_stream s) { naga::cospan::detail::hip_check(cs_hip_event_record(b_, s), "event"); }
    float elapsed_ms() {
        naga::cospan::detail::hip_check(cs_hip_event_synchronize(b_), "event");
        float ms = 0.0f;
        naga::cospan::detail::hip_check(cs_hip_event_elapsed_ms(&ms, a_, b_), "event");
        return ms;
    }

private:
    cs_hip_event a_ = nullptr, b_ = nullptr;
};

/// A timing event recorded between launches of one forward.
struct LaunchMark {
    cs_hip_event e = nullptr;
};

inline void mark(std::vector<LaunchMark>* marks, cs_hip_stream s) {
    if (!marks) return;
    LaunchMark m;
    naga::cospan::detail::hip_check(cs_hip_event_create(&m.e, 1), "event");
    naga::cospan::detail::hip_check(cs_hip_event_record(m.e, s), "event");
    marks->push_back(m);
}

/// The milliseconds between consecutive marks (the launches between them), after which the marks are destroyed.
inline std::vector<float> durations(std::vector<LaunchMark>& marks) {
    std::vector<float> d;
    if (!marks.empty()) naga::cospan::detail::hip_check(cs_hip_event_synchronize(marks.back().e), "event");
    for (std::size_t k = 1; k < marks.size(); ++k) {
        float ms = 0.0f;
        naga::cospan::detail::hip_check(cs_hip_event_elapsed_ms(&ms, marks[k - 1].e, marks[k].e), "event");
        d.push_back(ms);
    }
    for (LaunchMark& m : marks) cs_hip_event_destroy(m.e);
    marks.clear();
    return d;
}

}  // namespace wl
