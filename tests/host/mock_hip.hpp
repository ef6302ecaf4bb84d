// tests/host/mock_hip.hpp — a CPU implementation of include/cospan_hip.h for host-logic tests (TEST INFRASTRUCTURE ONLY).
//
// Device memory is host memory filled with a garbage byte at allocation (as fresh device memory may hold anything).
// Streams are queues that run only when something synchronises, in an adversarial order (the highest-numbered runnable
// stream first), blocked only by the event waits the host code asked for: a launch that misses a dependency reads data
// that is not written yet. Graph capture records the work of every stream and a replay runs it in recorded order.
// Kernels run through CPU emulations registered per host stub; the scan and the column build run on the CPU and check
// that their workspace is zero when they start (the contract of the real ones).
#pragma once

#include <cstddef>
#include <cstdint>
#include <functional>
#include <string>

namespace mock {

using KernelFn = std::function<void(const void* args, const std::uint32_t grid[3], const std::uint32_t block[3])>;

/// Registers the CPU emulation of the kernel whose host stub is `stub`; its one argument has `arg_bytes` bytes.
void register_kernel(const void* stub, std::size_t arg_bytes, KernelFn fn);

/// Launches of kernels with no registered emulation run `fn` instead (with null argument bytes) rather than failing:
/// for tests of host logic only (wiring, launch counts), whose kernels' results are not looked at.
void set_fallback_kernel(KernelFn fn);

/// True when p lies inside a live mock device allocation.
bool is_device(const void* p);

struct Stats {
    std::uint64_t launches = 0, graph_launches = 0, captures = 0, scans = 0, fills = 0, copies = 0;
};
Stats stats();

/// The first failure a mock check saw (empty when none): a non-zero scan workspace, an unknown kernel, ...
std::string failure();

}  // namespace mock
