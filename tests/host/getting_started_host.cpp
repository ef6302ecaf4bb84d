// tests/host/getting_started_host.cpp — examples/getting_started.cpp (the reference's one-component demo) on the CPU
// boundary of tests/host/mock_hip.cpp (TEST INFRASTRUCTURE ONLY). The demo's source is included as it is, main and all;
// a static initializer placed after it registers the emulation of its one kernel (the index-range group of Axpb), which
// calls the user's own compute_forward on the host for every index of the launch (the CS_COMPONENTS region makes it a
// host function too), and checks at exit that the CPU boundary saw no failure. The demo prints y.
#include "../../examples/getting_started.cpp"

#include <cstdio>
#include <cstdlib>

#include "mock_hip.hpp"

namespace {

using GAxpb = cs::detail::GroupArgs<Axpb>;

void check_mock_at_exit() {
    const std::string f = mock::failure();
    if (!f.empty()) {
        std::printf("mock: %s\n", f.c_str());
        std::fflush(stdout);
        std::_Exit(1);
    }
}

struct RegisterEmulation {
    RegisterEmulation() {
        mock::register_kernel(reinterpret_cast<const void*>(&cs::detail::index_kernel<Axpb>), sizeof(GAxpb),
                              [](const void* bytes, const std::uint32_t*, const std::uint32_t*) {
                                  const auto& a = static_cast<const GAxpb*>(bytes)->args.head;
                                  for (std::size_t i = 0; i < a.units[2]; ++i) Axpb::compute_forward(cs::id<1>{i}, a.in, a.out);
                              });
        std::atexit(check_mock_at_exit);
    }
} g_register;

}  // namespace
