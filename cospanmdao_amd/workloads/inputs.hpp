#pragma once

// =============================================================================
// cospanmdao_amd/workloads/inputs.hpp — deterministic synthetic inputs of the example models
// =============================================================================
//
// The counter hash of the reference's example data (examples/gemver-chain/common/chain_data.hpp:34-49: two mix32 rounds of
// (seed, index) giving 24 bits, exact in f32), usable on host and device, and the Fourier-mode fields of the NS residual:
// each of u, v, p is the sum of 8 modes a_m / |k_m|^2 cos(2 pi (kx x + ky y) + phi_m), integer k in [-4, 4]^2 \ {0},
// a_m ~ N(0, 1) (Box-Muller), phi_m ~ U[0, 2 pi), every draw from the hash (seed, field * 64 + m * 8 + t). The mode
// tables (per column cos/sin of 2 pi kx x, per row cos/sin of 2 pi ky y + phi) are computed on the host in FP64; the
// device sums a_m (cx cy - sx sy) in FP64 in mode order with no contraction, so the fields equal a host evaluation of the
// same formula bit for bit.

#include <cmath>
#include <cstdint>
#include <vector>

#include <hip/hip_runtime.h>

namespace inputs {

__host__ __device__ inline std::uint32_t mix32(std::uint32_t x) {
    x ^= x >> 16;
    x *= 0x7feb352du;
    x ^= x >> 15;
    x *= 0x846ca68bu;
    x ^= x >> 16;
    return x;
}

__host__ __device__ inline float hash_unit(std::uint32_t seed, std::uint64_t index) {
    const std::uint32_t lo = static_cast<std::uint32_t>(index), hi = static_cast<std::uint32_t>(index >> 32);
    const std::uint32_t h  = mix32(mix32(seed * 0x9e3779b9u + 0x85ebca6bu) ^ mix32(lo + mix32(hi + seed)));
    return static_cast<float>(h >> 8) * (1.0f / 8388608.0f) - 1.0f;
}

struct Mode {
    int    kx, ky;
    double amp, phi;
};

inline Mode ns_mode(std::uint32_t seed, int field, int m) {
#pragma clang fp contract(off)
    const double       pi = 3.14159265358979323846;
    const std::uint64_t b = static_cast<std::uint64_t>(field) * 64u + static_cast<std::uint64_t>(m) * 8u;
    Mode               md;
    md.kx = static_cast<int>(std::floor((hash_unit(seed, b + 0) + 1.0f) * 4.5f)) - 4;
    md.ky = static_cast<int>(std::floor((hash_unit(seed, b + 1) + 1.0f) * 4.5f)) - 4;
    if (md.kx == 0 && md.ky == 0) md.kx = 1;
    const double u1 = 1.0 - (static_cast<double>(hash_unit(seed, b + 2)) + 1.0) * 0.5;
    const double u2 = (static_cast<double>(hash_unit(seed, b + 3)) + 1.0) * 0.5;
    const double g  = std::sqrt(-2.0 * std::log(u1)) * std::cos(2.0 * pi * u2);
    md.amp          = g / static_cast<double>(md.kx * md.kx + md.ky * md.ky);
    md.phi          = 2.0 * pi * ((static_cast<double>(hash_unit(seed, b + 4)) + 1.0) * 0.5);
    return md;
}

/// The mode tables of one field: amp[8], then cx, sx (8 x nx), then cy, sy (8 x ny), as one FP64 array.
inline std::vector<double> ns_tables(std::uint32_t seed, int field, int nx, int ny) {
#pragma clang fp contract(off)
    const double        pi = 3.14159265358979323846;
    std::vector<double> t(8 + 16 * static_cast<std::size_t>(nx) + 16 * static_cast<std::size_t>(ny));
    double*             amp = t.data();
    double*             cx  = amp + 8;
    double*             sx  = cx + 8 * static_cast<std::size_t>(nx);
    double*             cy  = sx + 8 * static_cast<std::size_t>(nx);
    double*             sy  = cy + 8 * static_cast<std::size_t>(ny);
    for (int m = 0; m < 8; ++m) {
        const Mode md = ns_mode(seed, field, m);
        amp[m]        = md.amp;
        for (int j = 0; j < nx; ++j) {
            const double a = 2.0 * pi * static_cast<double>(md.kx) * static_cast<double>(j) / static_cast<double>(nx);
            cx[m * nx + j] = std::cos(a);
            sx[m * nx + j] = std::sin(a);
        }
        for (int i = 0; i < ny; ++i) {
            const double a = 2.0 * pi * static_cast<double>(md.ky) * static_cast<double>(i) / static_cast<double>(ny) + md.phi;
            cy[m * ny + i] = std::cos(a);
            sy[m * ny + i] = std::sin(a);
        }
    }
    return t;
}

}  // namespace inputs
