// cospanmdao_amd/csrc/fill.hip — fills, the hash initializer and the reduced-precision quantizers of libcospan_hip.
//
// Fills are kernels (never memset nodes): a captured graph that clears counters replays correctly.
// fill_hash_f32 is the device form of the host generators of examples/gemver-chain/common/chain_data.hpp:34-63: the same
// integer hash, the same exact float conversion, so device-generated inputs equal the host's bit for bit.

#include <hip/hip_runtime.h>

#include <cstdint>

#include "cospan_hip.h"
#include "internal.hpp"
#include "naga/cospan/device.hpp"
#include "naga/cospan/dtype.hpp"

namespace {

namespace cs = naga::cospan;

__global__ void fill_u32_kernel(uint32_t* p, uint32_t v, uint64_t n) {
    const uint64_t stride = static_cast<uint64_t>(gridDim.x) * blockDim.x;
    for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride) p[i] = v;
}

__global__ void fill_u8_kernel(uint8_t* p, uint8_t v, uint64_t n) {
    const uint64_t stride = static_cast<uint64_t>(gridDim.x) * blockDim.x;
    for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride) p[i] = v;
}

__device__ inline uint32_t mix32(uint32_t x) {
    x ^= x >> 16;
    x *= 0x7feb352du;
    x ^= x >> 15;
    x *= 0x846ca68bu;
    x ^= x >> 16;
    return x;
}

__device__ inline float hash_unit(uint32_t seed, uint64_t index) {
    const uint32_t lo = static_cast<uint32_t>(index), hi = static_cast<uint32_t>(index >> 32);
    const uint32_t h  = mix32(mix32(seed * 0x9e3779b9u + 0x85ebca6bu) ^ mix32(lo + mix32(hi + seed)));
    return static_cast<float>(h >> 8) * (1.0f / 8388608.0f) - 1.0f;
}

/// grid.x covers the columns, grid.y strides over the rows
__global__ void fill_hash_kernel(uint32_t seed, float* out, uint64_t rows, uint64_t cols, uint64_t ld) {
    const uint64_t j = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (j >= cols) return;
    for (uint64_t i = blockIdx.y; i < rows; i += gridDim.y) out[i * ld + j] = hash_unit(seed, i * cols + j);
}

__global__ void quantize_bf16_kernel(const float* in, uint16_t* out, uint64_t rows, uint64_t cols, uint64_t ld_in, uint64_t ld_out) {
    const uint64_t j = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (j >= cols) return;
    for (uint64_t i = blockIdx.y; i < rows; i += gridDim.y) out[i * ld_out + j] = cs::bf16::from_float(in[i * ld_in + j]).bits;
}

/// OCP E4M3, round to nearest even, saturating at +-448 (the algorithm of the CPU quantizer: the magnitude in units of its
/// binade's spacing, rounded by rint, re-encoded).
__device__ inline uint8_t to_e4m3(float f) {
    const uint32_t u    = __float_as_uint(f);
    const uint8_t  sign = static_cast<uint8_t>((u >> 31) << 7);
    if (__builtin_isnan(f)) return static_cast<uint8_t>(sign | 0x7f);
    const float a = fabsf(f);
    if (a >= 448.0f) return static_cast<uint8_t>(sign | 0x7e);
    if (a == 0.0f) return sign;
    int ex;
    (void)frexpf(a, &ex);
    int e_unb = ex - 1;
    if (e_unb < -6) e_unb = -6;
    const float spacing = ldexpf(1.0f, e_unb - 3);
    const float r       = rintf(a / spacing);
    const float mag     = r * spacing;
    if (mag >= 448.0f) return static_cast<uint8_t>(sign | 0x7e);
    if (mag == 0.0f) return sign;
    int ex2;
    (void)frexpf(mag, &ex2);
    const int e2 = ex2 - 1;
    if (e2 < -6) return static_cast<uint8_t>(sign | static_cast<uint32_t>(mag / ldexpf(1.0f, -9)));
    const uint32_t m = static_cast<uint32_t>((mag / ldexpf(1.0f, e2) - 1.0f) * 8.0f);
    return static_cast<uint8_t>(sign | static_cast<uint8_t>((e2 + 7) << 3) | static_cast<uint8_t>(m));
}

/// one work-group per row: the row's max |a| (exact: max is order independent), then q = e4m3(a / scale)
__global__ __launch_bounds__(256) void quantize_e4m3_kernel(const float* in, uint8_t* out, float* scale, uint64_t rows,
                                                            uint64_t cols, uint64_t ld_in, uint64_t ld_out) {
    for (uint64_t i = blockIdx.x; i < rows; i += gridDim.x) {
        const float* a    = in + i * ld_in;
        float        amax = 0.0f;
        for (uint64_t j = threadIdx.x; j < cols; j += blockDim.x) amax = fmaxf(amax, fabsf(a[j]));
        amax          = cs::block::reduce(amax, cs::maximum<float>{});
        const float s = amax > 0.0f ? __fdiv_rn(amax, 448.0f) : 1.0f;
        if (threadIdx.x == 0) scale[i] = s;
        for (uint64_t j = threadIdx.x; j < cols; j += blockDim.x) out[i * ld_out + j] = to_e4m3(__fdiv_rn(a[j], s));
    }
}

uint32_t grid_for(uint64_t n, uint32_t threads) {
    const uint64_t g = (n + threads - 1) / threads;
    return static_cast<uint32_t>(g < 65536 ? (g ? g : 1) : 65536);
}

}  // namespace

extern "C" {

int cs_hip_fill_u32_async(void* dst, uint32_t value, size_t count, cs_hip_stream s) {
    if (count == 0) return CS_HIP_OK;
    if (!dst) return cs_hip::fail(CS_HIP_ERR_INVALID, "cs_hip_fill_u32_async: null");
    hipLaunchKernelGGL(fill_u32_kernel, dim3(grid_for(count, 256)), dim3(256), 0, cs_hip::raw(s), static_cast<uint32_t*>(dst), value,
                       static_cast<uint64_t>(count));
    return cs_hip::check(hipGetLastError(), "fill_u32 launch");
}

int cs_hip_fill_u8_async(void* dst, uint8_t value, size_t count, cs_hip_stream s) {
    if (count == 0) return CS_HIP_OK;
    if (!dst) return cs_hip::fail(CS_HIP_ERR_INVALID, "cs_hip_fill_u8_async: null");
    hipLaunchKernelGGL(fill_u8_kernel, dim3(grid_for(count, 256)), dim3(256), 0, cs_hip::raw(s), static_cast<uint8_t*>(dst), value,
                       static_cast<uint64_t>(count));
    return cs_hip::check(hipGetLastError(), "fill_u8 launch");
}

int cs_hip_fill_hash_f32(uint32_t seed, float* out, size_t rows, size_t cols, size_t ld, cs_hip_stream s) {
    if (rows == 0 || cols == 0) return CS_HIP_OK;
    if (!out || ld < cols) return cs_hip::fail(CS_HIP_ERR_INVALID, "cs_hip_fill_hash_f32: null output or ld < cols");
    const dim3 grid(static_cast<uint32_t>((cols + 255) / 256), static_cast<uint32_t>(rows < 1024 ? rows : 1024));
    hipLaunchKernelGGL(fill_hash_kernel, grid, dim3(256), 0, cs_hip::raw(s), seed, out, static_cast<uint64_t>(rows),
                       static_cast<uint64_t>(cols), static_cast<uint64_t>(ld));
    return cs_hip::check(hipGetLastError(), "fill_hash launch");
}

int cs_hip_quantize_bf16(const float* in, uint16_t* out, size_t rows, size_t cols, size_t ld_in, size_t ld_out, cs_hip_stream s) {
    if (rows == 0 || cols == 0) return CS_HIP_OK;
    if (!in || !out || ld_in < cols || ld_out < cols) return cs_hip::fail(CS_HIP_ERR_INVALID, "cs_hip_quantize_bf16: bad arguments");
    const dim3 grid(static_cast<uint32_t>((cols + 255) / 256), static_cast<uint32_t>(rows < 1024 ? rows : 1024));
    hipLaunchKernelGGL(quantize_bf16_kernel, grid, dim3(256), 0, cs_hip::raw(s), in, out, static_cast<uint64_t>(rows),
                       static_cast<uint64_t>(cols), static_cast<uint64_t>(ld_in), static_cast<uint64_t>(ld_out));
    return cs_hip::check(hipGetLastError(), "quantize_bf16 launch");
}

int cs_hip_quantize_e4m3_rows(const float* in, uint8_t* out, float* scale, size_t rows, size_t cols, size_t ld_in, size_t ld_out,
                              cs_hip_stream s) {
    if (rows == 0) return CS_HIP_OK;
    if (!in || !out || !scale || ld_in < cols || ld_out < cols)
        return cs_hip::fail(CS_HIP_ERR_INVALID, "cs_hip_quantize_e4m3_rows: bad arguments");
    const uint32_t groups = static_cast<uint32_t>(rows < 65536 ? rows : 65536);
    hipLaunchKernelGGL(quantize_e4m3_kernel, dim3(groups), dim3(256), 0, cs_hip::raw(s), in, out, scale, static_cast<uint64_t>(rows),
                       static_cast<uint64_t>(cols), static_cast<uint64_t>(ld_in), static_cast<uint64_t>(ld_out));
    return cs_hip::check(hipGetLastError(), "quantize_e4m3 launch");
}

}  // extern "C"
