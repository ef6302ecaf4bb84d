#pragma once

// =============================================================================
// naga/cospan/device.hpp — wave64 and work-group primitives for CDNA4 (gfx950)
// =============================================================================
//
// Wave-level: DPP inclusive scans (row_shr 1/2/4/8 inside each row of 16 lanes, then row_bcast 15 and 31 across rows),
// xor-butterfly reductions (a balanced pairwise tree over lanes in lane order: every lane ends with the same value),
// ballot and readlane. Group-level: the same composed through LDS, one slot per wave.
//
// These are the building blocks of the group functions of kernel.hpp (cs::reduce_over_group, ...), of the device scan and
// of the hand-written kernels.

#include <hip/hip_runtime.h>

#include <cstdint>
#include <type_traits>

#include "naga/cospan/vocab.hpp"

namespace naga::cospan::wave {

inline constexpr int kSize = 64;

__device__ inline int lane_id() { return static_cast<int>(__lane_id()); }
__device__ inline std::uint64_t ballot(bool p) { return __ballot(p); }

/// One DPP move of a 32-bit pattern: lanes whose source is outside the row (or in a row masked off) keep `old`.
template <int Ctrl, int RowMask = 0xf, int BankMask = 0xf>
__device__ inline std::uint32_t dpp32(std::uint32_t old, std::uint32_t x) {
    return static_cast<std::uint32_t>(
        __builtin_amdgcn_update_dpp(static_cast<int>(old), static_cast<int>(x), Ctrl, RowMask, BankMask, false));
}

/// The same for any 4- or 8-byte trivially copyable value.
template <int Ctrl, int RowMask = 0xf, int BankMask = 0xf, class T>
__device__ inline T dpp(T old, T x) {
    static_assert(sizeof(T) == 4 || sizeof(T) == 8, "cospan: DPP moves 4- or 8-byte values");
    if constexpr (sizeof(T) == 4) {
        return __builtin_bit_cast(T, dpp32<Ctrl, RowMask, BankMask>(__builtin_bit_cast(std::uint32_t, old),
                                                                     __builtin_bit_cast(std::uint32_t, x)));
    } else {
        const std::uint64_t o = __builtin_bit_cast(std::uint64_t, old), v = __builtin_bit_cast(std::uint64_t, x);
        const std::uint32_t lo = dpp32<Ctrl, RowMask, BankMask>(static_cast<std::uint32_t>(o), static_cast<std::uint32_t>(v));
        const std::uint32_t hi =
            dpp32<Ctrl, RowMask, BankMask>(static_cast<std::uint32_t>(o >> 32), static_cast<std::uint32_t>(v >> 32));
        return __builtin_bit_cast(T, (static_cast<std::uint64_t>(hi) << 32) | lo);
    }
}

// DPP controls (GFX9 encoding)
inline constexpr int kRowShr1 = 0x111, kRowShr2 = 0x112, kRowShr4 = 0x114, kRowShr8 = 0x118;
inline constexpr int kRowBcast15 = 0x142, kRowBcast31 = 0x143;

/// Inclusive scan over the 64 lanes in lane order (Hillis-Steele inside rows of 16, then the row broadcasts).
template <class T, class Op>
__device__ inline T inclusive_scan(T x, Op op, T identity) {
    x = op(x, dpp<kRowShr1>(identity, x));
    x = op(x, dpp<kRowShr2>(identity, x));
    x = op(x, dpp<kRowShr4>(identity, x));
    x = op(x, dpp<kRowShr8>(identity, x));
    x = op(x, dpp<kRowBcast15, 0xa>(identity, x));
    x = op(x, dpp<kRowBcast31, 0xc>(identity, x));
    return x;
}

template <class T>
__device__ inline T readlane(T x, int l) {
    static_assert(sizeof(T) == 4 || sizeof(T) == 8, "cospan: readlane moves 4- or 8-byte values");
    if constexpr (sizeof(T) == 4) {
        return __builtin_bit_cast(T, __builtin_amdgcn_readlane(__builtin_bit_cast(int, x), l));
    } else {
        const std::uint64_t v  = __builtin_bit_cast(std::uint64_t, x);
        const std::uint32_t lo = static_cast<std::uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(v), l));
        const std::uint32_t hi = static_cast<std::uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(v >> 32), l));
        return __builtin_bit_cast(T, (static_cast<std::uint64_t>(hi) << 32) | lo);
    }
}

/// x of lane `src` (any trivially copyable value, moved as 32-bit words through ds_bpermute). `src` must be an active lane.
template <class T>
__device__ inline T shfl(T x, int src) {
    static_assert(std::is_trivially_copyable_v<T>, "cospan: shuffles move trivially copyable values");
    constexpr int kWords = static_cast<int>((sizeof(T) + 3) / 4);
    std::uint32_t w[kWords] = {};
    __builtin_memcpy(w, &x, sizeof(T));
    for (int k = 0; k < kWords; ++k)
        w[k] = static_cast<std::uint32_t>(__builtin_amdgcn_ds_bpermute(src << 2, static_cast<int>(w[k])));
    T r;
    __builtin_memcpy(&r, w, sizeof(T));
    return r;
}

/// x of lane (lane ^ mask), through ds_bpermute.
template <class T>
__device__ inline T shfl_xor(T x, int mask) {
    static_assert(sizeof(T) == 4 || sizeof(T) == 8, "cospan: shuffles move 4- or 8-byte values");
    const int src = (lane_id() ^ mask) << 2;
    if constexpr (sizeof(T) == 4) {
        return __builtin_bit_cast(T, __builtin_amdgcn_ds_bpermute(src, __builtin_bit_cast(int, x)));
    } else {
        const std::uint64_t v  = __builtin_bit_cast(std::uint64_t, x);
        const std::uint32_t lo = static_cast<std::uint32_t>(__builtin_amdgcn_ds_bpermute(src, static_cast<int>(v)));
        const std::uint32_t hi = static_cast<std::uint32_t>(__builtin_amdgcn_ds_bpermute(src, static_cast<int>(v >> 32)));
        return __builtin_bit_cast(T, (static_cast<std::uint64_t>(hi) << 32) | lo);
    }
}

/// Reduction over the 64 lanes as an xor butterfly: a balanced pairwise tree in lane order; every lane gets the result.
template <class T, class Op>
__device__ inline T reduce(T x, Op op) {
    for (int m = 1; m < kSize; m <<= 1) x = op(x, shfl_xor(x, m));
    return x;
}

/// Reduction over lanes 0..n-1 (n uniform, 1 <= n <= 64; lanes n..63 may be inactive or hold anything): the same balanced
/// pairwise tree in lane order as reduce(), built by strided steps that only read lanes below n, then broadcast from
/// lane 0. Every active lane gets the result.
template <class T, class Op>
__device__ inline T reduce_first(T x, Op op, int n) {
    if (n == kSize) return reduce(x, op);
    const int l = lane_id();
    for (int width = 1; width < n; width <<= 1) {
        const int src = l + width < n ? l + width : l;
        const T   y   = shfl(x, src);
        if ((l & (2 * width - 1)) == 0 && l + width < n) x = op(x, y);
    }
    return shfl(x, 0);
}

}  // namespace naga::cospan::wave

namespace naga::cospan::block {

/// Work-group reduction (up to 16 waves): waves reduce by butterfly, then the wave results are combined as a balanced
/// pairwise tree in wave order. Every work-item gets the result. Uses one LDS slot per wave; ends with a barrier so the
/// slots may be reused by the next call.
template <class T, class Op>
__device__ inline T reduce_first(T x, Op op, int count);

template <class T, class Op>
__device__ inline T reduce(T x, Op op) {
    return reduce_first(x, op, static_cast<int>(blockDim.x * blockDim.y * blockDim.z));
}

/// The same over the values of work-items 0..count-1 in local linear order (count uniform, 1 <= count <= the group
/// size); the other work-items take part and get the result too. A wave that is only partly populated (a group size that
/// is not a multiple of 64) never reads its missing lanes.
template <class T, class Op>
__device__ inline T reduce_first(T x, Op op, int count) {
    __shared__ T slots[16];
    const int    tid    = static_cast<int>(threadIdx.x + blockDim.x * (threadIdx.y + blockDim.y * threadIdx.z));
    const int    w      = tid / wave::kSize;
    const int    nwaves = (count + wave::kSize - 1) / wave::kSize;
    const int    in_w   = count - w * wave::kSize;   // values held by this wave (may be <= 0 past the count)
    if (in_w > 0) x = wave::reduce_first(x, op, in_w < wave::kSize ? in_w : wave::kSize);
    const int nthreads = static_cast<int>(blockDim.x * blockDim.y * blockDim.z);
    if (nthreads <= wave::kSize) return x;   // one wave: reduce_first already broadcast lane 0's result
    if ((tid & (wave::kSize - 1)) == 0 && w < nwaves) slots[w] = x;
    __syncthreads();
    // the tree over the wave results, unrolled over the 16-wave bound so that v stays in registers (a loop bounded by
    // nwaves would index v at run time, which puts it in scratch memory: per work-item stores and loads to HBM)
    T v[16];
#pragma unroll
    for (int k = 0; k < 16; ++k) v[k] = slots[k < nwaves ? k : 0];
#pragma unroll
    for (int width = 1; width < 16; width <<= 1)
#pragma unroll
        for (int k = 0; k + width < 16; k += 2 * width)
            if (k + width < nwaves) v[k] = op(v[k], v[k + width]);
    __syncthreads();
    return v[0];
}

/// Work-group exclusive scan in local linear order (integers or anything whose op is associative enough for the caller):
/// wave DPP scans, then the wave totals through LDS. Every work-item gets its exclusive prefix; `total` (if given) gets
/// the group's total.
template <class T, class Op>
__device__ inline T exclusive_scan(T x, Op op, T identity, T* total = nullptr) {
    __shared__ T slots[16];
    const int    nthreads = static_cast<int>(blockDim.x * blockDim.y * blockDim.z);
    const int    tid      = static_cast<int>(threadIdx.x + blockDim.x * (threadIdx.y + blockDim.y * threadIdx.z));
    const int    nwaves   = (nthreads + wave::kSize - 1) / wave::kSize;
    const int    w        = tid / wave::kSize;
    const T      incl     = wave::inclusive_scan(x, op, identity);
    // lane-exclusive prefix: the inclusive value of the previous lane (lane 0 of the wave: identity)
    T prev = identity;
    {
        const int src = ((wave::lane_id() + wave::kSize - 1) & (wave::kSize - 1)) << 2;
        if constexpr (sizeof(T) == 4) {
            prev = __builtin_bit_cast(T, __builtin_amdgcn_ds_bpermute(src, __builtin_bit_cast(int, incl)));
        } else {
            const std::uint64_t v  = __builtin_bit_cast(std::uint64_t, incl);
            const std::uint32_t lo = static_cast<std::uint32_t>(__builtin_amdgcn_ds_bpermute(src, static_cast<int>(v)));
            const std::uint32_t hi = static_cast<std::uint32_t>(__builtin_amdgcn_ds_bpermute(src, static_cast<int>(v >> 32)));
            prev                   = __builtin_bit_cast(T, (static_cast<std::uint64_t>(hi) << 32) | lo);
        }
        if (wave::lane_id() == 0) prev = identity;
    }
    if ((tid & (wave::kSize - 1)) == wave::kSize - 1 || tid == nthreads - 1) slots[w] = incl;
    __syncthreads();
    T before = identity, all = identity;
    for (int k = 0; k < nwaves; ++k) {
        if (k < w) before = op(before, slots[k]);
        all = op(all, slots[k]);
    }
    __syncthreads();
    if (total) *total = all;
    return op(before, prev);
}

}  // namespace naga::cospan::block

namespace naga::cospan::bounded {

/// Polls `ready()` at most `limit` times with a short sleep between polls: true as soon as it holds, false when it never
/// did. A limit of 0 polls nothing and gives up at once (the host's way to force the give-up path).
template <class Ready>
__device__ inline bool poll(Ready&& ready, std::uint32_t limit) {
    for (std::uint32_t k = 0; k < limit; ++k) {
        if (ready()) return true;
        __builtin_amdgcn_s_sleep(1);
    }
    return false;
}

/// ORs `bits` (CS_HIP_FAULT_*) into the device's fault word; a null word reports nowhere.
__device__ inline void report(std::uint32_t* fault_word, std::uint32_t bits) {
    if (fault_word != nullptr) __hip_atomic_fetch_or(fault_word, bits, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

}  // namespace naga::cospan::bounded
