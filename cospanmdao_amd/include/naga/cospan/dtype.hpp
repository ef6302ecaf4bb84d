#pragma once

// =============================================================================
// naga/cospan/dtype.hpp — the closed set of port element types
// =============================================================================
//
// The reference's set (cospanmdao/include/naga/cospan/dtype.hpp:17-20: f32 f64 i32 u32 i64 u64 u8 and the atomic forms
// of the 4- and 8-byte types) plus two STORAGE types for reduced-precision ports: bf16 and e4m3 (OCP FP8, 4 exponent
// bits, 3 mantissa bits, finite-only). A storage type is read and written as its bit pattern; arithmetic widens it to
// f32 (bf16: exact; e4m3: the gfx950 conversion instruction, exact). Neither has an atomic form.

#include <cstddef>
#include <cstdint>

#include "naga/cospan/vocab.hpp"

namespace naga::cospan {

enum class DType : std::uint8_t {
    f32, f64, i32, u32, i64, u64, u8, bf16, e4m3,
    atomic_f32, atomic_f64, atomic_i32, atomic_u32, atomic_i64, atomic_u64
};

constexpr bool is_atomic(DType dt) noexcept { return dt >= DType::atomic_f32; }
constexpr DType value_dtype(DType dt) noexcept {
    return is_atomic(dt) ? static_cast<DType>(static_cast<std::uint8_t>(dt) - static_cast<std::uint8_t>(DType::atomic_f32))
                         : dt;
}
constexpr bool has_atomic_form(DType dt) noexcept { return value_dtype(dt) <= DType::u64; }
constexpr DType to_atomic(DType dt) noexcept {
    return has_atomic_form(dt) ? static_cast<DType>(static_cast<std::uint8_t>(value_dtype(dt)) + static_cast<std::uint8_t>(DType::atomic_f32))
                               : dt;
}
constexpr bool is_storage_only(DType dt) noexcept { return dt == DType::bf16 || dt == DType::e4m3; }
constexpr bool is_integer_value_dtype(DType dt) noexcept {
    const DType v = value_dtype(dt);
    return v == DType::i32 || v == DType::u32 || v == DType::i64 || v == DType::u64 || v == DType::u8;
}

constexpr std::size_t dtype_size(DType dt) noexcept {
    switch (value_dtype(dt)) {
        case DType::f32: case DType::i32: case DType::u32: return 4;
        case DType::f64: case DType::i64: case DType::u64: return 8;
        case DType::bf16: return 2;
        default: return 1;  // u8, e4m3
    }
}

constexpr bool dtype_compatible(DType a, DType b) noexcept { return value_dtype(a) == value_dtype(b); }

constexpr const char* dtype_name(DType dt) noexcept {
    switch (dt) {
        case DType::f32: return "f32";
        case DType::f64: return "f64";
        case DType::i32: return "i32";
        case DType::u32: return "u32";
        case DType::i64: return "i64";
        case DType::u64: return "u64";
        case DType::u8: return "u8";
        case DType::bf16: return "bf16";
        case DType::e4m3: return "e4m3";
        case DType::atomic_f32: return "atomic_f32";
        case DType::atomic_f64: return "atomic_f64";
        case DType::atomic_i32: return "atomic_i32";
        case DType::atomic_u32: return "atomic_u32";
        case DType::atomic_i64: return "atomic_i64";
        case DType::atomic_u64: return "atomic_u64";
    }
    return "?";
}

/// bf16 storage: the upper half of an f32.
struct bf16 {
    std::uint16_t bits;
    CS_HD float   to_float() const { return __builtin_bit_cast(float, static_cast<std::uint32_t>(bits) << 16); }
    /// round to nearest even; a NaN stays a quiet NaN
    static CS_HD bf16 from_float(float f) {
        const std::uint32_t u = __builtin_bit_cast(std::uint32_t, f);
        if ((u & 0x7fffffffu) > 0x7f800000u) return bf16{static_cast<std::uint16_t>((u >> 16) | 0x0040u)};
        return bf16{static_cast<std::uint16_t>((u + 0x7fffu + ((u >> 16) & 1u)) >> 16)};
    }
};

/// OCP FP8 E4M3 storage (bias 7, no infinities, |x| <= 448).
struct e4m3 {
    std::uint8_t bits;
};

template <DType DT>
inline constexpr DType value_dtype_of = value_dtype(DT);
template <DType DT>
inline constexpr bool is_atomic_v = is_atomic(DT);

namespace detail {
template <DType V> struct cpp_type;
template <> struct cpp_type<DType::f32> { using type = float; };
template <> struct cpp_type<DType::f64> { using type = double; };
template <> struct cpp_type<DType::i32> { using type = std::int32_t; };
template <> struct cpp_type<DType::u32> { using type = std::uint32_t; };
template <> struct cpp_type<DType::i64> { using type = std::int64_t; };
template <> struct cpp_type<DType::u64> { using type = std::uint64_t; };
template <> struct cpp_type<DType::u8>  { using type = std::uint8_t; };
template <> struct cpp_type<DType::bf16> { using type = bf16; };
template <> struct cpp_type<DType::e4m3> { using type = e4m3; };
}  // namespace detail

template <DType DT>
using cpp_type_t = typename detail::cpp_type<value_dtype_of<DT>>::type;

static_assert(value_dtype_of<DType::atomic_f32> == DType::f32);
static_assert(value_dtype_of<DType::atomic_u64> == DType::u64);
static_assert(to_atomic(DType::i64) == DType::atomic_i64);
static_assert(!has_atomic_form(DType::bf16) && !has_atomic_form(DType::u8));

}  // namespace naga::cospan
