// Storage formats of the distribution fields.
//
// EXACT64 keeps the reference's unshifted fp64 pdfs and its operation order
// (the parity mode). The other modes keep the shifted value g_q = f_q - w_q,
// multiplied by a per-direction power of two for the 16-bit formats, and do
// their arithmetic in the Calc type (float, or double for F64S).
// Conversions round to nearest even (v_cvt_f16_f32, v_cvt_pk_bf16_f32). The
// 16-bit formats are the compiler's arithmetic types _Float16 and __bf16.
#pragma once

#include <hip/hip_runtime.h>
#include <hip/hip_fp16.h>

#include <cstdint>

#include "dual.hpp"
#include "lbm_hip/lattice.hpp"

namespace lbm_amd {

// DUAL64 is internal: shifted values as (value, tangent) fp64 pairs for the
// forward-mode derivative run.
enum class Mode : int { EXACT64 = 0, F32 = 1, F16 = 2, BF16 = 3, F64S = 4, DUAL64 = 5 };

template <Mode M>
struct StorageOf;
template <>
struct StorageOf<Mode::EXACT64> {
    using T = double;
    using Calc = double;
    static constexpr bool shifted = false;
};
template <>
struct StorageOf<Mode::F64S> {
    using T = double;
    using Calc = double;
    static constexpr bool shifted = true;
};
template <>
struct StorageOf<Mode::DUAL64> {
    using T = Dual<double>;
    using Calc = Dual<double>;
    static constexpr bool shifted = true;
};
template <>
struct StorageOf<Mode::F32> {
    using T = float;
    using Calc = float;
    static constexpr bool shifted = true;
};
template <>
struct StorageOf<Mode::F16> {
    using T = _Float16;
    using Calc = float;
    static constexpr bool shifted = true;
};
template <>
struct StorageOf<Mode::BF16> {
    using T = __bf16;
    using Calc = float;
    static constexpr bool shifted = true;
};

// Velocity from momentum by reciprocal (16-bit storage) or by division.
template <Mode M>
constexpr bool kRcp = sizeof(typename StorageOf<M>::T) == 2;

// Scale of the stored values: fp16 stores g * 2^12, every other format g
// itself. fp16's normal range is [2^-14, 65504], so stored magnitudes of
// 2^-14 .. 2^16 cover |g| from 1.5e-8 to 16: stable flows keep |g_q| below
// w_q <= 1/3 (|g_q / w_q| < 1), far from the top, and values below the
// bottom are under the format's own rounding of the usual |g| ~ 1e-4 .. 1e-2.
// bf16 has fp32's exponent range and needs no scale. A power of two keeps the
// scaling exact, and one scale for all directions lets the collision work on
// the stored (scaled) values directly: scaled sums equal the scaled sums of
// the unscaled values, bit for bit.
template <Mode M>
constexpr int kStoreScaleExp = (M == Mode::F16) ? 12 : 0;

// Conversions of one stored value.
//   to_calc / from_calc: the shifted value g itself (boundary rules,
//     moments, field I/O);
//   load / store: the collision's representation, g * 2^SX (the stored
//     scale), which the step kernels pass through collide_shifted<..., SX>;
//   scaled: g -> g * 2^SX, for rule values entering a collision.
template <Mode M, class L>
struct Io {
    using S = StorageOf<M>;
    using T = typename S::T;
    using Calc = typename S::Calc;
    static constexpr int SX = kStoreScaleExp<M>;
    static constexpr float kScale = static_cast<float>(1u << SX);
    static constexpr float kInvScale = 1.0f / static_cast<float>(1u << SX);

    template <int q>
    __device__ __forceinline__ static Calc load(T v) {
        if constexpr (M == Mode::F16) {
            return static_cast<float>(v);
        } else if constexpr (M == Mode::BF16) {
            return static_cast<float>(v);
        } else {
            return static_cast<Calc>(v);
        }
    }
    template <int q>
    __device__ __forceinline__ static T store(Calc c) {
        if constexpr (M == Mode::F16) {
            // c is rounded to fp32 first: without this (empty) barrier the
            // compiler may fuse an fma producing c with the conversion into
            // one mixed-precision fma that rounds straight to fp16, in some
            // kernels and not in others, and their results would differ
            __asm__("" : "+v"(c));
            return static_cast<_Float16>(c);  // round to nearest even
        } else if constexpr (M == Mode::BF16) {
            return static_cast<__bf16>(c);
        } else {
            return static_cast<T>(c);
        }
    }
    template <int q>
    __device__ __forceinline__ static Calc scaled(Calc g) {
        if constexpr (SX != 0)
            return g * kScale;
        else
            return g;
    }
    template <int q>
    __device__ __forceinline__ static Calc to_calc(T v) {
        if constexpr (SX != 0)
            return load<q>(v) * kInvScale;
        else
            return load<q>(v);
    }
    template <int q>
    __device__ __forceinline__ static T from_calc(Calc c) {
        return store<q>(scaled<q>(c));
    }
    // A boundary rule's value g as the pull reads it: written into the
    // boundary cell in the storage format, then loaded (the reference's
    // boundary pass stores its values in the field before the stream pulls
    // them). An identity for fp32 and fp64 storage; for 16-bit storage it
    // rounds, in every kernel alike.
    template <int q>
    __device__ __forceinline__ static Calc through_storage(Calc g) {
        return load<q>(from_calc<q>(g));
    }
};

}  // namespace lbm_amd
