// Storage formats of the distribution fields.
//
// EXACT64 keeps the reference's unshifted fp64 pdfs and its operation order
// (the parity mode). The other modes keep the shifted value g_q = f_q - w_q,
// multiplied by a per-direction power of two for the 16-bit formats, and do
// their arithmetic in the Calc type (float, or double for F64S).
// Conversions round to nearest even (v_cvt_f16_f32, v_cvt_pk_bf16_f32).
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
    using T = __half;
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

// Per-direction scale of the 16-bit formats: 2^8 times the power of two
// nearest to 1/w_q. |g_q / w_q| stays below ~1 for stable flows, so stored
// magnitudes sit near 2^8, far from the fp16 subnormal range (< 2^-14) and
// with 2^8 of headroom below the fp16 maximum (65504). Powers of two keep the
// scaling itself exact.
template <class L>
LBM_HD constexpr int scale_exp(int q) {
    // smallest e with 2^e >= (wden / wnum) / sqrt(2), compared in integers
    int e = 0;
    const int num = L::wnum(q), den = L::wden;
    while (2 * (num << e) * (num << e) < den * den) ++e;
    return e + 8;
}

template <class L>
LBM_HD constexpr float scale_of(int q) {
    return static_cast<float>(1u << scale_exp<L>(q));
}

template <class L>
LBM_HD constexpr float inv_scale_of(int q) {
    return 1.0f / static_cast<float>(1u << scale_exp<L>(q));
}

// Load / store of one stored value as Calc, with the scale applied.
template <Mode M, class L>
struct Io {
    using S = StorageOf<M>;
    using T = typename S::T;
    using Calc = typename S::Calc;
    static constexpr bool scaled = (M == Mode::F16 || M == Mode::BF16);

    template <int q>
    __device__ __forceinline__ static Calc to_calc(T v) {
        if constexpr (M == Mode::F16) {
            return __half2float(v) * inv_scale_of<L>(q);
        } else if constexpr (M == Mode::BF16) {
            return static_cast<float>(v) * inv_scale_of<L>(q);
        } else {
            return static_cast<Calc>(v);
        }
    }
    template <int q>
    __device__ __forceinline__ static T from_calc(Calc c) {
        if constexpr (M == Mode::F16) {
            return __float2half_rn(c * scale_of<L>(q));
        } else if constexpr (M == Mode::BF16) {
            return static_cast<__bf16>(c * scale_of<L>(q));
        } else {
            return static_cast<T>(c);
        }
    }
};

}  // namespace lbm_amd
