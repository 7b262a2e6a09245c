// Lattice velocity sets D3Q15, D3Q19, D3Q27 for host and device code.
//
// The direction order, the integer velocities and the rational weights are the
// ones of the reference lattice models (include/model.h:13-50 for D3Q15,
// :53-91 for D3Q19, :96-139 for D3Q27). They have to be identical, not merely
// equivalent: the reference stores pdfs by direction index, sums moments in
// ascending direction order and bounces back with inv(q) = Q-1-q, so any other
// ordering changes both the memory format and the rounding of every sum.
//
// Weights are stored as numerator / denominator and the double weight is the
// correctly rounded quotient, which is how `k.0/n` literals evaluate.
#pragma once

#include <cstdint>
#include <type_traits>
#include <utility>

#if defined(__HIPCC__)
#define LBM_HD __host__ __device__
#else
#define LBM_HD
#endif

namespace lbm_amd {

// Speed of sound as the reference defines it: the long double literal
// 0.57735026919L (include/lbmdefinitions.h:47) rounded to double. The hex form
// avoids relying on the device compiler's long double; a host test checks that
// both spellings are the same double.
constexpr double kCS = 0x1.279a745904047p-1;

// ---------------------------------------------------------------- D3Q15 ----
struct D3Q15 {
    static constexpr int Q = 15;
    static constexpr int rest = 7;
    static constexpr int wden = 72;
    LBM_HD static constexpr int cx(int q) {
        constexpr int t[15] = {-1, 1, 0, -1, 1, 0, -1, 0, 1, 0, -1, 1, 0, -1, 1};
        return t[q];
    }
    LBM_HD static constexpr int cy(int q) {
        constexpr int t[15] = {-1, -1, 0, 1, 1, -1, 0, 0, 0, 1, -1, -1, 0, 1, 1};
        return t[q];
    }
    LBM_HD static constexpr int cz(int q) {
        constexpr int t[15] = {-1, -1, -1, -1, -1, 0, 0, 0, 0, 0, 1, 1, 1, 1, 1};
        return t[q];
    }
    LBM_HD static constexpr int wnum(int q) {
        constexpr int t[15] = {1, 1, 8, 1, 1, 8, 8, 16, 8, 8, 1, 1, 8, 1, 1};
        return t[q];
    }
    static constexpr const char* name = "D3Q15";
};

// ---------------------------------------------------------------- D3Q19 ----
struct D3Q19 {
    static constexpr int Q = 19;
    static constexpr int rest = 9;
    static constexpr int wden = 36;
    LBM_HD static constexpr int cx(int q) {
        constexpr int t[19] = {0, -1, 0, 1, 0, -1, 0, 1, -1, 0, 1, -1, 0, 1, 0, -1, 0, 1, 0};
        return t[q];
    }
    LBM_HD static constexpr int cy(int q) {
        constexpr int t[19] = {-1, 0, 0, 0, 1, -1, -1, -1, 0, 0, 0, 1, 1, 1, -1, 0, 0, 0, 1};
        return t[q];
    }
    LBM_HD static constexpr int cz(int q) {
        constexpr int t[19] = {-1, -1, -1, -1, -1, 0, 0, 0, 0, 0, 0, 0, 0, 0, 1, 1, 1, 1, 1};
        return t[q];
    }
    LBM_HD static constexpr int wnum(int q) {
        constexpr int t[19] = {1, 1, 2, 1, 1, 1, 2, 1, 2, 12, 2, 1, 2, 1, 1, 1, 2, 1, 1};
        return t[q];
    }
    static constexpr const char* name = "D3Q19";
};

// ---------------------------------------------------------------- D3Q27 ----
// Direction q = 9*(cz+1) + 3*(cy+1) + (cx+1): x fastest, z slowest.
struct D3Q27 {
    static constexpr int Q = 27;
    static constexpr int rest = 13;
    static constexpr int wden = 216;
    LBM_HD static constexpr int cx(int q) { return q % 3 - 1; }
    LBM_HD static constexpr int cy(int q) { return (q / 3) % 3 - 1; }
    LBM_HD static constexpr int cz(int q) { return q / 9 - 1; }
    LBM_HD static constexpr int wnum(int q) {
        // 64 for the rest direction, /4 per non-zero component.
        return 64 >> (2 * ((cx(q) != 0) + (cy(q) != 0) + (cz(q) != 0)));
    }
    static constexpr const char* name = "D3Q27";
};

template <class L>
LBM_HD constexpr double weight(int q) {
    return static_cast<double>(L::wnum(q)) / static_cast<double>(L::wden);
}

template <class L>
LBM_HD constexpr int inv(int q) { return L::Q - 1 - q; }

// Compile-time loop: f(std::integral_constant<int, I>{}) for I in [B, E).
template <int B, int E, class F>
LBM_HD inline void static_for(F&& f) {
    if constexpr (B < E) {
        f(std::integral_constant<int, B>{});
        static_for<B + 1, E>(static_cast<F&&>(f));
    }
}

}  // namespace lbm_amd
