// Forward-mode dual numbers (value, directional derivative) for the fp64
// tangent run: the collision and streaming templates instantiated with
// Dual<double> propagate d/ds of every quantity along a parameter direction.
#pragma once

#include "lbm_hip/lattice.hpp"

namespace lbm_amd {

template <class T>
struct Dual {
    T v, d;
    LBM_HD constexpr Dual() : v(0), d(0) {}
    LBM_HD constexpr Dual(T a) : v(a), d(0) {}  // constants have no derivative
    LBM_HD constexpr Dual(T a, T b) : v(a), d(b) {}

    LBM_HD friend constexpr Dual operator+(Dual a, Dual b) { return {a.v + b.v, a.d + b.d}; }
    LBM_HD friend constexpr Dual operator-(Dual a, Dual b) { return {a.v - b.v, a.d - b.d}; }
    LBM_HD friend constexpr Dual operator-(Dual a) { return {-a.v, -a.d}; }
    LBM_HD friend constexpr Dual operator*(Dual a, Dual b) { return {a.v * b.v, a.d * b.v + a.v * b.d}; }
    LBM_HD friend constexpr Dual operator/(Dual a, Dual b) {
        const T q = a.v / b.v;
        return {q, (a.d - q * b.d) / b.v};
    }
    LBM_HD Dual& operator+=(Dual b) { return *this = *this + b; }
    LBM_HD Dual& operator-=(Dual b) { return *this = *this - b; }
    LBM_HD Dual& operator*=(Dual b) { return *this = *this * b; }
};

}  // namespace lbm_amd
