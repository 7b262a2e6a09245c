// Collision operators and boundary rules, per cell, in registers.
#pragma once

#include <type_traits>

#include "domain_view.hpp"
#include "lbm_hip/lattice.hpp"
#include "storage.hpp"

namespace lbm_amd {

// ---------------------------------------------------------------------------
// Reference BGK in fp64 (include/collision.hpp:7-70), operation for operation:
//   rho = sum_q f_q (from 0, ascending q)
//   m_d = sum_q f_q * c_qd (from 0, ascending q), u_d = m_d / rho
//   feq_q = w_q * rho * (1 + cu/(C_S*C_S) + cu*cu/(2*C_S*C_S*C_S*C_S) - uu/(2*C_S*C_S))
//   f_q -= (f_q - feq_q) / tau
// Terms with c_qd = 0 add +-0 to a sum that starts at +0 and are left out:
// x + (+-0) == x for every x this sum can hold, so the result bits are the
// same. The translation unit that instantiates this is built without FMA
// contraction.
template <class L>
__device__ __forceinline__ void moments_exact(const double (&f)[L::Q], double& rho, double (&u)[3]) {
    rho = 0.0;
    static_for<0, L::Q>([&](auto qc) { rho += f[decltype(qc)::value]; });
    double m0 = 0.0, m1 = 0.0, m2 = 0.0;
    static_for<0, L::Q>([&](auto qc) {
        constexpr int q = decltype(qc)::value;
        if constexpr (L::cx(q) > 0) m0 += f[q];
        if constexpr (L::cx(q) < 0) m0 += -f[q];
    });
    static_for<0, L::Q>([&](auto qc) {
        constexpr int q = decltype(qc)::value;
        if constexpr (L::cy(q) > 0) m1 += f[q];
        if constexpr (L::cy(q) < 0) m1 += -f[q];
    });
    static_for<0, L::Q>([&](auto qc) {
        constexpr int q = decltype(qc)::value;
        if constexpr (L::cz(q) > 0) m2 += f[q];
        if constexpr (L::cz(q) < 0) m2 += -f[q];
    });
    u[0] = m0 / rho;
    u[1] = m1 / rho;
    u[2] = m2 / rho;
}

// Velocity with a caller-given density (Outflow/Pressure pass rho_ref,
// include/boundary.hpp:144,209).
template <class L>
__device__ __forceinline__ void velocity_exact(const double (&f)[L::Q], double dens, double (&u)[3]) {
    double m0 = 0.0, m1 = 0.0, m2 = 0.0;
    static_for<0, L::Q>([&](auto qc) {
        constexpr int q = decltype(qc)::value;
        if constexpr (L::cx(q) > 0) m0 += f[q];
        if constexpr (L::cx(q) < 0) m0 += -f[q];
    });
    static_for<0, L::Q>([&](auto qc) {
        constexpr int q = decltype(qc)::value;
        if constexpr (L::cy(q) > 0) m1 += f[q];
        if constexpr (L::cy(q) < 0) m1 += -f[q];
    });
    static_for<0, L::Q>([&](auto qc) {
        constexpr int q = decltype(qc)::value;
        if constexpr (L::cz(q) > 0) m2 += f[q];
        if constexpr (L::cz(q) < 0) m2 += -f[q];
    });
    u[0] = m0 / dens;
    u[1] = m1 / dens;
    u[2] = m2 / dens;
}

// One equilibrium value, reference expression tree.
template <class L, int q>
__device__ __forceinline__ double feq_exact(double rho, const double (&u)[3]) {
    double cu = 0.0;
    if constexpr (L::cx(q) > 0) cu += u[0];
    if constexpr (L::cx(q) < 0) cu += -u[0];
    if constexpr (L::cy(q) > 0) cu += u[1];
    if constexpr (L::cy(q) < 0) cu += -u[1];
    if constexpr (L::cz(q) > 0) cu += u[2];
    if constexpr (L::cz(q) < 0) cu += -u[2];
    double uu = 0.0;
    uu += u[0] * u[0];
    uu += u[1] * u[1];
    uu += u[2] * u[2];
    constexpr double cs = kCS;
    return weight<L>(q) * rho *
           (1 + cu / (cs * cs) + (cu) * (cu) / (2 * cs * cs * cs * cs) - uu / (2 * cs * cs));
}

template <class L>
__device__ __forceinline__ void collide_exact(double (&f)[L::Q], double tau) {
    double rho, u[3];
    moments_exact<L>(f, rho, u);
    static_for<0, L::Q>([&](auto qc) {
        constexpr int q = decltype(qc)::value;
        const double feq = feq_exact<L, q>(rho, u);
        f[q] -= (f[q] - feq) / tau;
    });
}

// ---------------------------------------------------------------------------
// Coefficient 3 w_q of the odd (momentum-carrying) part of the equilibrium.
//
// In fp32 the constant for each velocity class (|c|^2 = 1, 2, 3) is rounded
// to the side that keeps the first-moment sum rule sum_q c_qx^2 (3 w_q) = 1
// exact in real arithmetic. Rounded to nearest, the D3Q15/D3Q19 constants
// (mantissa 4/3) all sit 2^-25 above their value and the sum is 1 + 3e-8, so
// every collision would scale the momentum by 1 + omega * 3e-8: a velocity
// drift of 2.4e-5 relative after 500 steps of a Taylor-Green box. The
// rounding errors of the two directions are +2^-25 and -2^-24 (mantissa 4/3)
// or +2^-27 and -2^-24 (mantissa 16/9), and the class shares of the sum are
// D3Q15 2/3, 1/3; D3Q19 1/3, 2/3; D3Q27 4/9, 4/9, 1/9, which fixes the choice
// below. fp64 keeps the nearest constant (its drift is 1e-16 per step).
template <class L>
struct OddCoef32;
template <>
struct OddCoef32<D3Q15> {  // axis 1/3 up, diagonal 1/24 down
    static constexpr float c[4] = {0.0f, 0x1.555556p-2f, 0.0f, 0x1.555554p-5f};
};
template <>
struct OddCoef32<D3Q19> {  // axis 1/6 down, diagonal 1/12 up
    static constexpr float c[4] = {0.0f, 0x1.555554p-3f, 0x1.555556p-4f, 0.0f};
};
template <>
struct OddCoef32<D3Q27> {  // axis 2/9 up, edge 1/18 up, corner 1/72 down
    static constexpr float c[4] = {0.0f, 0x1.c71c72p-3f, 0x1.c71c72p-5f, 0x1.c71c70p-7f};
};

template <class L>
LBM_HD constexpr int speed_class(int q) {
    return L::cx(q) * L::cx(q) + L::cy(q) * L::cy(q) + L::cz(q) * L::cz(q);
}

template <class L, class Calc>
LBM_HD constexpr Calc odd_coef(int q) {
    if constexpr (std::is_same<Calc, float>::value)
        return OddCoef32<L>::c[speed_class<L>(q)];
    else
        return static_cast<Calc>(3.0 * weight<L>(q));
}

// The fp32 constants are within one ulp of 3 w_q and satisfy the sum rule
// exactly (the double sum of these few floats is exact).
template <class L>
constexpr bool odd_coef32_ok() {
    double sum = 0;
    for (int q = 0; q < L::Q; ++q) {
        if (q == L::rest) continue;
        const double exact = 3.0 * weight<L>(q);
        const double v = static_cast<double>(odd_coef<L, float>(q));
        const double d = v > exact ? v - exact : exact - v;
        if (!(d < exact * 0x1p-23)) return false;
        sum += static_cast<double>(L::cx(q) * L::cx(q)) * v;
    }
    return sum == 1.0;
}
static_assert(odd_coef32_ok<D3Q15>() && odd_coef32_ok<D3Q19>() && odd_coef32_ok<D3Q27>(),
              "fp32 odd-part coefficients break the momentum sum rule");

// ---------------------------------------------------------------------------
// a * b + c rounded once (v_fma_f32 / v_fma_f64). Written out where it is
// wanted: the kernels are built without implicit contraction, so every
// kernel that runs a cell rounds its arithmetic the same way. For dual
// numbers the value is the fp64 fma and the tangent follows the product rule.
template <class Calc>
__device__ __forceinline__ Calc fmadd(Calc a, Calc b, Calc c) {
    if constexpr (std::is_same<Calc, float>::value)
        return __builtin_fmaf(a, b, c);
    else if constexpr (std::is_same<Calc, double>::value)
        return __builtin_fma(a, b, c);
    else
        return Calc(__builtin_fma(a.v, b.v, c.v), a.d * b.v + a.v * b.d + c.d);
}

// Index of the first direction with a nonzero component along axis d (0, 1,
// 2): sums over directions start from that term instead of from zero.
template <class L>
constexpr int first_dir(int d) {
    for (int q = 0; q < L::Q; ++q)
        if ((d == 0 ? L::cx(q) : (d == 1 ? L::cy(q) : L::cz(q))) != 0) return q;
    return -1;
}

// Accumulates s +- v for direction q along axis d, starting at the first
// contributing direction.
template <class L, int d, int q, class Calc>
__device__ __forceinline__ void acc_axis(Calc& s, Calc v) {
    constexpr int c = d == 0 ? L::cx(q) : (d == 1 ? L::cy(q) : L::cz(q));
    if constexpr (q == first_dir<L>(d)) {
        s = c > 0 ? v : -v;
    } else {
        if constexpr (c > 0) s += v;
        if constexpr (c < 0) s -= v;
    }
}

// Shifted BGK: g_q = f_q - w_q, arithmetic in Calc (float or double),
// 1/c_s^2 = 3 exactly. With a Brinkman field the equilibrium velocity is
// (1 - alpha) * (u + tau * g) (porosity scaling of the equilibrium velocity
// plus a velocity-shift body force).
//   drho = sum g, rho = 1 + drho, u = sum c g / rho
//   geq_q = w_q * (drho + rho * (3 cu + 4.5 cu^2 - 1.5 u.u))
//   g_q  += omega * (geq_q - g_q)
// evaluated per pair (q, Q-1-q), which shares the even part, as
//   E = w_q * (base + (4.5 rho) cu cu),  O = (w3_q rho) cu
//   g_q = (1 - omega) g_q + omega (E + O),  g_qbar = (1 - omega) g_qbar + omega (E - O)
// with base = drho - 1.5 rho u.u, w3_q = 3 w_q, and fused multiply-adds where
// a product meets a sum. The relaxation rate multiplies the equilibrium as a
// whole: folding it into E and O separately measured a systematic velocity
// drift (2.8e-6 relative after 500 fp32 steps of a Taylor-Green box,
// against 1.6e-7).
// u is a true quotient: a reciprocal of rho multiplied into j rounds twice,
// and over a flow's smooth density field those roundings bias the momentum
// (5e-6 relative drift after 500 fp32 steps of a Taylor-Green box, against
// 6e-7 with the division). RCP selects the reciprocal form (the hardware
// reciprocal for fp32), used with 16-bit storage, whose own rounding (~1e-3)
// dwarfs that bias.
// SX: g[] holds the values scaled by 2^SX (the stored scale of the format),
// in and out. The sums are taken on the scaled values and unscaled once,
// and the equilibrium constants carry the scale; powers of two commute with
// rounding away from the subnormal range, so this is the unscaled
// computation scaled by 2^SX.
// Returns base = drho - 1.5 rho u.u of the incoming values: it is NaN or
// infinite whenever one of them is, which the batched step uses to flag
// domains that have diverged (callers that do not need it ignore it).
template <class L, class Calc, bool BRINK, bool RCP = false, int SX = 0>
__device__ __forceinline__ Calc collide_shifted(Calc (&g)[L::Q], Calc omega, Calc alpha,
                                                Calc tgx, Calc tgy, Calc tgz) {
    static_assert(SX == 0 || RCP, "scaled values take the reciprocal form");
    constexpr Calc kS = static_cast<Calc>(static_cast<double>(1u << SX));
    constexpr Calc kInvS = static_cast<Calc>(1.0 / static_cast<double>(1u << SX));
    Calc drho = g[0], jx{}, jy{}, jz{};
    static_for<0, L::Q>([&](auto qc) {
        constexpr int q = decltype(qc)::value;
        if constexpr (q > 0) drho += g[q];
        acc_axis<L, 0, q>(jx, g[q]);
        acc_axis<L, 1, q>(jy, g[q]);
        acc_axis<L, 2, q>(jz, g[q]);
    });
    if constexpr (SX != 0) drho = drho * kInvS;
    const Calc rho = Calc(1) + drho;
    Calc ux, uy, uz;
    if constexpr (RCP) {
        Calc inv_rho;
        if constexpr (std::is_same<Calc, float>::value)
            inv_rho = __builtin_amdgcn_rcpf(rho);
        else
            inv_rho = Calc(1) / rho;
        if constexpr (SX != 0) inv_rho = inv_rho * kInvS;  // also unscales j
        ux = jx * inv_rho;
        uy = jy * inv_rho;
        uz = jz * inv_rho;
    } else {
        ux = jx / rho;
        uy = jy / rho;
        uz = jz / rho;
    }
    if constexpr (BRINK) {
        const Calc s = Calc(1) - alpha;
        ux = s * (ux + tgx);
        uy = s * (uy + tgy);
        uz = s * (uz + tgz);
    }
    const Calc uu = fmadd(ux, ux, fmadd(uy, uy, uz * uz));
    const Calc base = fmadd(Calc(-1.5) * rho, uu, drho);
    const Calc one_m_omega = Calc(1) - omega;
    const Calc t = Calc(4.5) * rho;
    // Pairs (q, Q-1-q) share the even part of the equilibrium.
    static_for<0, L::Q / 2>([&](auto qc) {
        constexpr int q = decltype(qc)::value;
        constexpr int qb = L::Q - 1 - q;
        constexpr Calc w = static_cast<Calc>(weight<L>(q)) * kS;
        constexpr Calc w3 = odd_coef<L, Calc>(q) * kS;
        Calc cu{};
        if constexpr (L::cx(q) != 0) cu = L::cx(q) > 0 ? ux : -ux;
        if constexpr (L::cy(q) != 0) {
            if constexpr (L::cx(q) == 0)
                cu = L::cy(q) > 0 ? uy : -uy;
            else if constexpr (L::cy(q) > 0)
                cu += uy;
            else
                cu -= uy;
        }
        if constexpr (L::cz(q) != 0) {
            if constexpr (L::cx(q) == 0 && L::cy(q) == 0)
                cu = L::cz(q) > 0 ? uz : -uz;
            else if constexpr (L::cz(q) > 0)
                cu += uz;
            else
                cu -= uz;
        }
        const Calc e = w * fmadd(t * cu, cu, base);
        const Calc o = (w3 * rho) * cu;
        g[q] = fmadd(one_m_omega, g[q], omega * (e + o));
        g[qb] = fmadd(one_m_omega, g[qb], omega * (e - o));
    });
    constexpr int r = L::rest;
    constexpr Calc wr = static_cast<Calc>(weight<L>(r)) * kS;
    g[r] = fmadd(one_m_omega, g[r], omega * (wr * base));
    return base;
}

// False when v is NaN or infinite (the value part of a dual number).
template <class Calc>
__device__ __forceinline__ bool finite_calc(Calc v) {
    if constexpr (std::is_same<Calc, float>::value || std::is_same<Calc, double>::value)
        return __builtin_isfinite(v);
    else
        return __builtin_isfinite(v.v);
}

}  // namespace lbm_amd
