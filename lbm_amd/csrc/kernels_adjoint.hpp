// Discrete adjoint of the shifted-storage time step, and the objectives.
//
// Forward step t (pull form, as k_bulk / k_near):
//   h^t(x)_q = g^{t-1}_q(x - c_q)            (or a boundary rule of x's own g^{t-1})
//   g^t(x)   = C(h^t(x); alpha(x))           (BGK, optional Brinkman + body force)
// With lambda^t = dJ/dg^t and mu^t = dJ/dh^t = (dC/dh)^T lambda^t, streaming
// transposes to a pull in the opposite direction:
//   lambda^t_k(x) = mu^{t+1}_k(x + c_k)      if x + c_k is fluid,
//                 = mu^{t+1}_kbar(x)         if x + c_k is a no-slip or moving wall
//                                            (the adjoint of bounce-back),
//                 + sum_{moving-wall links q of x} mu^{t+1}_q(x) 6 w_q c_q.u_w
//                                            (the wall rule's density coupling),
//                 = 0                        if x + c_k is an inflow cell,
//                 = -mu^{t+1}_kbar(x) + c_k.W  if x + c_k is an outflow or
//                                            pressure cell (below).
// An outflow or pressure link j of x (x - c_j such a cell, reference density
// r) pulls h_j = 2 w_j ((r - 1) + r (4.5 (c_j.u)^2 - 1.5 u.u)) - g_jbar(x)
// with u = m / r, m the momentum of x's own g^t (include/boundary.hpp:
// 140-147, 206-211). Its derivative by g_k(x) is
// -delta(k, jbar) + 2 w_j (9 (c_j.u)(c_j.c_k) - 3 u.c_k), so every g_k(x)
// also receives c_k.W with W = sum_j mu_j(x) 2 w_j (9 (c_j.u_j) c_j - 3 u_j).
// The kernels therefore keep mu (not lambda) in memory: each backward step
// reads g^{t-1} (pull), mu^{t+1} (reverse pull) and writes mu^t, and adds
// dJ/dalpha(x) and dJ/du_wall of step t into fp64 accumulators.
//
// Transposed BGK Jacobian (s = 1 - alpha, v = u + tau g, ue = s v):
//   geq_q = w_q (drho + rho A_q(ue)),  A_q(v) = 3 c.v + 4.5 (c.v)^2 - 1.5 v.v
//   mu_k = (1 - omega) lambda_k + omega (P + s V.(c_k - u))
//   P = sum_q lambda_q w_q (1 + A_q),  V = sum_q lambda_q w_q (3 c_q + 9 (c_q.ue) c_q - 3 ue)
//   dJ/dalpha += -omega rho V.v
#pragma once

#include "collide.hpp"
#include "domain_view.hpp"
#include "kernels.hpp"
#include "kernels_fields.hpp"
#include "lbm_hip/lattice.hpp"
#include "ops_adjoint.hpp"
#include "storage.hpp"

namespace lbm_amd {

// dJ/dg of the objective at one cell, from the cell's final values g.
template <class L, class Calc>
__device__ __forceinline__ void objective_grad(const AdjArgs& a, int x, int y, int z,
                                               const Calc (&g)[L::Q], Calc (&lam)[L::Q]) {
    Calc drho = Calc(0), jx = Calc(0), jy = Calc(0), jz = Calc(0);
    static_for<0, L::Q>([&](auto qc) {
        constexpr int q = decltype(qc)::value;
        drho += g[q];
        jx += Calc(L::cx(q)) * g[q];
        jy += Calc(L::cy(q)) * g[q];
        jz += Calc(L::cz(q)) * g[q];
    });
    const Calc rho = Calc(1) + drho;
    const Calc ux = jx / rho, uy = jy / rho, uz = jz / rho;
    Calc ex = Calc(0), ey = Calc(0), ez = Calc(0);
    if (a.objective == OBJ_PLANE_MISFIT) {
        if (x == a.plane_x) {
            const double* t = a.target + 3 * ((static_cast<int64_t>(z) - 1) * a.L.yl + (y - 1));
            ex = Calc(2) * (ux - static_cast<Calc>(t[0]));
            ey = Calc(2) * (uy - static_cast<Calc>(t[1]));
            ez = Calc(2) * (uz - static_cast<Calc>(t[2]));
        }
    } else {
        ex = ux;
        ey = uy;
        ez = uz;
    }
    // dJ/dg_k = e.(c_k - u) / rho
    const Calc eu = ex * ux + ey * uy + ez * uz;
    static_for<0, L::Q>([&](auto qc) {
        constexpr int q = decltype(qc)::value;
        lam[q] = (Calc(L::cx(q)) * ex + Calc(L::cy(q)) * ey + Calc(L::cz(q)) * ez - eu) / rho;
    });
}

// mu = (dC/dh)^T lam at the pre-collision state h; returns dJ/dalpha.
template <class L, class Calc, bool BRINK>
__device__ __forceinline__ Calc adjoint_collide(const Calc (&h)[L::Q], const Calc (&lam)[L::Q],
                                                Calc (&mu)[L::Q], Calc omega, Calc alpha, Calc tgx,
                                                Calc tgy, Calc tgz) {
    Calc drho = Calc(0), jx = Calc(0), jy = Calc(0), jz = Calc(0);
    static_for<0, L::Q>([&](auto qc) {
        constexpr int q = decltype(qc)::value;
        drho += h[q];
        if constexpr (L::cx(q) > 0) jx += h[q];
        if constexpr (L::cx(q) < 0) jx -= h[q];
        if constexpr (L::cy(q) > 0) jy += h[q];
        if constexpr (L::cy(q) < 0) jy -= h[q];
        if constexpr (L::cz(q) > 0) jz += h[q];
        if constexpr (L::cz(q) < 0) jz -= h[q];
    });
    const Calc rho = Calc(1) + drho;
    const Calc ux = jx / rho, uy = jy / rho, uz = jz / rho;
    Calc s = Calc(1), vx = ux, vy = uy, vz = uz;
    if constexpr (BRINK) {
        s = Calc(1) - alpha;
        vx = ux + tgx;
        vy = uy + tgy;
        vz = uz + tgz;
    }
    const Calc ex = s * vx, ey = s * vy, ez = s * vz;  // equilibrium velocity
    const Calc uu = ex * ex + ey * ey + ez * ez;
    Calc P = Calc(0), S = Calc(0), Vx = Calc(0), Vy = Calc(0), Vz = Calc(0);
    static_for<0, L::Q>([&](auto qc) {
        constexpr int q = decltype(qc)::value;
        const Calc lw = lam[q] * static_cast<Calc>(weight<L>(q));
        const Calc cu = Calc(L::cx(q)) * ex + Calc(L::cy(q)) * ey + Calc(L::cz(q)) * ez;
        P += lw * (Calc(1) + Calc(3) * cu + Calc(4.5) * cu * cu - Calc(1.5) * uu);
        const Calc k = lw * (Calc(3) + Calc(9) * cu);
        if constexpr (L::cx(q) != 0) Vx += Calc(L::cx(q)) * k;
        if constexpr (L::cy(q) != 0) Vy += Calc(L::cy(q)) * k;
        if constexpr (L::cz(q) != 0) Vz += Calc(L::cz(q)) * k;
        S += lw;
    });
    Vx -= Calc(3) * S * ex;
    Vy -= Calc(3) * S * ey;
    Vz -= Calc(3) * S * ez;
    const Calc base = P - s * (Vx * ux + Vy * uy + Vz * uz);
    const Calc one_m_omega = Calc(1) - omega;
    static_for<0, L::Q>([&](auto qc) {
        constexpr int q = decltype(qc)::value;
        const Calc vc = Calc(L::cx(q)) * Vx + Calc(L::cy(q)) * Vy + Calc(L::cz(q)) * Vz;
        mu[q] = one_m_omega * lam[q] + omega * (base + s * vc);
    });
    if constexpr (BRINK)
        return -omega * rho * (Vx * vx + Vy * vy + Vz * vz);
    else
        return Calc(0);
}

// Backward step of one BULK cell (all Q neighbours fluid).
template <class L, Mode M, bool BRINK>
__device__ __forceinline__ void adj_bulk_cell(const AdjArgs& a, int x, int y, int z, int64_t e) {
    using IO = Io<M, L>;
    using T = typename IO::T;
    using Calc = typename IO::Calc;
    constexpr int Q = L::Q;
    const FieldLayout& Ly = a.L;
    const int xm = wrap1(x - 1, Ly.xl), xp = wrap1(x + 1, Ly.xl);
    const int ym = wrap1(y - 1, Ly.yl), yp = wrap1(y + 1, Ly.yl);
    const int zm = wrap1(z - 1, Ly.zl), zp = wrap1(z + 1, Ly.zl);
    const T* __restrict__ gp = static_cast<const T*>(a.gprev);
    const Calc* __restrict__ mi = static_cast<const Calc*>(a.mu_in);

    Calc h[Q], lam[Q];
    static_for<0, Q>([&](auto qc) {
        constexpr int q = decltype(qc)::value;
        constexpr int cx = L::cx(q), cy = L::cy(q), cz = L::cz(q);
        // forward pull from x - c_q
        const int sx = cx > 0 ? xm : (cx < 0 ? xp : x);
        const int sy = cy > 0 ? ym : (cy < 0 ? yp : y);
        const int sz = cz > 0 ? zm : (cz < 0 ? zp : z);
        h[q] = IO::template to_calc<q>(gp[q * Ly.plane + cell_elem(Ly, sx, sy, sz)]);
    });
    if (a.terminal) {
        const T* __restrict__ gc = static_cast<const T*>(a.gcur);
        Calc gf[Q];
        static_for<0, Q>([&](auto qc) {
            constexpr int q = decltype(qc)::value;
            gf[q] = IO::template to_calc<q>(gc[q * Ly.plane + e]);
        });
        objective_grad<L, Calc>(a, x, y, z, gf, lam);
    } else {
        static_for<0, Q>([&](auto qc) {
            constexpr int q = decltype(qc)::value;
            constexpr int cx = L::cx(q), cy = L::cy(q), cz = L::cz(q);
            // reverse pull from x + c_q
            const int nx = cx > 0 ? xp : (cx < 0 ? xm : x);
            const int ny = cy > 0 ? yp : (cy < 0 ? ym : y);
            const int nz = cz > 0 ? zp : (cz < 0 ? zm : z);
            lam[q] = mi[q * Ly.plane + cell_elem(Ly, nx, ny, nz)];
        });
    }
    Calc alpha = Calc(0);
    if constexpr (BRINK) alpha = static_cast<const Calc*>(a.alpha)[e];
    Calc mu[Q];
    const Calc da = adjoint_collide<L, Calc, BRINK>(
        h, lam, mu, static_cast<Calc>(a.omega), alpha, static_cast<Calc>(a.tau * a.g[0]),
        static_cast<Calc>(a.tau * a.g[1]), static_cast<Calc>(a.tau * a.g[2]));
    Calc* __restrict__ mo = static_cast<Calc*>(a.mu_out);
    static_for<0, Q>([&](auto qc) {
        constexpr int q = decltype(qc)::value;
        mo[q * Ly.plane + e] = mu[q];
    });
    if constexpr (BRINK) a.grad_alpha[e] += static_cast<double>(da);
}

// Backward step for BULK cells: one cell per thread, blocks and grid as
// k_bulk with V = 1.
template <class L, Mode M, bool BRINK>
__global__ void __launch_bounds__(256) k_adj_bulk(AdjArgs a, int nbx, int nby) {
    const FieldLayout& Ly = a.L;
    const int b = blockIdx.x;
    const int bx = b % nbx;
    const int t = b / nbx;
    const int by = t % nby;
    const int z = 1 + t / nby;
    const int x = 1 + bx * blockDim.x + threadIdx.x;
    const int y = 1 + by * blockDim.y + threadIdx.y;
    if (x > Ly.xl || y > Ly.yl) return;
    const int64_t e = cell_elem(Ly, x, y, z);
    if (kind_at(a.kind, e) != KIND_BULK) return;
    adj_bulk_cell<L, M, BRINK>(a, x, y, z, e);
}

// Backward step of one fluid cell of any kind, with the adjoint boundary
// rules of NoSlip, MovingWall, Inflow, Outflow and Pressure sources (FreeSlip
// is rejected on the host). gw (3 doubles) accumulates dJ/du_wall of this
// cell's moving-wall links.
template <class L, Mode M, bool BRINK>
__device__ __forceinline__ void adj_cell(const AdjArgs& a, int x, int y, int z, double* __restrict__ gw) {
    using IO = Io<M, L>;
    using T = typename IO::T;
    using Calc = typename IO::Calc;
    constexpr int Q = L::Q;
    const FieldLayout& Ly = a.L;
    const int64_t e = cell_elem(Ly, x, y, z);
    const T* __restrict__ gp = static_cast<const T*>(a.gprev);
    const Calc* __restrict__ mi = static_cast<const Calc*>(a.mu_in);

    // forward pre-collision state h^t(x), boundary rules on x's own g^{t-1}
    Calc own[Q];
    static_for<0, Q>([&](auto qc) {
        constexpr int q = decltype(qc)::value;
        own[q] = IO::template to_calc<q>(gp[q * Ly.plane + e]);
    });
    StepArgs sa{};
    sa.L = Ly;
    sa.src = a.gprev;
    sa.kind = a.kind;
    sa.conds = a.conds;
    Calc h[Q];
    Calc rho_own = Calc(1);
    static_for<0, Q>([&](auto qc) { rho_own += own[decltype(qc)::value]; });
    static_for<0, Q>([&](auto qc) {
        constexpr int q = decltype(qc)::value;
        const int sx = wrap_if(x - L::cx(q), Ly.xl, Ly.per[0]);
        const int sy = wrap_if(y - L::cy(q), Ly.yl, Ly.per[1]);
        const int sz = wrap_if(z - L::cz(q), Ly.zl, Ly.per[2]);
        const int64_t se = cell_elem(Ly, sx, sy, sz);
        const uint8_t k = kind_at(a.kind, se);
        if (k < KIND_COND0 || a.materialised) {
            h[q] = IO::template to_calc<q>(gp[q * Ly.plane + se]);
        } else {
            bool stale;
            // (the adjoint takes no FreeSlip condition)
            h[q] = pull_nonfluid<L, M, q, false>(sa, cond_at(a.conds, k - KIND_COND0), own, x, y, z, sx, sy, sz, se, stale);
        }
    });

    Calc lam[Q];
    if (a.terminal) {
        const T* __restrict__ gc = static_cast<const T*>(a.gcur);
        Calc gf[Q];
        static_for<0, Q>([&](auto qc) {
            constexpr int q = decltype(qc)::value;
            gf[q] = IO::template to_calc<q>(gc[q * Ly.plane + e]);
        });
        objective_grad<L, Calc>(a, x, y, z, gf, lam);
    } else {
        Calc wall_term = Calc(0);
        uint32_t outlinks = 0;  // bit q: x pulls q from an outflow or pressure cell
        static_for<0, Q>([&](auto qc) {
            constexpr int q = decltype(qc)::value;
            // consumer of g_q(x) in step t+1: the cell x + c_q, or x itself via a wall
            const int nx = wrap_if(x + L::cx(q), Ly.xl, Ly.per[0]);
            const int ny = wrap_if(y + L::cy(q), Ly.yl, Ly.per[1]);
            const int nz = wrap_if(z + L::cz(q), Ly.zl, Ly.per[2]);
            const int64_t ne = cell_elem(Ly, nx, ny, nz);
            const uint8_t k = kind_at(a.kind, ne);
            if (k < KIND_COND0) {
                lam[q] = mi[q * Ly.plane + ne];
            } else {
                const int kind = cond_at(a.conds, k - KIND_COND0).kind;
                const Calc back = mi[(Q - 1 - q) * Ly.plane + e];
                lam[q] = (kind == BC_NOSLIP || kind == BC_MOVINGWALL)
                             ? back
                             : ((kind == BC_OUTFLOW || kind == BC_PRESSURE) ? -back : Calc(0));
            }
            // x pulls direction q from the wall at x - c_q: density coupling
            const int sx = wrap_if(x - L::cx(q), Ly.xl, Ly.per[0]);
            const int sy = wrap_if(y - L::cy(q), Ly.yl, Ly.per[1]);
            const int sz = wrap_if(z - L::cz(q), Ly.zl, Ly.per[2]);
            const uint8_t ks = kind_at(a.kind, cell_elem(Ly, sx, sy, sz));
            if (ks >= KIND_COND0 && ks < KIND_INERT) {
                const GCond& cd = cond_at(a.conds, ks - KIND_COND0);
                if (cd.kind == BC_MOVINGWALL) {
                    const double cu = L::cx(q) * cd.vel[0] + L::cy(q) * cd.vel[1] + L::cz(q) * cd.vel[2];
                    wall_term += mi[q * Ly.plane + e] * static_cast<Calc>(6.0 * weight<L>(q) * cu);
                }
                if (cd.kind == BC_OUTFLOW || cd.kind == BC_PRESSURE) outlinks |= 1u << q;
            }
        });
        Calc Wx = Calc(0), Wy = Calc(0), Wz = Calc(0);
        if (outlinks) {
            // momentum of x's own g^t: the forward collision of h, as stored
            Calc gt[Q];
            static_for<0, Q>([&](auto qc) { gt[decltype(qc)::value] = h[decltype(qc)::value]; });
            Calc al = Calc(0);
            if constexpr (BRINK) al = static_cast<const Calc*>(a.alpha)[e];
            collide_shifted<L, Calc, BRINK, kRcp<M>, IO::SX>(gt, static_cast<Calc>(a.omega), al,
                                                            static_cast<Calc>(a.tau * a.g[0]),
                                                            static_cast<Calc>(a.tau * a.g[1]),
                                                            static_cast<Calc>(a.tau * a.g[2]));
            OwnAcc<L, Calc> acc;
            static_for<0, Q>([&](auto qc) {
                constexpr int q = decltype(qc)::value;
                acc.template add<q>(IO::template to_calc<q>(IO::template store<q>(gt[q])));
            });
            static_for<0, Q>([&](auto qc) {
                constexpr int q = decltype(qc)::value;
                if (outlinks & (1u << q)) {
                    const int sx = wrap_if(x - L::cx(q), Ly.xl, Ly.per[0]);
                    const int sy = wrap_if(y - L::cy(q), Ly.yl, Ly.per[1]);
                    const int sz = wrap_if(z - L::cz(q), Ly.zl, Ly.per[2]);
                    const GCond& cd = cond_at(a.conds, kind_at(a.kind, cell_elem(Ly, sx, sy, sz)) - KIND_COND0);
                    const Calc ir = Calc(1) / static_cast<Calc>(cd.rho);
                    const Calc ux = acc.r.m0 * ir, uy = acc.r.m1 * ir, uz = acc.r.m2 * ir;
                    const Calc cu = Calc(L::cx(q)) * ux + Calc(L::cy(q)) * uy + Calc(L::cz(q)) * uz;
                    const Calc f = mi[q * Ly.plane + e] * static_cast<Calc>(2.0 * weight<L>(q));
                    Wx += f * (Calc(9) * cu * Calc(L::cx(q)) - Calc(3) * ux);
                    Wy += f * (Calc(9) * cu * Calc(L::cy(q)) - Calc(3) * uy);
                    Wz += f * (Calc(9) * cu * Calc(L::cz(q)) - Calc(3) * uz);
                }
            });
        }
        static_for<0, Q>([&](auto qc) {
            constexpr int q = decltype(qc)::value;
            lam[q] += wall_term + (Calc(L::cx(q)) * Wx + Calc(L::cy(q)) * Wy + Calc(L::cz(q)) * Wz);
        });
    }

    Calc alpha = Calc(0);
    if constexpr (BRINK) alpha = static_cast<const Calc*>(a.alpha)[e];
    Calc mu[Q];
    const Calc da = adjoint_collide<L, Calc, BRINK>(
        h, lam, mu, static_cast<Calc>(a.omega), alpha, static_cast<Calc>(a.tau * a.g[0]),
        static_cast<Calc>(a.tau * a.g[1]), static_cast<Calc>(a.tau * a.g[2]));
    Calc* __restrict__ mo = static_cast<Calc*>(a.mu_out);
    static_for<0, Q>([&](auto qc) {
        constexpr int q = decltype(qc)::value;
        mo[q * Ly.plane + e] = mu[q];
    });
    if constexpr (BRINK) a.grad_alpha[e] += static_cast<double>(da);

    // dJ/du_wall of this step: h_q = g_qbar + 6 w_q rho c_q.u_w on moving-wall links
    if (!a.materialised) {
        bool walls = false;
        double gw0 = 0, gw1 = 0, gw2 = 0;
        static_for<0, Q>([&](auto qc) {
            constexpr int q = decltype(qc)::value;
            const int sx = wrap_if(x - L::cx(q), Ly.xl, Ly.per[0]);
            const int sy = wrap_if(y - L::cy(q), Ly.yl, Ly.per[1]);
            const int sz = wrap_if(z - L::cz(q), Ly.zl, Ly.per[2]);
            const uint8_t ks = kind_at(a.kind, cell_elem(Ly, sx, sy, sz));
            if (ks >= KIND_COND0 && ks < KIND_INERT && cond_at(a.conds, ks - KIND_COND0).kind == BC_MOVINGWALL) {
                const double f = static_cast<double>(mu[q] * rho_own) * 6.0 * weight<L>(q);
                gw0 += f * L::cx(q);
                gw1 += f * L::cy(q);
                gw2 += f * L::cz(q);
                walls = true;
            }
        });
        if (walls) {
            Glob<double>* g = glob(gw);
            g[0] += gw0;
            g[1] += gw1;
            g[2] += gw2;
        }
    }
}

// Backward step of one fluid cell of any kind for kernels that give every
// cell of a region to a lane (batched cell tiles). All lanes pull h and
// reverse-pull lambda plainly, collide and store together; a NEAR lane then
// replaces the values on its boundary links and adds the wall terms, with the
// operations of adj_cell in the same order, so mu and the gradients carry
// the same bits as k_adj_bulk / k_adj_near. B32 indexes with 32-bit byte
// offsets (planes of the field and of mu under 4 GiB).
template <class L, Mode M, bool BRINK, bool B32 = false>
__device__ __forceinline__ void adj_any_cell(const AdjArgs& a, int x, int y, int z, typename ElemIdx<B32>::type e,
                                             bool near, double* __restrict__ gw) {
    using IO = Io<M, L>;
    using T = typename IO::T;
    using Calc = typename IO::Calc;
    using I = typename ElemIdx<B32>::type;
    static_assert(!B32 || sizeof(T) == sizeof(Calc), "mu planes must be no larger than field planes");
    constexpr int Q = L::Q;
    const FieldLayout& Ly = a.L;
    const T* __restrict__ gp = static_cast<const T*>(a.gprev);
    const Calc* __restrict__ mi = static_cast<const Calc*>(a.mu_in);
    const bool wraps = (Ly.per[0] && (x == 1 || x == Ly.xl)) || (Ly.per[1] && (y == 1 || y == Ly.yl)) ||
                       (Ly.per[2] && (z == 1 || z == Ly.zl));
    // element offset of x + s c_q (s = -1: source, +1: consumer)
    auto nb_elem = [&](auto qc, int sgn) -> I {
        constexpr int q = decltype(qc)::value;
        if (!wraps)
            return e + static_cast<I>(sgn * (L::cx(q) + L::cy(q) * Ly.px + L::cz(q) * Ly.pxy));
        return elem_at<B32>(Ly, wrap_if(x + sgn * L::cx(q), Ly.xl, Ly.per[0]),
                            wrap_if(y + sgn * L::cy(q), Ly.yl, Ly.per[1]), wrap_if(z + sgn * L::cz(q), Ly.zl, Ly.per[2]));
    };

    Calc h[Q];
    static_for<0, Q>([&](auto qc) {
        constexpr int q = decltype(qc)::value;
        h[q] = IO::template to_calc<q>(plane_ref<B32>(gp, q, Ly.plane, nb_elem(qc, -1)));
    });
    if (near && !a.materialised) {
        StepArgs sa{};
        sa.L = Ly;
        sa.src = a.gprev;
        sa.kind = a.kind;
        sa.conds = a.conds;
        const OwnSums<Calc> os = own_sums_at<L, M, B32>(gp, Ly.plane, e);
        auto own = [&](auto kc) {
            constexpr int k = decltype(kc)::value;
            return IO::template to_calc<k>(plane_ref<B32>(gp, k, Ly.plane, e));
        };
        static_for<0, Q>([&](auto qc) {
            constexpr int q = decltype(qc)::value;
            const int sx = wrap_if(x - L::cx(q), Ly.xl, Ly.per[0]);
            const int sy = wrap_if(y - L::cy(q), Ly.yl, Ly.per[1]);
            const int sz = wrap_if(z - L::cz(q), Ly.zl, Ly.per[2]);
            const I se = elem_at<B32>(Ly, sx, sy, sz);
            const uint8_t k = kind_at(a.kind, se);
            if (k >= KIND_COND0) {
                bool stale;
                h[q] = pull_boundary<L, M, q, false, B32>(sa, cond_at(a.conds, k - KIND_COND0), os, own, x, y, z, sx, sy, sz,
                                                          se, stale);
            }
        });
    }

    Calc lam[Q];
    if (a.terminal) {
        const T* __restrict__ gc = static_cast<const T*>(a.gcur);
        Calc gf[Q];
        static_for<0, Q>([&](auto qc) {
            constexpr int q = decltype(qc)::value;
            gf[q] = IO::template to_calc<q>(plane_ref<B32>(gc, q, Ly.plane, e));
        });
        objective_grad<L, Calc>(a, x, y, z, gf, lam);
    } else {
        static_for<0, Q>([&](auto qc) {
            constexpr int q = decltype(qc)::value;
            lam[q] = plane_ref<B32>(mi, q, Ly.plane, nb_elem(qc, 1));
        });
        if (near) {
            Calc wall_term = Calc(0);
            static_for<0, Q>([&](auto qc) {
                constexpr int q = decltype(qc)::value;
                const uint8_t k = kind_at(a.kind, nb_elem(qc, 1));
                if (k >= KIND_COND0) {
                    const int kind = cond_at(a.conds, k - KIND_COND0).kind;
                    lam[q] = (kind == BC_NOSLIP || kind == BC_MOVINGWALL) ? plane_ref<B32>(mi, Q - 1 - q, Ly.plane, e)
                                                                          : Calc(0);
                }
                const uint8_t ks = kind_at(a.kind, nb_elem(qc, -1));
                if (ks >= KIND_COND0 && ks < KIND_INERT) {
                    const GCond& cd = cond_at(a.conds, ks - KIND_COND0);
                    if (cd.kind == BC_MOVINGWALL) {
                        const double cu = L::cx(q) * cd.vel[0] + L::cy(q) * cd.vel[1] + L::cz(q) * cd.vel[2];
                        wall_term += plane_ref<B32>(mi, q, Ly.plane, e) * static_cast<Calc>(6.0 * weight<L>(q) * cu);
                    }
                }
            });
            static_for<0, Q>([&](auto qc) { lam[decltype(qc)::value] += wall_term; });
        }
    }

    Calc alpha = Calc(0);
    if constexpr (BRINK) alpha = static_cast<const Calc*>(a.alpha)[e];
    Calc mu[Q];
    const Calc da = adjoint_collide<L, Calc, BRINK>(
        h, lam, mu, static_cast<Calc>(a.omega), alpha, static_cast<Calc>(a.tau * a.g[0]),
        static_cast<Calc>(a.tau * a.g[1]), static_cast<Calc>(a.tau * a.g[2]));
    Calc* __restrict__ mo = static_cast<Calc*>(a.mu_out);
    static_for<0, Q>([&](auto qc) {
        constexpr int q = decltype(qc)::value;
        plane_ref<B32>(mo, q, Ly.plane, e) = mu[q];
    });
    if constexpr (BRINK) a.grad_alpha[e] += static_cast<double>(da);

    if (near && !a.materialised) {
        Calc rho_own = Calc(1);
        static_for<0, Q>([&](auto qc) {
            constexpr int q = decltype(qc)::value;
            rho_own += IO::template to_calc<q>(plane_ref<B32>(gp, q, Ly.plane, e));
        });
        bool walls = false;
        double gw0 = 0, gw1 = 0, gw2 = 0;
        static_for<0, Q>([&](auto qc) {
            constexpr int q = decltype(qc)::value;
            const uint8_t ks = kind_at(a.kind, nb_elem(qc, -1));
            if (ks >= KIND_COND0 && ks < KIND_INERT && cond_at(a.conds, ks - KIND_COND0).kind == BC_MOVINGWALL) {
                const double f = static_cast<double>(mu[q] * rho_own) * 6.0 * weight<L>(q);
                gw0 += f * L::cx(q);
                gw1 += f * L::cy(q);
                gw2 += f * L::cz(q);
                walls = true;
            }
        });
        if (walls) {
            Glob<double>* g = glob(gw);
            g[0] += gw0;
            g[1] += gw1;
            g[2] += gw2;
        }
    }
}

// Backward step for NEAR cells (listed): gradients per NEAR cell index.
template <class L, Mode M, bool BRINK>
__global__ void __launch_bounds__(128) k_adj_near(AdjArgs a) {
    const int64_t i = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
    if (i >= a.nb) return;
    const CellCoord c = a.blist[i];
    adj_cell<L, M, BRINK>(a, c.x, c.y, c.z, a.grad_wall + 3 * i);
}

// Objective value over interior fluid cells in fp64 (component 0), and for
// the dual mode its directional derivative (component 1): per-block
// partials, then k_sum_partials.
template <class L, Mode M>
__global__ void __launch_bounds__(256) k_objective(FieldLayout Ly, const void* src, const uint8_t* kind,
                                                   int objective, int plane_x, const double* target,
                                                   double* partial) {
    using IO = Io<M, L>;
    using T = typename IO::T;
    using Calc = typename IO::Calc;
    const int64_t n = static_cast<int64_t>(Ly.xl) * Ly.yl * Ly.zl;
    double s[4] = {0.0, 0.0, 0.0, 0.0};
    const T* g = static_cast<const T*>(src);
    for (int64_t i = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; i < n;
         i += static_cast<int64_t>(gridDim.x) * blockDim.x) {
        const int x = 1 + static_cast<int>(i % Ly.xl);
        const int y = 1 + static_cast<int>((i / Ly.xl) % Ly.yl);
        const int z = 1 + static_cast<int>(i / (static_cast<int64_t>(Ly.xl) * Ly.yl));
        if (objective == OBJ_PLANE_MISFIT && x != plane_x) continue;
        const int64_t e = cell_elem(Ly, x, y, z);
        if (kind[e] >= KIND_COND0) continue;
        Calc drho = Calc(0), jx = Calc(0), jy = Calc(0), jz = Calc(0);
        static_for<0, L::Q>([&](auto qc) {
            constexpr int q = decltype(qc)::value;
            const Calc v = IO::template to_calc<q>(g[q * Ly.plane + e]);
            drho += v;
            jx += Calc(L::cx(q)) * v;
            jy += Calc(L::cy(q)) * v;
            jz += Calc(L::cz(q)) * v;
        });
        const Calc rho = Calc(1) + drho;
        Calc ux = jx / rho, uy = jy / rho, uz = jz / rho;
        Calc val;
        if (objective == OBJ_PLANE_MISFIT) {
            const double* t = target + 3 * ((static_cast<int64_t>(z) - 1) * Ly.yl + (y - 1));
            ux -= static_cast<Calc>(t[0]);
            uy -= static_cast<Calc>(t[1]);
            uz -= static_cast<Calc>(t[2]);
            val = ux * ux + uy * uy + uz * uz;
        } else {
            val = Calc(0.5) * (ux * ux + uy * uy + uz * uz);
        }
        if constexpr (M == Mode::DUAL64) {
            s[0] += val.v;
            s[1] += val.d;
        } else {
            s[0] += static_cast<double>(val);
        }
    }
    block_sum4(s);
    if (threadIdx.x == 0)
#pragma unroll
        for (int k = 0; k < 4; ++k) partial[4 * blockIdx.x + k] = s[k];
}

// Storage-mode field -> dual field (value = stored shifted value, tangent 0).
template <class L, Mode M>
__global__ void k_to_dual(FieldLayout Ly, const void* src, Dual<double>* dst, int q0) {
    using IO = Io<M, L>;
    using T = typename IO::T;
    const int64_t n = Ly.plane;
    const int64_t i = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
    if (i >= n) return;
    const T* s = static_cast<const T*>(src);
    static_for<0, L::Q>([&](auto qc) {
        constexpr int q = decltype(qc)::value;
        dst[q * n + i] = Dual<double>(static_cast<double>(IO::template to_calc<q>(s[q * n + i])), 0.0);
    });
}

}  // namespace lbm_amd
