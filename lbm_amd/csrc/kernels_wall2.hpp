// Two time steps per pass for domains with boundary cells: walls, solid
// obstacles, inflow and outflow faces (NoSlip, MovingWall, Inflow, Outflow,
// Pressure; fp32 storage). A domain with a FreeSlip condition steps singly,
// since that rule reads cells other than the one it serves.
//
// The march is that of k_bulk2 (kernels_tb.hpp): a block owns a TX x TY
// tile and a chunk of z planes; per plane it does step 1 on the tile and its
// one-cell ring, passes g1 through LDS, and step 2 on the tile one plane
// behind. What boundaries add:
//   - before the pass, k_emit_rules writes into every boundary link (b, q)
//     the value the reference's boundary pass leaves there
//     (include/domain.hpp:157-165): the rule of b for direction q, from the
//     stored values of the fluid cell x = b + c_q. Each link has one such x,
//     so one thread per near-boundary cell x writes its links. Step 1 then
//     pulls every value plainly, as the single-step cell kernel does after
//     patching its pulls with the same rule values;
//   - non-fluid cells (ghost layer, solids) do neither step and send nothing;
//   - step 2 needs the rule values of x's non-fluid sources b = x - c_k from
//     x's own step-1 values. No neighbour sends those (b is not fluid), so
//     after step 1 the thread that computed x evaluates, for every link k of
//     x's mask (lmask: bit k when x - c_k is not fluid), the rule value of
//     (b, k) and writes it into x's own slot k. Nothing else writes that
//     slot: its only regular sender would be b.
// The storage format is fp32, the arithmetic type, so a rule value written
// to memory or LDS is the value the single step computes, and a pass equals
// two single steps bit for bit on every fluid cell.
//
// Coordinates along periodic axes wrap as in k_bulk2; along walled axes the
// ring reaches the ghost layer (0 and l + 1), whose cells are not fluid.
#pragma once

#include "kernels.hpp"
#include "kernels_tb.hpp"
#include "ops.hpp"

namespace lbm_amd {

__device__ __forceinline__ int wrap_axis(int v, int l, int per) { return per ? wrap1(v, l) : v; }

// One thread per near-boundary cell x (the domain's list): for every
// direction q whose source b = x - c_q is not fluid, b[q] in src <- the rule
// value of (b, q) from x's stored values. Reads fluid cells, writes
// non-fluid cells only.
template <class L, Mode M, int RULES>
__global__ void __launch_bounds__(256) k_emit_rules(StepArgs a) {
    using IO = Io<M, L>;
    using T = typename IO::T;
    using Calc = typename IO::Calc;
    constexpr int Q = L::Q;
    const int64_t i = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
    if (i >= a.nb) return;
    const CellCoord c = a.blist[i];
    const FieldLayout& Ly = a.L;
    T* __restrict__ buf = static_cast<T*>(a.dst);
    const uint32_t e = elem_at<true>(Ly, c.x, c.y, c.z);
    const uint32_t lm = glob(a.lmask)[e];
    const OwnSums<Calc> os = own_sums_at<L, M, true>(buf, Ly.plane, e);
    static_for<0, Q>([&](auto qc) {
        constexpr int q = decltype(qc)::value;
        constexpr int qb = Q - 1 - q;
        if (lm & (1u << q)) {
            const uint32_t se = elem_at<true>(Ly, wrap_if(c.x - L::cx(q), Ly.xl, Ly.per[0]),
                                              wrap_if(c.y - L::cy(q), Ly.yl, Ly.per[1]),
                                              wrap_if(c.z - L::cz(q), Ly.zl, Ly.per[2]));
            const Calc finv = IO::template to_calc<qb>(plane_ref<true>(buf, qb, Ly.plane, e));
            plane_ref<true>(buf, q, Ly.plane, se) = IO::template from_calc<q>(
                rule_value<L, M, q, RULES>(cond_at(a.conds, kind_at(a.kind, se) - KIND_COND0), os, finv));
        }
    });
}

template <class L, Mode M, int RULES>
__global__ void __launch_bounds__(256) k_wall2(StepArgs a, int ntx, int nty, int zc, int nxcd) {
    using IO = Io<M, L>;
    using T = typename IO::T;
    using Calc = typename IO::Calc;
    using S = TbSlots<L>;
    constexpr int Q = L::Q;
    static_assert(sizeof(T) == sizeof(Calc), "boundary values cross memory in the arithmetic type");
    __shared__ T lds[S::NSLOT][TB_CELLS];

    const FieldLayout& Ly = a.L;
    const int tid = threadIdx.x;
    int b = tb_logical_block(static_cast<int>(blockIdx.x), static_cast<int>(gridDim.x), nxcd);
    const int x0 = 1 + (b % ntx) * TB_TX;
    b /= ntx;
    const int y0 = 1 + (b % nty) * TB_TY;
    b /= nty;
    const int z0 = 1 + b * zc;
    const int z1 = min(z0 + zc - 1, Ly.zl);
    const T* __restrict__ src = static_cast<const T*>(a.src);
    T* __restrict__ dst = static_cast<T*>(a.dst);
    // byte offsets within a plane fit 32 bits (the host checks plane bytes)
    constexpr uint32_t TS = sizeof(T);
    const uint32_t py = static_cast<uint32_t>(Ly.px) * TS, pz = static_cast<uint32_t>(Ly.pxy) * TS;
    const uint32_t off = static_cast<uint32_t>(Ly.off) * TS;

    // this thread's patch cells: tile-local (pi, pj) in [-1, TX] x [-1, TY];
    // gx, gy are their coordinates in the ghosted box
    int pi[TB_PASS], pj[TB_PASS], gx[TB_PASS], gy[TB_PASS];
    bool live[TB_PASS];
#pragma unroll
    for (int k = 0; k < TB_PASS; ++k) {
        const int p = k * 256 + tid;
        pi[k] = p % TB_PX - 1;
        pj[k] = p / TB_PX - 1;
        const int ux = x0 + pi[k], uy = y0 + pj[k];
        live[k] = p < TB_NP && ux <= Ly.xl + 1 && uy <= Ly.yl + 1;
        gx[k] = wrap_axis(ux, Ly.xl, Ly.per[0]);
        gy[k] = wrap_axis(uy, Ly.yl, Ly.per[1]);
    }

    T hold[TB_PASS][Q];  // step-1 values of this plane's patch cells
    T prev[TB_PASS][Q];  // the c_z = +1 values of the previous plane's
    bool fl[TB_PASS], pfl[TB_PASS];  // patch cell fluid on this / the previous plane
    uint32_t wm[TB_PASS];            // link mask of a near-boundary tile cell (bit k: x - c_k not fluid)
    // Cell kinds are loaded one plane ahead (patch cells) or at the top of
    // the iteration (step-2 cells), so the field loads do not wait on them.
    uint8_t kcur[TB_PASS], knext[TB_PASS];
    const uint32_t rowb[TB_PASS] = {py * static_cast<uint32_t>(gy[0]) + TS * static_cast<uint32_t>(gx[0]),
                                    py * static_cast<uint32_t>(gy[1]) + TS * static_cast<uint32_t>(gx[1]),
                                    py * static_cast<uint32_t>(gy[2]) + TS * static_cast<uint32_t>(gx[2])};
    static_assert(TB_PASS == 3, "rowb lists one entry per patch cell");
    auto plane_off = [&](int zz) -> uint32_t {
        return off + pz * static_cast<uint32_t>(wrap_axis(zz, Ly.zl, Ly.per[2]));
    };
#pragma unroll
    for (int k = 0; k < TB_PASS; ++k) {
        pfl[k] = false;
        static_for<0, Q>([&](auto qc) { prev[k][decltype(qc)::value] = T{}; });
        knext[k] = live[k] ? kind_at(a.kind, (plane_off(z0 - 1) + rowb[k]) / TS) : uint8_t(KIND_INERT);
    }
    for (int zs = z0 - 1; zs <= z1 + 1; ++zs) {
#pragma unroll
        for (int k = 0; k < TB_PASS; ++k) {
            kcur[k] = knext[k];
            knext[k] = (live[k] && zs + 1 <= z1 + 1) ? kind_at(a.kind, (plane_off(zs + 1) + rowb[k]) / TS)
                                                      : uint8_t(KIND_INERT);
        }
        uint8_t kout[TB_OUT];
#pragma unroll
        for (int k = 0; k < TB_OUT; ++k) {
            const int c = k * 256 + tid;
            const int ox = x0 + c % TB_TX, oy = y0 + c / TB_TX;
            kout[k] = (zs - 1 >= z0 && ox <= Ly.xl && oy <= Ly.yl)
                          ? kind_at(a.kind, static_cast<uint32_t>(Ly.off) + ox + static_cast<uint32_t>(Ly.pxy) * (zs - 1) +
                                                static_cast<uint32_t>(Ly.px) * oy)
                          : uint8_t(KIND_INERT);
        }
        // (the empty asm keeps the Q plane bases from being hoisted, as in k_bulk2)
        int64_t plane = Ly.plane;
        asm volatile("" : "+s"(plane));
        // (1) step 1 on plane zs
        const int gz = wrap_axis(zs, Ly.zl, Ly.per[2]);
        // planes gz - 1 .. gz + 1; only fluid cells (1 <= gz <= zl) use them
        const uint32_t zs3[3] = {off + pz * static_cast<uint32_t>(wrap_axis(gz - 1, Ly.zl, Ly.per[2])),
                                 off + pz * static_cast<uint32_t>(gz),
                                 off + pz * static_cast<uint32_t>(wrap_axis(gz + 1, Ly.zl, Ly.per[2]))};
#pragma unroll
        for (int k = 0; k < TB_PASS; ++k) {
            fl[k] = false;
            wm[k] = 0;
            if (!live[k]) continue;
            const uint32_t e = (zs3[1] + rowb[k]) / TS;
            const uint8_t kd = kcur[k];
            if (kd >= KIND_COND0) continue;
            fl[k] = true;
            const uint32_t xs[3] = {TS * static_cast<uint32_t>(wrap_axis(gx[k] - 1, Ly.xl, Ly.per[0])),
                                    TS * static_cast<uint32_t>(gx[k]),
                                    TS * static_cast<uint32_t>(wrap_axis(gx[k] + 1, Ly.xl, Ly.per[0]))};
            const uint32_t ys[3] = {py * static_cast<uint32_t>(wrap_axis(gy[k] - 1, Ly.yl, Ly.per[1])),
                                    py * static_cast<uint32_t>(gy[k]),
                                    py * static_cast<uint32_t>(wrap_axis(gy[k] + 1, Ly.yl, Ly.per[1]))};
            Calc f[Q];
            static_for<0, Q>([&](auto qc) {
                constexpr int q = decltype(qc)::value;
                // source x - c_q: index 0 for c = +1, 1 for 0, 2 for -1
                constexpr int ix = 1 - L::cx(q), iy = 1 - L::cy(q), iz = 1 - L::cz(q);
                f[q] = IO::template load<q>(tb_at(src + q * plane, zs3[iz] + ys[iy] + xs[ix]));
            });
            // links whose rule value step 2 will need from this cell
            if (kd == KIND_NEAR && pi[k] >= 0 && pi[k] < TB_TX && pj[k] >= 0 && pj[k] < TB_TY)
                wm[k] = glob(a.lmask)[e];
            collide_shifted<L, Calc, false, kRcp<M>, IO::SX>(f, static_cast<Calc>(a.omega), Calc(0), Calc(0), Calc(0),
                                                            Calc(0));
            static_for<0, Q>([&](auto qc) {
                constexpr int q = decltype(qc)::value;
                const T v = IO::template store<q>(f[q]);
                hold[k][q] = v;
                if constexpr (L::cz(q) < 0) {
                    const int di = pi[k] + L::cx(q), dj = pj[k] + L::cy(q);
                    if (di >= 0 && di < TB_TX && dj >= 0 && dj < TB_TY)
                        lds[S::template slot<q>()][dj * TB_TX + di] = v;
                }
            });
        }
        __syncthreads();
        // (2) step 2 on plane zs - 1
        const int zo = zs - 1;
        if (zo >= z0) {
#pragma unroll
            for (int k = 0; k < TB_OUT; ++k) {
                const int c = k * 256 + tid;
                const int ox = x0 + c % TB_TX, oy = y0 + c / TB_TX;
                if (ox <= Ly.xl && oy <= Ly.yl) {
                    const uint32_t eb = TS * (static_cast<uint32_t>(Ly.off) + ox) + pz * zo + py * oy;
                    if (kout[k] < KIND_COND0) {
                        Calc f[Q];
                        static_for<0, Q>([&](auto qc) {
                            constexpr int q = decltype(qc)::value;
                            f[q] = IO::template load<q>(lds[S::template slot<q>()][c]);
                        });
                        collide_shifted<L, Calc, false, kRcp<M>, IO::SX>(f, static_cast<Calc>(a.omega), Calc(0),
                                                                        Calc(0), Calc(0), Calc(0));
                        static_for<0, Q>([&](auto qc) {
                            constexpr int q = decltype(qc)::value;
                            store_stream1<T>(tb_ptr(dst + q * plane, eb), IO::template store<q>(f[q]));
                        });
                    }
                }
            }
        }
        __syncthreads();
        // sends read by step 2 of the next iteration: c_z = 0 of this plane,
        // c_z = +1 of the previous plane; then the rule values of this
        // plane's fluid tile cells next to non-fluid cells
#pragma unroll
        for (int k = 0; k < TB_PASS; ++k) {
            if (!live[k]) continue;
            static_for<0, Q>([&](auto qc) {
                constexpr int q = decltype(qc)::value;
                if constexpr (L::cz(q) >= 0) {
                    const int di = pi[k] + L::cx(q), dj = pj[k] + L::cy(q);
                    const bool from = L::cz(q) == 0 ? fl[k] : pfl[k];
                    if (from && di >= 0 && di < TB_TX && dj >= 0 && dj < TB_TY)
                        lds[S::template slot<q>()][dj * TB_TX + di] = L::cz(q) == 0 ? hold[k][q] : prev[k][q];
                }
                if constexpr (L::cz(q) > 0) prev[k][q] = hold[k][q];
            });
            pfl[k] = fl[k];
        }
#pragma unroll
        for (int k = 0; k < TB_PASS; ++k) {
            if (!wm[k]) continue;
            OwnAcc<L, Calc> acc;
            static_for<0, Q>([&](auto qc) {
                constexpr int q = decltype(qc)::value;
                acc.template add<q>(IO::template to_calc<q>(hold[k][q]));
            });
            const int cell = pj[k] * TB_TX + pi[k];
            static_for<0, Q>([&](auto qc) {
                constexpr int q = decltype(qc)::value;  // pull direction from the boundary cell b = x - c_q
                constexpr int qb = Q - 1 - q;
                if (wm[k] & (1u << q)) {
                    const uint32_t be =
                        (off + pz * static_cast<uint32_t>(wrap_axis(gz - L::cz(q), Ly.zl, Ly.per[2])) +
                         py * static_cast<uint32_t>(wrap_axis(gy[k] - L::cy(q), Ly.yl, Ly.per[1])) +
                         TS * static_cast<uint32_t>(wrap_axis(gx[k] - L::cx(q), Ly.xl, Ly.per[0]))) /
                        TS;
                    const GCond& cd = cond_at(a.conds, kind_at(a.kind, be) - KIND_COND0);
                    lds[S::template slot<q>()][cell] = IO::template store<q>(IO::template scaled<q>(
                        rule_value<L, M, q, RULES>(cd, acc.r, IO::template to_calc<qb>(hold[k][qb]))));
                }
            });
        }
    }
}

}  // namespace lbm_amd
