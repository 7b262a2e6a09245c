// Two time steps per pass for domains with boundary cells: walls, solid
// obstacles, inflow and outflow faces (NoSlip, MovingWall, Inflow, Outflow,
// Pressure; fp32, fp16 or bf16 storage). A domain with a FreeSlip condition steps singly,
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
//     the thread that computed x keeps all of x's step-1 values in
//     registers (the tile cells of a thread are the cells it also finishes
//     in step 2) and, once step 2 of the plane before is done, evaluates for
//     every link k of x's mask (lmask: bit k when x - c_k is not fluid) the
//     rule value of (b, k) and writes it into x's own slot k. Nothing else
//     writes that slot: its only regular sender would be b. Where all links
//     share one condition (lcond) its data is loaded once for all of them.
// A rule value crosses memory or LDS in the storage format. The single step
// takes its rule values through the storage format too (Io::through_storage:
// as the reference's boundary pass stores them in the field), so a pass
// equals two single steps bit for bit on every fluid cell, in fp32, fp16 and
// bf16 storage.
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
// non-fluid cells only. x's Q values are loaded once, every value is
// computed before the first store, and where all links share one condition
// its data is loaded once, so a thread waits for memory about twice instead
// of twice per link.
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
    if (lm == 0) return;
    const int lc = glob(a.lcond)[e];
    if (lc != LCOND_MIXED && cond_at(a.conds, lc).kind == BC_NOSLIP) {
        // Every link NoSlip: the rule value of link q is the cell's own value
        // of direction q-bar, so only the linked directions are read (a cell
        // on an x wall has Q values in Q different lines, of which the other
        // kinds' rules need all for the density). An unlinked direction reads
        // the first linked one's element instead, a line fetched anyway, so
        // that every load goes out unconditionally ahead of the masked stores.
        const int q0 = __builtin_ctz(lm);
        T v[Q];
        static_for<0, Q>([&](auto qc) {
            constexpr int q = decltype(qc)::value;
            v[q] = plane_ref<true>(buf, ((lm >> q) & 1u) ? Q - 1 - q : Q - 1 - q0, Ly.plane, e);
        });
        static_for<0, Q>([&](auto qc) {
            constexpr int q = decltype(qc)::value;
            if (lm & (1u << q)) {
                const uint32_t se = elem_at<true>(Ly, wrap_if(c.x - L::cx(q), Ly.xl, Ly.per[0]),
                                                  wrap_if(c.y - L::cy(q), Ly.yl, Ly.per[1]),
                                                  wrap_if(c.z - L::cz(q), Ly.zl, Ly.per[2]));
                plane_ref<true>(buf, q, Ly.plane, se) =
                    IO::template from_calc<q>(IO::template to_calc<Q - 1 - q>(v[q]));
            }
        });
        return;
    }
    Calc own[Q];
    OwnAcc<L, Calc> acc;
    static_for<0, Q>([&](auto qc) {
        constexpr int q = decltype(qc)::value;
        own[q] = IO::template to_calc<q>(plane_ref<true>(buf, q, Ly.plane, e));
        acc.template add<q>(own[q]);
    });
    T rv[Q];
    auto put = [&](auto qc, Calc v) {
        constexpr int q = decltype(qc)::value;
        rv[q] = IO::template from_calc<q>(v);
    };
    auto finv = [&](auto qc) { return own[Q - 1 - decltype(qc)::value]; };
    if (lc != LCOND_MIXED) {
        const GCond& cd = cond_at(a.conds, lc);
        const int ck = cd.kind;
        if ((RULES & RULE_MOVINGWALL) != 0 && ck == BC_MOVINGWALL) {
            static_for<0, Q>([&](auto qc) {
                put(qc, moving_wall_rule<L, Calc, decltype(qc)::value>(acc.r, finv(qc), cd));
            });
        } else if ((RULES & RULE_INFLOW) != 0 && ck == BC_INFLOW) {
            static_for<0, Q>([&](auto qc) { put(qc, static_cast<Calc>(cd.geq[decltype(qc)::value])); });
        } else if ((RULES & RULE_OUTFLOW) != 0 && (ck == BC_OUTFLOW || ck == BC_PRESSURE)) {
            static_for<0, Q>([&](auto qc) {
                put(qc, outflow_rule<L, Calc, decltype(qc)::value>(acc.r, finv(qc), cd));
            });
        } else {  // NoSlip
            static_for<0, Q>([&](auto qc) { put(qc, finv(qc)); });
        }
    } else {
        static_for<0, Q>([&](auto qc) {
            constexpr int q = decltype(qc)::value;
            rv[q] = T{};
            if (lm & (1u << q)) {
                const uint32_t se = elem_at<true>(Ly, wrap_if(c.x - L::cx(q), Ly.xl, Ly.per[0]),
                                                  wrap_if(c.y - L::cy(q), Ly.yl, Ly.per[1]),
                                                  wrap_if(c.z - L::cz(q), Ly.zl, Ly.per[2]));
                put(qc, rule_value<L, M, q, RULES>(cond_at(a.conds, kind_at(a.kind, se) - KIND_COND0), acc.r,
                                                   finv(qc)));
            }
        });
    }
    static_for<0, Q>([&](auto qc) {
        constexpr int q = decltype(qc)::value;
        if (lm & (1u << q)) {
            const uint32_t se = elem_at<true>(Ly, wrap_if(c.x - L::cx(q), Ly.xl, Ly.per[0]),
                                              wrap_if(c.y - L::cy(q), Ly.yl, Ly.per[1]),
                                              wrap_if(c.z - L::cz(q), Ly.zl, Ly.per[2]));
            plane_ref<true>(buf, q, Ly.plane, se) = rv[q];
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
    static_assert(TB_PASS == TB_OUT + 1 && TB_NP - TB_CELLS <= 256, "tile cells on k < TB_OUT, the ring on the last k");
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

    // This thread's patch cells, tile-local (pi, pj) in [-1, TX] x [-1, TY]:
    // k < TB_OUT are the tile cells c = k * 256 + tid, the cells whose step
    // 2 this thread also does; the last k is ring cell tid (the rows pj = -1
    // and pj = TY, then the columns pi = -1 and pi = TX). gx, gy are the
    // coordinates in the ghosted box.
    int pi[TB_PASS], pj[TB_PASS], gx[TB_PASS], gy[TB_PASS];
    bool live[TB_PASS];
#pragma unroll
    for (int k = 0; k < TB_PASS; ++k) {
        bool in = true;
        if (k < TB_OUT) {
            const int c = k * 256 + tid;
            pi[k] = c % TB_TX;
            pj[k] = c / TB_TX;
        } else if (tid < TB_PX) {
            pi[k] = tid - 1;
            pj[k] = -1;
        } else if (tid < 2 * TB_PX) {
            pi[k] = tid - TB_PX - 1;
            pj[k] = TB_TY;
        } else if (tid < 2 * TB_PX + TB_TY) {
            pi[k] = -1;
            pj[k] = tid - 2 * TB_PX;
        } else {
            pi[k] = TB_TX;
            pj[k] = tid - 2 * TB_PX - TB_TY;
            in = tid < TB_NP - TB_CELLS;
        }
        const int ux = x0 + pi[k], uy = y0 + pj[k];
        live[k] = in && ux <= Ly.xl + 1 && uy <= Ly.yl + 1;
        gx[k] = wrap_axis(ux, Ly.xl, Ly.per[0]);
        gy[k] = wrap_axis(uy, Ly.yl, Ly.per[1]);
    }

    T hold[TB_PASS][Q];  // step-1 values of this plane's patch cells (ring cells: c_z >= 0 only)
    T prev[TB_PASS][Q];  // the c_z = +1 values of the previous plane's
    bool fl[TB_PASS], pfl[TB_PASS];     // patch cell fluid on this / the previous plane
    uint32_t wm[TB_OUT];  // link masks of this plane's tile cells next to non-fluid cells
    int lcs[TB_OUT];      // their shared condition
#pragma unroll
    for (int k = 0; k < TB_PASS; ++k) {
        pfl[k] = false;
        static_for<0, Q>([&](auto qc) { prev[k][decltype(qc)::value] = T{}; });
    }
    for (int zs = z0 - 1; zs <= z1 + 1; ++zs) {
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
            if (k < TB_OUT) wm[k & (TB_OUT - 1)] = 0;
            if (!live[k]) continue;
            const uint32_t e = (zs3[1] + py * static_cast<uint32_t>(gy[k]) + TS * static_cast<uint32_t>(gx[k])) / TS;
            const uint8_t kd = kind_at(a.kind, e);
            if (kd >= KIND_COND0) continue;
            fl[k] = true;
            // a tile cell of the chunk next to non-fluid cells: its links
            uint32_t lm = 0;
            int lc = LCOND_MIXED;
            if (k < TB_OUT && kd == KIND_NEAR && zs >= z0 && zs <= z1) {
                lm = glob(a.lmask)[e];
                lc = glob(a.lcond)[e];
            }
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
            collide_shifted<L, Calc, false, kRcp<M>, IO::SX>(f, static_cast<Calc>(a.omega), Calc(0), Calc(0), Calc(0),
                                                            Calc(0));
            static_for<0, Q>([&](auto qc) {
                constexpr int q = decltype(qc)::value;
                const T v = IO::template store<q>(f[q]);
                if constexpr (L::cz(q) < 0) {
                    const int di = pi[k] + L::cx(q), dj = pj[k] + L::cy(q);
                    if (di >= 0 && di < TB_TX && dj >= 0 && dj < TB_TY)
                        lds[S::template slot<q>()][dj * TB_TX + di] = v;
                }
                // tile cells keep every value (their rule values are evaluated
                // in phase 3), ring cells the ones sent later
                if (k < TB_OUT || L::cz(q) >= 0) hold[k][q] = v;
            });
            if (k < TB_OUT) {
                wm[k & (TB_OUT - 1)] = lm;
                lcs[k & (TB_OUT - 1)] = lc;
            }
        }
        __syncthreads();
        // (2) step 2 on plane zs - 1: this thread's tile cells c = k * 256 + tid
        const int zo = zs - 1;
        if (zo >= z0) {
#pragma unroll
            for (int k = 0; k < TB_OUT; ++k) {
                if (!pfl[k]) continue;
                const int c = k * 256 + tid;
                const uint32_t eb = TS * (static_cast<uint32_t>(Ly.off) + x0 + pi[k]) + pz * zo + py * (y0 + pj[k]);
                Calc f[Q];
                static_for<0, Q>([&](auto qc) {
                    constexpr int q = decltype(qc)::value;
                    f[q] = IO::template load<q>(lds[S::template slot<q>()][c]);
                });
                collide_shifted<L, Calc, false, kRcp<M>, IO::SX>(f, static_cast<Calc>(a.omega), Calc(0), Calc(0),
                                                                Calc(0), Calc(0));
                static_for<0, Q>([&](auto qc) {
                    constexpr int q = decltype(qc)::value;
                    store_stream1<T>(tb_ptr(dst + q * plane, eb), IO::template store<q>(f[q]));
                });
            }
        }
        __syncthreads();
        // sends read by step 2 of the next iteration: c_z = 0 of this plane,
        // c_z = +1 of the previous plane (held one more plane in registers)
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
        // rule values of this plane's tile cells next to non-fluid cells,
        // from their step-1 values, into their own slots (written by no
        // neighbour: the senders are not fluid); read by step 2 next
#pragma unroll
        for (int k = 0; k < TB_OUT; ++k) {
            const uint32_t lm = wm[k];
            if (!lm) continue;
            const int c = k * 256 + tid;
            OwnAcc<L, Calc> acc;
            static_for<0, Q>([&](auto qc) {
                constexpr int q = decltype(qc)::value;
                acc.template add<q>(IO::template to_calc<q>(hold[k][q]));
            });
            auto finv = [&](auto qc) {
                constexpr int qb = Q - 1 - decltype(qc)::value;
                return IO::template to_calc<qb>(hold[k][qb]);
            };
            T rv[Q];
            auto put = [&](auto qc, Calc v) {
                constexpr int q = decltype(qc)::value;
                rv[q] = IO::template store<q>(IO::template scaled<q>(v));
            };
            const int lc = lcs[k];
            if (lc != LCOND_MIXED) {
                const GCond& cd = cond_at(a.conds, lc);
                const int ck = cd.kind;
                if ((RULES & RULE_MOVINGWALL) != 0 && ck == BC_MOVINGWALL) {
                    static_for<0, Q>([&](auto qc) {
                        put(qc, moving_wall_rule<L, Calc, decltype(qc)::value>(acc.r, finv(qc), cd));
                    });
                } else if ((RULES & RULE_INFLOW) != 0 && ck == BC_INFLOW) {
                    static_for<0, Q>([&](auto qc) { put(qc, static_cast<Calc>(cd.geq[decltype(qc)::value])); });
                } else if ((RULES & RULE_OUTFLOW) != 0 && (ck == BC_OUTFLOW || ck == BC_PRESSURE)) {
                    static_for<0, Q>([&](auto qc) {
                        put(qc, outflow_rule<L, Calc, decltype(qc)::value>(acc.r, finv(qc), cd));
                    });
                } else {  // NoSlip
                    static_for<0, Q>([&](auto qc) { put(qc, finv(qc)); });
                }
            } else {
                static_for<0, Q>([&](auto qc) {
                    constexpr int q = decltype(qc)::value;  // pull direction from b = x - c_q
                    rv[q] = T{};
                    if (lm & (1u << q)) {
                        const uint32_t be =
                            (off + pz * static_cast<uint32_t>(wrap_axis(gz - L::cz(q), Ly.zl, Ly.per[2])) +
                             py * static_cast<uint32_t>(wrap_axis(gy[k] - L::cy(q), Ly.yl, Ly.per[1])) +
                             TS * static_cast<uint32_t>(wrap_axis(gx[k] - L::cx(q), Ly.xl, Ly.per[0]))) /
                            TS;
                        put(qc, rule_value<L, M, q, RULES>(cond_at(a.conds, kind_at(a.kind, be) - KIND_COND0),
                                                           acc.r, finv(qc)));
                    }
                });
            }
            static_for<0, Q>([&](auto qc) {
                constexpr int q = decltype(qc)::value;
                if (lm & (1u << q)) lds[S::template slot<q>()][c] = rv[q];
            });
        }
    }
}

}  // namespace lbm_amd
