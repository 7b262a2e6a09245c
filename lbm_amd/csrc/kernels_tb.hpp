// Two time steps per pass (temporal blocking) for domains whose every
// interior cell is BULK: periodic along all three axes, no boundary cells.
//
// A pass reads the field g0 once and writes g2 once, where two single steps
// read and write it twice each. Per cell the arithmetic is that of k_bulk
// and the intermediate g1 is rounded to the storage type, so a pass equals
// two single steps bit for bit.
//
// A block of 256 threads owns a TX x TY tile of cells in x-y and a chunk of
// z planes [z0, z1]. It marches through the planes zs = z0 - 1 .. z1 + 1:
//   (1) step 1 on plane zs for the tile and the ring of cells around it in
//       x-y (the (TX + 2) x (TY + 2) patch), pulling g0 from global memory;
//       every patch cell sends g1_q to the tile cell x + c_q of plane
//       zs + cz_q (the cell that pulls it in step 2) through LDS;
//   (2) step 2 on plane zs - 1, whose Q values have all arrived: collide and
//       store g2.
// A value sent with cz = -1 is read in the same iteration, cz = 0 in the next
// and cz = +1 in the one after. The cz >= 0 values are held in registers
// across (2), and the cz = +1 values one plane longer, and written after (2)
// of the iteration before they are read, so LDS holds one plane of slots per
// direction: Q x TX x TY values.
//
// With 16-bit storage the values travel in gather form: every patch cell
// writes its values into its own slots (patch index p = (pj + 1) (TX + 2) +
// pi + 1) with no test, and step 2 reads slot p(x - c_q) of direction q.
// That takes (TX + 2)(TY + 2) slots per direction instead of TX TY, and
// leaves out the per-direction bounds tests of the scatter form, which
// weigh on the 16-bit pass (bound by instruction issue, not by memory).
//
// Coordinates are kept unwrapped (x0 - 1 .. x0 + TX, zs = z0 - 1 .. z1 + 1)
// for the LDS bookkeeping and wrapped onto the periodic box only to address
// global memory, which keeps boxes of one or two cells along an axis exact.
#pragma once

#include "kernels.hpp"
#include "ops.hpp"

namespace lbm_amd {

constexpr int TB_PX = TB_TX + 2;
constexpr int TB_PY = TB_TY + 2;
constexpr int TB_NP = TB_PX * TB_PY;            // patch cells of one plane
constexpr int TB_PASS = (TB_NP + 255) / 256;    // step-1 cells per thread
constexpr int TB_CELLS = TB_TX * TB_TY;
constexpr int TB_OUT = TB_CELLS / 256;          // step-2 cells per thread
static_assert(TB_CELLS % 256 == 0, "step 2 gives every thread the same number of cells");

// LDS slot row of direction q (one plane per direction, see above).
template <class L>
struct TbSlots {
    static constexpr int count(int cz) {
        int n = 0;
        for (int q = 0; q < L::Q; ++q) n += (L::cz(q) == cz) ? 1 : 0;
        return n;
    }
    static constexpr int NM = count(-1), N0 = count(0), NP = count(1);
    static constexpr int NSLOT = NM + N0 + NP;
    static constexpr int rank(int q) {
        int r = 0;
        for (int k = 0; k < q; ++k) r += (L::cz(k) == L::cz(q)) ? 1 : 0;
        return r;
    }
    static constexpr int base(int q) {
        return L::cz(q) < 0 ? rank(q) : (L::cz(q) == 0 ? NM + rank(q) : NM + N0 + rank(q));
    }
    template <int q>
    static __device__ __forceinline__ constexpr int slot() {
        return base(q);
    }
};

// LDS rows of one direction: with 16-bit storage one slot per patch cell
// (the gather form below), else one per tile cell.
template <class L, Mode M>
struct TbCols {
    static constexpr bool gather = sizeof(typename Io<M, L>::T) == 2;
    static constexpr int n = gather ? TB_NP : TB_CELLS;
};

template <class L, Mode M>
constexpr size_t tb_lds_bytes() {
    return static_cast<size_t>(TbSlots<L>::NSLOT) * TbCols<L, M>::n * sizeof(typename Io<M, L>::T);
}

// base[byte_off]: a uniform base and a 32-bit byte offset, the form of a
// global access with a scalar base and one offset register per lane.
template <class T>
__device__ __forceinline__ const T& tb_at(const T* base, uint32_t byte_off) {
    return *reinterpret_cast<const T*>(reinterpret_cast<const char*>(base) + byte_off);
}
template <class T>
__device__ __forceinline__ T* tb_ptr(T* base, uint32_t byte_off) {
    return reinterpret_cast<T*>(reinterpret_cast<char*>(base) + byte_off);
}

template <class L, Mode M, bool BRINK>
__device__ __forceinline__ void tb_collide(const StepArgs& a, typename Io<M, L>::Calc (&f)[L::Q],
                                           typename Io<M, L>::Calc alpha) {
    using Calc = typename Io<M, L>::Calc;
    collide_shifted<L, Calc, BRINK, kRcp<M>, Io<M, L>::SX>(f, static_cast<Calc>(a.omega), BRINK ? alpha : Calc(0),
                                             static_cast<Calc>(a.tau * a.g[0]), static_cast<Calc>(a.tau * a.g[1]),
                                             static_cast<Calc>(a.tau * a.g[2]));
}

// Logical block of hardware block b when the hardware deals blocks to nxcd
// L2 domains in turn (b mod nxcd): each domain gets one contiguous range of
// logical blocks, so tiles that share halo lines share an L2. nxcd <= 1
// keeps the identity.
__device__ __forceinline__ int tb_logical_block(int b, int nb, int nxcd) {
    if (nxcd <= 1) return b;
    const int x = b % nxcd, slot = b / nxcd;
    const int per = nb / nxcd, rem = nb % nxcd;
    return x < rem ? x * (per + 1) + slot : rem * (per + 1) + (x - rem) * per + slot;
}

// Logical block = tx + ntx * (ty + nty * chunk); chunk c covers planes
// 1 + c * zc .. min((c + 1) * zc, zl).
template <class L, Mode M, bool BRINK>
__global__ void __launch_bounds__(256) k_bulk2(StepArgs a, int ntx, int nty, int zc, int nxcd) {
    using IO = Io<M, L>;
    using T = typename IO::T;
    using Calc = typename IO::Calc;
    using S = TbSlots<L>;
    constexpr int Q = L::Q;
    constexpr bool GATHER = TbCols<L, M>::gather;
    __shared__ T lds[S::NSLOT][TbCols<L, M>::n];

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
    const Calc* __restrict__ alpha = static_cast<const Calc*>(a.alpha);
    // byte offsets within a plane fit 32 bits (the host checks plane bytes)
    constexpr uint32_t TS = sizeof(T);
    const uint32_t py = static_cast<uint32_t>(Ly.px) * TS, pz = static_cast<uint32_t>(Ly.pxy) * TS;

    // this thread's patch cells: tile-local (pi, pj) in [-1, TX] x [-1, TY]
    int pi[TB_PASS], pj[TB_PASS], gx[TB_PASS], gy[TB_PASS];
    bool live[TB_PASS];
#pragma unroll
    for (int k = 0; k < TB_PASS; ++k) {
        const int p = k * 256 + tid;
        pi[k] = p % TB_PX - 1;
        pj[k] = p / TB_PX - 1;
        const int ux = x0 + pi[k], uy = y0 + pj[k];
        live[k] = p < TB_NP && ux <= Ly.xl + 1 && uy <= Ly.yl + 1;
        gx[k] = wrap1(ux, Ly.xl);
        gy[k] = wrap1(uy, Ly.yl);
    }

    // where patch cell k of this thread sends value q (scatter form: the
    // tile cell x + c_q, if in the tile; gather form: its own slot)
    auto send = [&](auto qc, int k, T v) {
        constexpr int q = decltype(qc)::value;
        if constexpr (GATHER) {
            lds[S::template slot<q>()][k * 256 + tid] = v;
        } else {
            const int di = pi[k] + L::cx(q), dj = pj[k] + L::cy(q);
            if (di >= 0 && di < TB_TX && dj >= 0 && dj < TB_TY) lds[S::template slot<q>()][dj * TB_TX + di] = v;
        }
    };

    T hold[TB_PASS][Q];
    // c_z = +1 values of the previous plane (only those are used); the first
    // iteration sends them for plane z0 - 1, which no step 2 reads
    T prev[TB_PASS][Q] = {};
    for (int zs = z0 - 1; zs <= z1 + 1; ++zs) {
        // The plane stride is redefined in each iteration (an empty asm that
        // emits nothing): the Q plane bases are then recomputed with a few
        // scalar instructions where they are used instead of being hoisted
        // out of the loop, where 2 Q of them overflow the scalar registers.
        int64_t plane = Ly.plane;
        asm volatile("" : "+s"(plane));
        // (1) step 1 on plane zs
        const int gz = wrap1(zs, Ly.zl);
        const uint32_t off = static_cast<uint32_t>(Ly.off) * TS;
        const uint32_t zs3[3] = {off + pz * wrap1(gz - 1, Ly.zl), off + pz * gz, off + pz * wrap1(gz + 1, Ly.zl)};
#pragma unroll
        for (int k = 0; k < TB_PASS; ++k) {
            if (!live[k]) continue;
            // 32-bit byte offsets within a plane (the plane base is uniform)
            const uint32_t xs[3] = {TS * wrap1(gx[k] - 1, Ly.xl), TS * gx[k], TS * wrap1(gx[k] + 1, Ly.xl)};
            const uint32_t ys[3] = {py * wrap1(gy[k] - 1, Ly.yl), py * gy[k], py * wrap1(gy[k] + 1, Ly.yl)};
            Calc f[Q];
            static_for<0, Q>([&](auto qc) {
                constexpr int q = decltype(qc)::value;
                // source x - c_q: index 0 for c = +1, 1 for 0, 2 for -1
                constexpr int ix = 1 - L::cx(q), iy = 1 - L::cy(q), iz = 1 - L::cz(q);
                const uint32_t o = zs3[iz] + ys[iy] + xs[ix];
                f[q] = IO::template load<q>(tb_at(src + q * plane, o));
            });
            Calc al = Calc(0);
            if constexpr (BRINK) al = alpha[(zs3[1] + ys[1] + xs[1]) / TS];
            tb_collide<L, M, BRINK>(a, f, al);
            static_for<0, Q>([&](auto qc) {
                constexpr int q = decltype(qc)::value;
                const T v = IO::template store<q>(f[q]);
                if constexpr (L::cz(q) < 0) {
                    send(qc, k, v);
                } else {
                    hold[k][q] = v;
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
                    Calc f[Q];
                    const int pc = (c / TB_TX + 1) * TB_PX + c % TB_TX + 1;  // c's patch index
                    static_for<0, Q>([&](auto qc) {
                        constexpr int q = decltype(qc)::value;
                        if constexpr (GATHER)
                            f[q] = IO::template load<q>(lds[S::template slot<q>()][pc - L::cx(q) - TB_PX * L::cy(q)]);
                        else
                            f[q] = IO::template load<q>(lds[S::template slot<q>()][c]);
                    });
                    const uint32_t eb = TS * (static_cast<uint32_t>(Ly.off) + ox) + pz * zo + py * oy;
                    Calc al = Calc(0);
                    if constexpr (BRINK) al = alpha[eb / TS];
                    tb_collide<L, M, BRINK>(a, f, al);
                    static_for<0, Q>([&](auto qc) {
                        constexpr int q = decltype(qc)::value;
                        store_stream1<T>(tb_ptr(dst + q * plane, eb), IO::template store<q>(f[q]));
                    });
                }
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
                if constexpr (L::cz(q) >= 0) send(qc, k, L::cz(q) == 0 ? hold[k][q] : prev[k][q]);
                if constexpr (L::cz(q) > 0) prev[k][q] = hold[k][q];
            });
        }
    }
}

}  // namespace lbm_amd
