// In-place time step (AA pattern): one distribution buffer M instead of two.
//
// The post-collision state is held in one of two representations:
//   E (natural): g_q(x) in M[q][x]
//   O (swapped): g_q(x) in M[qbar][x + c_q]   (already streamed, opposite slot)
// and steps alternate between them:
//   E -> O (k_aa_eo): h_q(x) = M[q][x - c_q]; collide; g'_k -> M[kbar][x + c_k].
//   O -> E (k_aa_oe): h_q(x) = M[qbar][x];     collide; g'_k -> M[k][x].
// A cell reads and writes the same set of locations in either step, and no
// two cells share a location, so the update is in place.
//
// Boundary links. The value a fluid cell x pulls over a link q whose source
// b = x - c_q is not fluid is the boundary rule evaluated on x's own
// post-collision values of the previous step (as in the two-buffer kernels).
// Each step therefore leaves the rule value in the slot the next step reads:
//   after E -> O, in M[k][x] for every k with x + c_k not fluid (the link
//     kbar of x), which O -> E reads as h_kbar(x) = M[k][x];
//   after O -> E, in M[q][b] for every link q of x with a non-fluid source b,
//     which E -> O reads as h_q(x) = M[q][x - c_q].
// E -> O also stores g'_k into the boundary cell's slot M[kbar][x + c_k], so
// the O representation is complete for every fluid cell; the first step after
// set-up reads the boundary cells' stored values, as the two-buffer kernels do.
//
// k_aa_swap turns O into E without a collision, for the readers (moments,
// sums, field export): it swaps each pair M[q][x] <-> M[qbar][x + c_q]. For a
// link to a boundary cell this also leaves the rule value stored in E -> O in
// the boundary cell's slot, where the next E -> O step reads it.
//
// FreeSlip reads other fluid cells' values and has no in-place form here; the
// host rejects it for this layout.
#pragma once

#include "kernels.hpp"

namespace lbm_amd {


// Element of the neighbour x + sgn c_q (wrapped along periodic axes).
template <class L, int q, class I>
__device__ __forceinline__ I aa_nb(const FieldLayout& Ly, bool wraps, I e, int x, int y, int z, int sgn) {
    if (!wraps) return e + static_cast<I>(sgn * (L::cx(q) + L::cy(q) * Ly.px + L::cz(q) * Ly.pxy));
    return static_cast<I>(Ly.off) + static_cast<I>(wrap_if(x + sgn * L::cx(q), Ly.xl, Ly.per[0])) +
           static_cast<I>(Ly.px) * static_cast<I>(wrap_if(y + sgn * L::cy(q), Ly.yl, Ly.per[1])) +
           static_cast<I>(Ly.pxy) * static_cast<I>(wrap_if(z + sgn * L::cz(q), Ly.zl, Ly.per[2]));
}

__device__ __forceinline__ bool aa_wraps(const FieldLayout& Ly, int x, int y, int z) {
    return (Ly.per[0] && (x == 1 || x == Ly.xl)) || (Ly.per[1] && (y == 1 || y == Ly.yl)) ||
           (Ly.per[2] && (z == 1 || z == Ly.zl));
}

// Cell (x, y, z) of this lane in the k_cells tiling; false past the domain.
__device__ __forceinline__ bool aa_tile_cell(const FieldLayout& Ly, int tpp, double rxl, int& x, int& y, int& z) {
    const int64_t w = static_cast<int64_t>(blockIdx.x) * (blockDim.x >> 6) + (threadIdx.x >> 6);
    z = 1 + static_cast<int>(w / tpp);
    if (z > Ly.zl) return false;
    const int64_t c = (w % tpp) * 64 + (threadIdx.x & 63);
    const int64_t xl = Ly.xl;
    if (c >= xl * Ly.yl) return false;
    int64_t yy = static_cast<int64_t>(static_cast<double>(c) * rxl);
    int64_t xx = c - yy * xl;
    if (xx < 0) {
        --yy;
        xx += xl;
    } else if (xx >= xl) {
        ++yy;
        xx -= xl;
    }
    x = 1 + static_cast<int>(xx);
    y = 1 + static_cast<int>(yy);
    return true;
}

template <class L, Mode M>
__device__ __forceinline__ void aa_collide(const StepArgs& a, int64_t e, typename Io<M, L>::Calc (&f)[L::Q]) {
    using Calc = typename Io<M, L>::Calc;
    if constexpr (M == Mode::EXACT64) {
        collide_exact<L>(f, a.tau);
    } else if (a.alpha) {
        const Calc alpha = static_cast<const Calc*>(a.alpha)[e];
        collide_shifted<L, Calc, true, kRcp<M>, Io<M, L>::SX>(f, static_cast<Calc>(a.omega), alpha,
                                                              static_cast<Calc>(a.tau * a.g[0]),
                                                static_cast<Calc>(a.tau * a.g[1]), static_cast<Calc>(a.tau * a.g[2]));
    } else {
        collide_shifted<L, Calc, false, kRcp<M>, Io<M, L>::SX>(f, static_cast<Calc>(a.omega), Calc(0), Calc(0), Calc(0),
                                                               Calc(0));
    }
}

// x's post-collision values (f: the collision's scaled representation) as
// they would be stored (rounded to the storage format), which is what the
// two-buffer kernels' rules read back.
template <class L, Mode M>
__device__ __forceinline__ void aa_own(const typename Io<M, L>::Calc (&f)[L::Q], typename Io<M, L>::Calc (&own)[L::Q]) {
    using IO = Io<M, L>;
    static_for<0, L::Q>([&](auto qc) {
        constexpr int q = decltype(qc)::value;
        own[q] = IO::template to_calc<q>(IO::template store<q>(f[q]));
    });
}

template <class L, Mode M, bool B32>
__global__ void __launch_bounds__(256) k_aa_eo(StepArgs a, int tpp, double rxl) {
    using IO = Io<M, L>;
    using T = typename IO::T;
    using Calc = typename IO::Calc;
    constexpr int Q = L::Q;
    const FieldLayout& Ly = a.L;
    int x, y, z;
    if (!aa_tile_cell(Ly, tpp, rxl, x, y, z)) return;
    using I = typename ElemIdx<B32>::type;
    const I e = static_cast<I>(cell_elem(Ly, x, y, z));
    const uint8_t k = kind_at(a.kind, e);
    if (k >= KIND_COND0) return;
    T* __restrict__ m = static_cast<T*>(a.dst);
    const bool wraps = aa_wraps(Ly, x, y, z);
    Calc f[Q];
    static_for<0, Q>([&](auto qc) {
        constexpr int q = decltype(qc)::value;
        f[q] = IO::template load<q>(plane_ref<B32>(m, q, Ly.plane, aa_nb<L, q>(Ly, wraps, e, x, y, z, -1)));
    });
    aa_collide<L, M>(a, static_cast<int64_t>(e), f);
    static_for<0, Q>([&](auto qc) {
        constexpr int q = decltype(qc)::value;
        plane_ref<B32>(m, Q - 1 - q, Ly.plane, aa_nb<L, q>(Ly, wraps, e, x, y, z, 1)) = IO::template store<q>(f[q]);
    });
    if (k == KIND_NEAR) {
        Calc own[Q];
        aa_own<L, M>(f, own);
        const OwnSums<Calc> os = own_sums<L, Calc>(own);
        static_for<0, Q>([&](auto qc) {
            constexpr int q = decltype(qc)::value;
            constexpr int j = Q - 1 - q;  // the link of x whose source is x + c_q
            const uint8_t kb = kind_at(a.kind, aa_nb<L, q>(Ly, wraps, e, x, y, z, 1));
            if (kb >= KIND_COND0 && kb < KIND_INERT)
                plane_ref<B32>(m, q, Ly.plane, e) =
                    IO::template from_calc<j>(rule_value<L, M, j>(cond_at(a.conds, kb - KIND_COND0), os, own[q]));
        });
    }
}

template <class L, Mode M, bool B32>
__global__ void __launch_bounds__(256) k_aa_oe(StepArgs a, int tpp, double rxl) {
    using IO = Io<M, L>;
    using T = typename IO::T;
    using Calc = typename IO::Calc;
    constexpr int Q = L::Q;
    const FieldLayout& Ly = a.L;
    int x, y, z;
    if (!aa_tile_cell(Ly, tpp, rxl, x, y, z)) return;
    using I = typename ElemIdx<B32>::type;
    const I e = static_cast<I>(cell_elem(Ly, x, y, z));
    const uint8_t k = kind_at(a.kind, e);
    if (k >= KIND_COND0) return;
    T* __restrict__ m = static_cast<T*>(a.dst);
    Calc f[Q];
    static_for<0, Q>([&](auto qc) {
        constexpr int q = decltype(qc)::value;
        f[q] = IO::template load<q>(plane_ref<B32>(m, Q - 1 - q, Ly.plane, e));
    });
    aa_collide<L, M>(a, static_cast<int64_t>(e), f);
    static_for<0, Q>([&](auto qc) {
        constexpr int q = decltype(qc)::value;
        plane_ref<B32>(m, q, Ly.plane, e) = IO::template store<q>(f[q]);
    });
    if (k == KIND_NEAR) {
        const bool wraps = aa_wraps(Ly, x, y, z);
        Calc own[Q];
        aa_own<L, M>(f, own);
        const OwnSums<Calc> os = own_sums<L, Calc>(own);
        static_for<0, Q>([&](auto qc) {
            constexpr int q = decltype(qc)::value;
            const I be = aa_nb<L, q>(Ly, wraps, e, x, y, z, -1);
            const uint8_t kb = kind_at(a.kind, be);
            if (kb >= KIND_COND0 && kb < KIND_INERT)
                plane_ref<B32>(m, q, Ly.plane, be) =
                    IO::template from_calc<q>(rule_value<L, M, q>(cond_at(a.conds, kb - KIND_COND0), os, own[Q - 1 - q]));
        });
    }
}

// O -> E without a collision: swap M[q][x] <-> M[qbar][x + c_q]. A fluid cell
// takes the pairs of its first Q/2 directions, and those of the others whose
// partner cell is not fluid (that partner takes no pairs).
template <class L, Mode M, bool B32>
__global__ void __launch_bounds__(256) k_aa_swap(StepArgs a, int tpp, double rxl) {
    using T = typename Io<M, L>::T;
    constexpr int Q = L::Q;
    const FieldLayout& Ly = a.L;
    int x, y, z;
    if (!aa_tile_cell(Ly, tpp, rxl, x, y, z)) return;
    using I = typename ElemIdx<B32>::type;
    const I e = static_cast<I>(cell_elem(Ly, x, y, z));
    if (kind_at(a.kind, e) >= KIND_COND0) return;
    T* __restrict__ m = static_cast<T*>(a.dst);
    const bool wraps = aa_wraps(Ly, x, y, z);
    static_for<0, Q>([&](auto qc) {
        constexpr int q = decltype(qc)::value;
        if constexpr (q != L::rest) {
            const I pe = aa_nb<L, q>(Ly, wraps, e, x, y, z, 1);
            if (q < Q / 2 || kind_at(a.kind, pe) >= KIND_COND0) {
                Glob<T>* p0 = &plane_ref<B32>(m, q, Ly.plane, e);
                Glob<T>* p1 = &plane_ref<B32>(m, Q - 1 - q, Ly.plane, pe);
                const T v0 = *p0;
                *p0 = *p1;
                *p1 = v0;
            }
        }
    });
}

}  // namespace lbm_amd
