// Fused pull-stream + collide kernels and the per-field kernels of one mode.
//
// One time step of the reference (src/main.cpp:50-52: stream, swap, collide)
// is one pass here: every interior fluid cell x pulls f_q from x - c_q in the
// source buffer (post-collision values of the previous step), collides in
// registers and writes its Q values to the destination buffer.
//
// The reference's second collide pass (boundary rules on non-fluid cells,
// include/domain.hpp:157-165) writes into a non-fluid cell b exactly those
// directions q for which b + c_q is an interior fluid cell, and the only
// reader of that value is the pull of that fluid cell in the next step. So a
// fluid cell x that pulls direction q from a non-fluid b = x - c_q evaluates
// b's rule for (b, q) itself, from its own post-collision values (all rules
// read only the neighbour b + c_q = x, except FreeSlip, which reads fluid
// cells next to b). Right after initialisation or after the host wrote the
// field, non-fluid cells hold values that the next pull reads as stored
// (the first step of the reference pulls the initial weights from every
// boundary cell); the `materialised` flag selects that.
//
// Two kernels per step:
//   k_bulk - every interior cell of kind BULK (all Q sources fluid), V cells
//            per thread along x, one plane-contiguous load per direction.
//   k_near - the listed cells of kind NEAR (fluid with a non-fluid source),
//            one thread per cell, evaluates boundary rules.
#pragma once

#include <type_traits>

#include "collide.hpp"
#include "domain_view.hpp"
#include "lbm_hip/lattice.hpp"
#include "storage.hpp"

namespace lbm_amd {

template <class T, int V>
struct alignas(sizeof(T) * V) VecT {
    T v[V];
};

// Streaming (non-temporal) stores for the bulk output: every value written
// in a step is read once in the next step, long after it would have been
// evicted, so keeping it in L2 only displaces source lines.
#ifndef LBM_NT_STORE
#define LBM_NT_STORE 1
#endif

template <int BYTES>
struct RawOf;
template <>
struct RawOf<1> { using type = uint8_t; };
template <>
struct RawOf<2> { using type = uint16_t; };
template <>
struct RawOf<4> { using type = uint32_t; };
template <>
struct RawOf<8> { using type = uint32_t __attribute__((ext_vector_type(2))); };
template <>
struct RawOf<16> { using type = uint32_t __attribute__((ext_vector_type(4))); };
template <>
struct RawOf<32> { using type = uint32_t __attribute__((ext_vector_type(8))); };

template <class T, int V>
__device__ __forceinline__ void store_stream(T* p, const VecT<T, V>& v) {
#if LBM_NT_STORE
    using R = typename RawOf<sizeof(T) * V>::type;
    R r;
    __builtin_memcpy(&r, &v, sizeof(R));
    __builtin_nontemporal_store(r, reinterpret_cast<R*>(p));
#else
    *reinterpret_cast<VecT<T, V>*>(p) = v;
#endif
}

template <class T>
__device__ __forceinline__ void store_stream1(T* p, T v) {
#if LBM_NT_STORE
    using R = typename RawOf<sizeof(T)>::type;
    R r;
    __builtin_memcpy(&r, &v, sizeof(R));
    __builtin_nontemporal_store(r, reinterpret_cast<R*>(p));
#else
    *p = v;
#endif
}

template <class T>
__device__ __forceinline__ void store_stream1(__attribute__((address_space(1))) T* p, T v) {
#if LBM_NT_STORE
    using R = typename RawOf<sizeof(T)>::type;
    R r;
    __builtin_memcpy(&r, &v, sizeof(R));
    __builtin_nontemporal_store(r, reinterpret_cast<__attribute__((address_space(1))) R*>(p));
#else
    *p = v;
#endif
}

// Element indices within a plane: 32-bit where a plane holds less than
// 4 GiB (B32), so that an access is a uniform direction-plane base plus one
// 32-bit byte offset per lane; 64-bit otherwise.
template <bool B32>
struct ElemIdx {
    using type = int64_t;
};
template <>
struct ElemIdx<true> {
    using type = uint32_t;
};

// Objects in the global address space. Every field, kind array and
// condition table lives in device (global) memory, but where a kernel reads
// their pointers from a table in memory (the batched kernels) the compiler
// cannot prove it and addresses them with flat instructions, which take no
// scalar base and wait on the LDS counter too. Accesses through Glob<T>
// references are global loads and stores in every kernel.
template <class T>
using Glob = __attribute__((address_space(1))) T;

template <class T>
__device__ __forceinline__ Glob<T>* glob(T* p) {
    return (Glob<T>*)p;
}

// Dual-number fields (the fp64 tangent run) keep generic references: a
// class with user-declared constructors is not copied out of an
// address-space-qualified object.
template <class T>
struct IsDual : std::false_type {};
template <class U>
struct IsDual<Dual<U>> : std::true_type {};
template <class T>
using PlaneElem = std::conditional_t<IsDual<std::remove_const_t<T>>::value, T, Glob<T>>;

// Element e of direction plane q of field m (global memory).
template <bool B32, class T, class I>
__device__ __forceinline__ PlaneElem<T>& plane_ref(T* m, int q, int64_t plane, I e) {
    using Byte = std::conditional_t<std::is_const_v<T>, const char, char>;
    using E = PlaneElem<T>;
    using EB = std::conditional_t<IsDual<std::remove_const_t<T>>::value, Byte, Glob<Byte>>;
    E* base = (E*)(m + q * plane);
    if constexpr (B32)
        return *reinterpret_cast<E*>(reinterpret_cast<EB*>(base) +
                                     static_cast<uint32_t>(e) * static_cast<uint32_t>(sizeof(T)));
    else
        return base[e];
}

// Cell kind at element e, and condition i of a domain's table.
template <class I>
__device__ __forceinline__ uint8_t kind_at(const uint8_t* kind, I e) {
    return glob(kind)[e];
}
using GCond = Glob<const Cond>;
__device__ __forceinline__ const GCond& cond_at(const Cond* conds, int i) { return glob(conds)[i]; }

// cell_elem in the index type of plane_ref. With B32 every element index of
// the field is below 2^32 and all terms are non-negative, so 32-bit
// arithmetic gives it exactly.
template <bool B32>
__device__ __forceinline__ typename ElemIdx<B32>::type elem_at(const FieldLayout& L, int x, int y, int z) {
    if constexpr (B32)
        return static_cast<uint32_t>(L.off) + static_cast<uint32_t>(x) +
               static_cast<uint32_t>(L.px) * static_cast<uint32_t>(y) +
               static_cast<uint32_t>(L.pxy) * static_cast<uint32_t>(z);
    else
        return cell_elem(L, x, y, z);
}

__device__ __forceinline__ int wrap1(int v, int l) { return v == 0 ? l : (v == l + 1 ? 1 : v); }

__device__ __forceinline__ int wrap_if(int v, int l, int per) {
    return per ? wrap1(v, l) : v;
}

// ---------------------------------------------------------------------------
// Bulk kernel. Block (BX, BY) threads covers BX*V cells of BY rows at one z.
// Grid is flattened: blockIdx.x = bx + nbx * (by + nby * (z - 1)).
//
// Source rows are wrapped unconditionally: along a periodic axis that is the
// periodic image, and along a walled axis any cell next to the wall pulls
// from a ghost cell, is NEAR and is masked out before its values are used.
// V cells (x0 .. x0+V-1, y, z) of one thread. Returns false when a cell it
// updated read a NaN or infinite value (finite_calc of the collision's base).
template <class L, Mode M, int V, bool FLAGS, bool BRINK>
__device__ __forceinline__ bool bulk_cells(const StepArgs& a, int x0, int y, int z,
                                           bool* near_lanes = nullptr) {
    using IO = Io<M, L>;
    using T = typename IO::T;
    using Calc = typename IO::Calc;
    constexpr int Q = L::Q;
    const FieldLayout& Ly = a.L;
    if (x0 > Ly.xl || y > Ly.yl) return true;

    const int64_t e0 = cell_elem(Ly, x0, y, z);
    bool act[V];
    bool any = false;
    if constexpr (FLAGS) {
        const VecT<uint8_t, V> kv = *reinterpret_cast<const VecT<uint8_t, V>*>(a.kind + e0);
#pragma unroll
        for (int i = 0; i < V; ++i) {
            act[i] = (kv.v[i] == KIND_BULK) && (x0 + i <= Ly.xl);
            any |= act[i];
            if (near_lanes) near_lanes[i] = (kv.v[i] == KIND_NEAR) && (x0 + i <= Ly.xl);
        }
    } else {
#pragma unroll
        for (int i = 0; i < V; ++i) {
            act[i] = (x0 + i <= Ly.xl);
            any |= act[i];
        }
    }
    const int ym = wrap1(y - 1, Ly.yl), yp = wrap1(y + 1, Ly.yl);
    const int zm = wrap1(z - 1, Ly.zl), zp = wrap1(z + 1, Ly.zl);
    const int xm = wrap1(x0 - 1, Ly.xl);  // source of lane 0 for cx = +1
    const int xpv = wrap1(x0 + V, Ly.xl); // source of lane V-1 for cx = -1
    const bool partial = (x0 + V - 1 > Ly.xl);

    const T* __restrict__ src = static_cast<const T*>(a.src);
    T* __restrict__ dst = static_cast<T*>(a.dst);

    Calc f[V][Q];
    static_for<0, Q>([&](auto qc) {
        constexpr int q = decltype(qc)::value;
        constexpr int cx = L::cx(q), cy = L::cy(q), cz = L::cz(q);
        const int sy = cy > 0 ? ym : (cy < 0 ? yp : y);
        const int sz = cz > 0 ? zm : (cz < 0 ? zp : z);
        const T* __restrict__ row = src + q * Ly.plane + Ly.off + Ly.px * sy + Ly.pxy * sz;
        T v[V];
        if constexpr (V == 1) {
            v[0] = row[cx == 0 ? x0 : (cx > 0 ? xm : xpv)];
        } else {
            const VecT<T, V> A = *reinterpret_cast<const VecT<T, V>*>(row + x0);
            if constexpr (cx == 0) {
#pragma unroll
                for (int i = 0; i < V; ++i) v[i] = A.v[i];
            } else if constexpr (cx > 0) {
                v[0] = row[xm];
#pragma unroll
                for (int i = 1; i < V; ++i) v[i] = A.v[i - 1];
            } else {
#pragma unroll
                for (int i = 0; i < V - 1; ++i) v[i] = A.v[i + 1];
                v[V - 1] = row[xpv];
                // last, partially filled vector of a periodic row: lane of x = xl
                // takes x = 1 (on a walled row that lane is NEAR and masked)
                if (partial) {
#pragma unroll
                    for (int i = 0; i < V - 1; ++i)
                        if (x0 + i == Ly.xl) v[i] = row[1];
                }
            }
        }
#pragma unroll
        for (int i = 0; i < V; ++i) f[i][q] = IO::template load<q>(v[i]);
    });

    // the loads above are issued before this test (every row they touch lies
    // inside the ghosted box), so the kind lookup costs no extra round trip
    if (!any) return true;

    Calc alpha[V];
    if constexpr (BRINK) {
        const Calc* al = static_cast<const Calc*>(a.alpha);
#pragma unroll
        for (int i = 0; i < V; ++i) alpha[i] = act[i] ? al[e0 + i] : Calc(0);
    }

    bool finite = true;
#pragma unroll
    for (int i = 0; i < V; ++i) {
        if constexpr (M == Mode::EXACT64) {
            collide_exact<L>(f[i], a.tau);
        } else {
            const Calc base = collide_shifted<L, Calc, BRINK, kRcp<M>, IO::SX>(f[i], static_cast<Calc>(a.omega),
                                            BRINK ? alpha[i] : Calc(0),
                                            static_cast<Calc>(a.tau * a.g[0]),
                                            static_cast<Calc>(a.tau * a.g[1]),
                                            static_cast<Calc>(a.tau * a.g[2]));
            finite &= !act[i] || finite_calc(base);
        }
    }

    bool all = true;
#pragma unroll
    for (int i = 0; i < V; ++i) all &= act[i];
    static_for<0, Q>([&](auto qc) {
        constexpr int q = decltype(qc)::value;
        T* __restrict__ out = dst + q * Ly.plane + e0;
        if (all) {
            VecT<T, V> o;
#pragma unroll
            for (int i = 0; i < V; ++i) o.v[i] = IO::template store<q>(f[i][q]);
            store_stream<T, V>(out, o);
        } else {
#pragma unroll
            for (int i = 0; i < V; ++i)
                if (act[i]) store_stream1<T>(out + i, IO::template store<q>(f[i][q]));
        }
    });
    return finite;
}

template <class L, Mode M, int V, bool FLAGS, bool BRINK>
__global__ void __launch_bounds__(256) k_bulk(StepArgs a, int nbx, int nby) {
    const int b = blockIdx.x;
    const int bx = b % nbx;
    const int t = b / nbx;
    const int by = t % nby;
    const int z = 1 + t / nby;
    const int xv = bx * blockDim.x + threadIdx.x;
    const int y = 1 + by * blockDim.y + threadIdx.y;
    bulk_cells<L, M, V, FLAGS, BRINK>(a, 1 + xv * V, y, z);
}

// ---------------------------------------------------------------------------
// Boundary rules, evaluated by the fluid cell x for a non-fluid source b.
// own[] holds x's post-collision values (the values the reference's rule
// reads from its fluid neighbour), in the Calc representation of the mode.

// The rules read x's own values only through their plain sum, their three
// momentum sums and the single value own[qbar]. OwnSums holds the sums, each
// accumulated from 0 over ascending q: the order of the reference's density()
// and velocity() (include/collision.hpp:7-31), which EXACT64 keeps bit for
// bit (adding -v equals subtracting v; directions with c_d = 0 add nothing).
template <class Calc>
struct OwnSums {
    Calc s, m0, m1, m2;
};

template <class L, class Calc>
struct OwnAcc {
    OwnSums<Calc> r{Calc(0), Calc(0), Calc(0), Calc(0)};
    template <int q>
    __device__ __forceinline__ void add(Calc v) {
        r.s += v;
        if constexpr (L::cx(q) > 0) r.m0 += v;
        if constexpr (L::cx(q) < 0) r.m0 -= v;
        if constexpr (L::cy(q) > 0) r.m1 += v;
        if constexpr (L::cy(q) < 0) r.m1 -= v;
        if constexpr (L::cz(q) > 0) r.m2 += v;
        if constexpr (L::cz(q) < 0) r.m2 -= v;
    }
};

template <class L, class Calc>
__device__ __forceinline__ OwnSums<Calc> own_sums(const Calc (&own)[L::Q]) {
    OwnAcc<L, Calc> acc;
    static_for<0, L::Q>([&](auto qc) {
        constexpr int q = decltype(qc)::value;
        acc.template add<q>(own[q]);
    });
    return acc.r;
}

// Own values read from the source field at element e (one pass, nothing kept).
template <class L, Mode M, bool B32 = false>
__device__ __forceinline__ OwnSums<typename Io<M, L>::Calc> own_sums_at(
    const typename Io<M, L>::T* __restrict__ src, int64_t plane, typename ElemIdx<B32>::type e) {
    using IO = Io<M, L>;
    OwnAcc<L, typename IO::Calc> acc;
    static_for<0, L::Q>([&](auto qc) {
        constexpr int q = decltype(qc)::value;
        acc.template add<q>(IO::template to_calc<q>(plane_ref<B32>(src, q, plane, e)));
    });
    return acc.r;
}

// c ? a : b by bit masks. A plain select between two array elements can be
// turned into a select between their addresses, which moves both arrays
// out of registers into scratch memory.
template <class T>
__device__ __forceinline__ T pick_bits(bool c, T a, T b) {
    using U = typename std::conditional<sizeof(T) == 8, uint64_t, uint32_t>::type;
    static_assert(sizeof(T) == sizeof(U), "float or double");
    const U m = U(0) - static_cast<U>(c);
    return __builtin_bit_cast(T, (__builtin_bit_cast(U, a) & m) | (__builtin_bit_cast(U, b) & ~m));
}

// The MovingWall and Outflow/Pressure values of the fp32/fp64 storage modes,
// shared by rule_value and rule_value_flat so that both round alike.
// MovingWall (include/boundary.hpp:55-65) with c_q . u_wall tabulated per
// condition.
template <class L, class Calc, int q>
__device__ __forceinline__ Calc moving_wall_rule(const OwnSums<Calc>& os, Calc finv, const GCond& cd) {
    const Calc rho = Calc(1) + os.s;
    Calc cu;
    if constexpr (std::is_same<Calc, float>::value)
        cu = cd.cu32[q];
    else
        cu = static_cast<Calc>(cd.cu64[q]);
    return finv + static_cast<Calc>(6.0 * weight<L>(q)) * rho * cu;
}

// Outflow / Pressure (include/boundary.hpp:140-147, 206-211):
// (feq_q - w) + (feq_qbar - w) - finv at the condition's density and x's
// velocity; the odd part of the two equilibria cancels.
template <class L, class Calc, int q>
__device__ __forceinline__ Calc outflow_rule(const OwnSums<Calc>& os, Calc finv, const GCond& cd) {
    const Calc rr = static_cast<Calc>(cd.rho);
    const Calc ir = Calc(1) / rr;
    const Calc ux = os.m0 * ir, uy = os.m1 * ir, uz = os.m2 * ir;
    const Calc cu = Calc(L::cx(q)) * ux + Calc(L::cy(q)) * uy + Calc(L::cz(q)) * uz;
    const Calc uu = ux * ux + uy * uy + uz * uz;
    const Calc even = (rr - Calc(1)) + rr * (Calc(4.5) * cu * cu - Calc(1.5) * uu);
    return static_cast<Calc>(2.0 * weight<L>(q)) * even - finv;
}

// Value of a NoSlip, MovingWall, Inflow, Outflow or Pressure rule for
// direction q, from x's own sums and own[qbar] (finv).
template <class L, Mode M, int q, int RULES = RULES_ALL>
__device__ __forceinline__ typename Io<M, L>::Calc rule_value(const GCond& cd,
                                                             const OwnSums<typename Io<M, L>::Calc>& os,
                                                             typename Io<M, L>::Calc finv) {
    using Calc = typename Io<M, L>::Calc;
    constexpr int Q = L::Q;
    const int kind = cd.kind;
    if constexpr ((RULES & RULE_MOVINGWALL) != 0) {
        if (kind == BC_MOVINGWALL) {  // boundary.hpp:55-65
            if constexpr (M == Mode::EXACT64) {
                double c_dot_u = 0;
                c_dot_u += static_cast<double>(L::cx(q)) * cd.vel[0];
                c_dot_u += static_cast<double>(L::cy(q)) * cd.vel[1];
                c_dot_u += static_cast<double>(L::cz(q)) * cd.vel[2];
                constexpr double cs = kCS;
                return finv + 2.0 * weight<L>(q) * os.s * c_dot_u / (cs * cs);
            } else {
                return moving_wall_rule<L, Calc, q>(os, finv, cd);
            }
        }
    }
    if constexpr ((RULES & RULE_INFLOW) != 0) {
        if (kind == BC_INFLOW)  // boundary.hpp:176-178, feq precomputed on the host
            return (M == Mode::EXACT64) ? static_cast<Calc>(cd.feq[q]) : static_cast<Calc>(cd.geq[q]);
    }
    if constexpr ((RULES & RULE_OUTFLOW) != 0) {
        if (kind == BC_OUTFLOW || kind == BC_PRESSURE) {  // boundary.hpp:140-147, 206-211
            if constexpr (M == Mode::EXACT64) {
                const double u[3] = {os.m0 / cd.rho, os.m1 / cd.rho, os.m2 / cd.rho};
                const double fq = feq_exact<L, q>(cd.rho, u);
                const double fb = feq_exact<L, Q - 1 - q>(cd.rho, u);
                return fq + fb - finv;
            } else {
                return outflow_rule<L, Calc, q>(os, finv, cd);
            }
        }
    }
    return finv;  // BC_NOSLIP, boundary.hpp:26-28
}

// rule_value without branches (fp32/fp64 storage modes): every family built
// in is evaluated and the condition's kind selects one. cd must be a valid
// condition even where the result is not used, so a caller can issue the
// loads of all its links at once instead of one link after another.
template <class L, Mode M, int q, int RULES = RULES_ALL>
__device__ __forceinline__ typename Io<M, L>::Calc rule_value_flat(const GCond& cd,
                                                                  const OwnSums<typename Io<M, L>::Calc>& os,
                                                                  typename Io<M, L>::Calc finv) {
    static_assert(M != Mode::EXACT64, "the reference-order mode keeps rule_value");
    using Calc = typename Io<M, L>::Calc;
    const int kind = cd.kind;
    Calc v = finv;  // BC_NOSLIP
    if constexpr ((RULES & RULE_MOVINGWALL) != 0) {
        const Calc w = moving_wall_rule<L, Calc, q>(os, finv, cd);
        v = kind == BC_MOVINGWALL ? w : v;
    }
    if constexpr ((RULES & RULE_INFLOW) != 0) {
        const Calc g = static_cast<Calc>(cd.geq[q]);
        v = kind == BC_INFLOW ? g : v;
    }
    if constexpr ((RULES & RULE_OUTFLOW) != 0) {
        const Calc o = outflow_rule<L, Calc, q>(os, finv, cd);
        v = (kind == BC_OUTFLOW || kind == BC_PRESSURE) ? o : v;
    }
    return v;
}

// FreeSlip candidates (include/boundary.hpp:93-110), in order: which of the
// boundary cell's coordinates (x, y, z) are replaced by the fluid cell's, i.e.
// which velocity components the candidate's population reflects.
constexpr int kFsSwap[6][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}, {0, 1, 1}, {1, 0, 1}, {1, 1, 0}};

// Direction index of velocity (cx, cy, cz) (include/model.h velocity_index),
// evaluated at compile time.
template <class L>
constexpr int dir_of(int cx, int cy, int cz) {
    for (int q = 0; q < L::Q; ++q)
        if (L::cx(q) == cx && L::cy(q) == cy && L::cz(q) == cz) return q;
    return -1;
}


// Value that fluid cell x = (x,y,z) pulls in direction q from the non-fluid
// cell b = (sx,sy,sz) (element se) with condition cd: the value the reference's
// rule writes into b[q] in collide pass 2 (include/boundary.hpp). os holds
// the sums of x's own post-collision values and own(k) returns own value k.
// Sets stale when the rule writes nothing (FreeSlip with no fluid candidate);
// the value returned is then b's stored value, which is what the reference
// reads too. FS = false builds the rule without the FreeSlip case, for
// domains that have no FreeSlip condition. B32 indexes with 32-bit offsets
// (planes under 4 GiB).
template <class L, Mode M, int q, bool FS = true, bool B32 = false, class OwnFn>
__device__ __forceinline__ typename Io<M, L>::Calc pull_boundary(
    const StepArgs& a, const GCond& cd, const OwnSums<typename Io<M, L>::Calc>& os, OwnFn own, int x,
    int y, int z, int sx, int sy, int sz, typename ElemIdx<B32>::type se, bool& stale) {
    using IO = Io<M, L>;
    using T = typename IO::T;
    constexpr int Q = L::Q;
    constexpr int cx = L::cx(q), cy = L::cy(q), cz = L::cz(q);
    const FieldLayout& Ly = a.L;
    const T* __restrict__ src = static_cast<const T*>(a.src);
    stale = false;
    if constexpr (FS) {
        if (cd.kind == BC_FREESLIP) {  // boundary.hpp:91-110 with b = (sx,sy,sz), b + c_q = x
            // Candidates in the reference's order; candidate j replaces the
            // coordinates flagged in kFsSwap[j] by x's and reflects those
            // velocity components. The reflected directions are compile-time.
            typename IO::Calc v{};
            bool found = false;
            static_for<0, 6>([&](auto jc) {
                constexpr int j = decltype(jc)::value;
                constexpr int rq = dir_of<L>(kFsSwap[j][0] ? -cx : cx, kFsSwap[j][1] ? -cy : cy,
                                             kFsSwap[j][2] ? -cz : cz);
                static_assert(rq >= 0, "reflected velocity outside the lattice");
                if (!found) {
                    const auto ce = elem_at<B32>(Ly, kFsSwap[j][0] ? x : sx, kFsSwap[j][1] ? y : sy,
                                                 kFsSwap[j][2] ? z : sz);
                    if (kind_at(a.kind, ce) < KIND_COND0) {
                        v = IO::template to_calc<rq>(plane_ref<B32>(src, rq, Ly.plane, ce));
                        found = true;
                    }
                }
            });
            if (found) return v;
            stale = true;
            return IO::template to_calc<q>(plane_ref<B32>(src, q, Ly.plane, se));
        }
    }
    if (cd.kind < BC_NOSLIP || cd.kind > BC_PRESSURE) {
        stale = true;
        return IO::template to_calc<q>(plane_ref<B32>(src, q, Ly.plane, se));
    }
    return rule_value<L, M, q>(cd, os, own(std::integral_constant<int, Q - 1 - q>{}));
}

// The same with x's own values already in registers.
template <class L, Mode M, int q, bool FS = true>
__device__ __forceinline__ typename Io<M, L>::Calc pull_nonfluid(
    const StepArgs& a, const GCond& cd, const typename Io<M, L>::Calc (&own)[L::Q], int x, int y,
    int z, int sx, int sy, int sz, int64_t se, bool& stale) {
    using Calc = typename Io<M, L>::Calc;
    const OwnSums<Calc> os = own_sums<L, Calc>(own);
    return pull_boundary<L, M, q, FS>(a, cd, os, [&](auto kc) { return own[decltype(kc)::value]; }, x, y, z,
                                  sx, sy, sz, se, stale);
}

// ---------------------------------------------------------------------------
// One NEAR cell (fluid with a non-fluid source). Returns false when the cell
// read a NaN or infinite value.
template <class L, Mode M>
__device__ __forceinline__ bool near_cell(const StepArgs& a, int x, int y, int z) {
    using IO = Io<M, L>;
    using T = typename IO::T;
    using Calc = typename IO::Calc;
    constexpr int Q = L::Q;
    const FieldLayout& Ly = a.L;
    const int64_t e = cell_elem(Ly, x, y, z);
    const T* __restrict__ src = static_cast<const T*>(a.src);
    T* __restrict__ dst = static_cast<T*>(a.dst);

    // x's own post-collision values (the rules' inputs): sums now, single
    // values reloaded when a rule needs one
    const OwnSums<Calc> os = own_sums_at<L, M>(src, Ly.plane, e);
    auto own = [&](auto kc) {
        constexpr int k = decltype(kc)::value;
        return IO::template to_calc<k>(src[k * Ly.plane + e]);
    };

    Calc f[Q];
    static_for<0, Q>([&](auto qc) {
        constexpr int q = decltype(qc)::value;
        constexpr int cx = L::cx(q), cy = L::cy(q), cz = L::cz(q);
        const int sx = wrap_if(x - cx, Ly.xl, Ly.per[0]);
        const int sy = wrap_if(y - cy, Ly.yl, Ly.per[1]);
        const int sz = wrap_if(z - cz, Ly.zl, Ly.per[2]);
        const int64_t se = cell_elem(Ly, sx, sy, sz);
        const uint8_t k = kind_at(a.kind, se);
        if (k < KIND_COND0 || a.materialised) {
            f[q] = IO::template load<q>(src[q * Ly.plane + se]);
        } else {
            bool stale;
            f[q] = IO::template through_storage<q>(pull_boundary<L, M, q>(a, cond_at(a.conds, k - KIND_COND0), os, own,
                                                                          x, y, z, sx, sy, sz, se, stale));
        }
    });

    bool finite = true;
    if constexpr (M == Mode::EXACT64) {
        collide_exact<L>(f, a.tau);
    } else {
        Calc alpha = Calc(0);
        if (a.alpha) alpha = static_cast<const Calc*>(a.alpha)[e];
        if (a.alpha) {
            finite = finite_calc(collide_shifted<L, Calc, true, kRcp<M>, IO::SX>(f, static_cast<Calc>(a.omega), alpha,
                                           static_cast<Calc>(a.tau * a.g[0]),
                                           static_cast<Calc>(a.tau * a.g[1]),
                                           static_cast<Calc>(a.tau * a.g[2])));
        } else {
            finite = finite_calc(collide_shifted<L, Calc, false, kRcp<M>, IO::SX>(f, static_cast<Calc>(a.omega), Calc(0),
                                                                          Calc(0), Calc(0), Calc(0)));
        }
    }
    static_for<0, Q>([&](auto qc) {
        constexpr int q = decltype(qc)::value;
        dst[q * Ly.plane + e] = IO::template store<q>(f[q]);
    });
    return finite;
}

// ---------------------------------------------------------------------------
// One interior cell of any kind, for kernels that hand every cell of a region
// to a thread (the batched cell tiles). Non-fluid cells return at once. Every
// direction is first pulled plainly (sources wrapped along periodic axes;
// along a walled axis the source may be a ghost or solid cell); a NEAR cell
// then replaces the values pulled from non-fluid sources by their boundary
// values (as near_cell), and all lanes collide and store together. The bits
// equal those of bulk_cells / near_cell for the same cell. B32 takes the
// cell's own and plain-pull addresses as 32-bit byte offsets (planes under
// 4 GiB). Returns false when a fluid cell read a NaN or infinite value.
// The pulls of a NEAR cell x (element e, ghosted-box coordinates x, y, z)
// from non-fluid sources, f[q] for every link q of its mask, replaced by
// their boundary values, from x's stored values in a.src (the reference's
// boundary pass folded into the pull).
template <class L, Mode M, bool FS, bool B32>
__device__ __forceinline__ void near_rules(const StepArgs& a, typename ElemIdx<B32>::type e, int x, int y, int z,
                                           typename Io<M, L>::Calc (&f)[L::Q]) {
    using IO = Io<M, L>;
    using T = typename IO::T;
    using Calc = typename IO::Calc;
    using I = typename ElemIdx<B32>::type;
    constexpr int Q = L::Q;
    const FieldLayout& Ly = a.L;
    const T* __restrict__ src = static_cast<const T*>(a.src);
    const OwnSums<Calc> os = own_sums_at<L, M, B32>(src, Ly.plane, e);
    auto own = [&](auto kc) {
        constexpr int kk = decltype(kc)::value;
        return IO::template to_calc<kk>(plane_ref<B32>(src, kk, Ly.plane, e));
    };
    const uint32_t lm = glob(a.lmask)[e];
    if constexpr (M != Mode::EXACT64 && M != Mode::DUAL64) {
        // Every link under one condition (not FreeSlip, whose rule reads
        // other cells): one branch on its kind, then that family's loads go
        // out together. The values are rule_value's (same functions). (A
        // caller without the shared-condition table takes the path below.)
        const int lc = a.lcond ? static_cast<int>(glob(a.lcond)[e]) : static_cast<int>(LCOND_MIXED);
        if (lc != LCOND_MIXED) {
            const GCond& cd = cond_at(a.conds, lc);
            const int ck = cd.kind;
            if (!FS || ck != BC_FREESLIP) {
                auto finv = [&](auto qc) { return own(std::integral_constant<int, Q - 1 - decltype(qc)::value>{}); };
                auto put = [&](auto qc, Calc v) {
                    constexpr int q = decltype(qc)::value;
                    f[q] = pick_bits(((lm >> q) & 1u) != 0, IO::template through_storage<q>(v), f[q]);
                };
                if (ck == BC_MOVINGWALL) {
                    static_for<0, Q>([&](auto qc) {
                        put(qc, moving_wall_rule<L, Calc, decltype(qc)::value>(os, finv(qc), cd));
                    });
                } else if (ck == BC_INFLOW) {
                    static_for<0, Q>([&](auto qc) { put(qc, static_cast<Calc>(cd.geq[decltype(qc)::value])); });
                } else if (ck == BC_OUTFLOW || ck == BC_PRESSURE) {
                    static_for<0, Q>([&](auto qc) {
                        put(qc, outflow_rule<L, Calc, decltype(qc)::value>(os, finv(qc), cd));
                    });
                } else {  // NoSlip
                    static_for<0, Q>([&](auto qc) { put(qc, finv(qc)); });
                }
                return;
            }
        }
    }
    static_for<0, Q>([&](auto qc) {
        constexpr int q = decltype(qc)::value;
        constexpr int cx = L::cx(q), cy = L::cy(q), cz = L::cz(q);
        if (lm & (1u << q)) {
            const int sx = wrap_if(x - cx, Ly.xl, Ly.per[0]);
            const int sy = wrap_if(y - cy, Ly.yl, Ly.per[1]);
            const int sz = wrap_if(z - cz, Ly.zl, Ly.per[2]);
            const I se = elem_at<B32>(Ly, sx, sy, sz);
            const uint8_t sk = kind_at(a.kind, se);
            bool stale;
            f[q] = IO::template through_storage<q>(pull_boundary<L, M, q, FS, B32>(
                a, cond_at(a.conds, sk - KIND_COND0), os, own, x, y, z, sx, sy, sz, se, stale));
        }
    });
}

template <class L, Mode M, bool FS = true, bool B32 = false>
__device__ __forceinline__ bool any_cell(const StepArgs& a, int x, int y, int z) {
    using IO = Io<M, L>;
    using T = typename IO::T;
    using Calc = typename IO::Calc;
    using I = typename ElemIdx<B32>::type;
    constexpr int Q = L::Q;
    const FieldLayout& Ly = a.L;
    const I e = elem_at<B32>(Ly, x, y, z);
    const uint8_t k = kind_at(a.kind, e);
    const T* __restrict__ src = static_cast<const T*>(a.src);
    T* __restrict__ dst = static_cast<T*>(a.dst);

    // Pulls are issued before the kind test (sources of any interior cell lie
    // inside the ghosted box), so the kind lookup adds no round trip; a
    // non-fluid cell discards them.
    Calc f[Q];
    // Away from the faces of periodic axes every source is a fixed element
    // offset c_x + c_y px + c_z pxy away; on those faces the source wraps.
    const bool wraps = (Ly.per[0] && (x == 1 || x == Ly.xl)) || (Ly.per[1] && (y == 1 || y == Ly.yl)) ||
                       (Ly.per[2] && (z == 1 || z == Ly.zl));
    if (!wraps) {
        static_for<0, Q>([&](auto qc) {
            constexpr int q = decltype(qc)::value;
            constexpr int cx = L::cx(q), cy = L::cy(q), cz = L::cz(q);
            const int64_t d = cx + (cy > 0 ? Ly.px : (cy < 0 ? -Ly.px : 0)) +
                              (cz > 0 ? Ly.pxy : (cz < 0 ? -Ly.pxy : 0));
            f[q] = IO::template load<q>(plane_ref<B32>(src, q, Ly.plane, e - static_cast<I>(d)));
        });
    } else {
        static_for<0, Q>([&](auto qc) {
            constexpr int q = decltype(qc)::value;
            constexpr int cx = L::cx(q), cy = L::cy(q), cz = L::cz(q);
            const int sx = wrap_if(x - cx, Ly.xl, Ly.per[0]);
            const int sy = wrap_if(y - cy, Ly.yl, Ly.per[1]);
            const int sz = wrap_if(z - cz, Ly.zl, Ly.per[2]);
            f[q] = IO::template load<q>(plane_ref<B32>(src, q, Ly.plane, elem_at<B32>(Ly, sx, sy, sz)));
        });
    }
    if (k >= KIND_COND0) return true;
    if (k == KIND_NEAR && !a.materialised) near_rules<L, M, FS, B32>(a, e, x, y, z, f);

    bool finite = true;
    if constexpr (M == Mode::EXACT64) {
        collide_exact<L>(f, a.tau);
    } else if (a.alpha) {
        const Calc alpha = static_cast<const Calc*>(a.alpha)[e];
        finite = finite_calc(collide_shifted<L, Calc, true, kRcp<M>, IO::SX>(f, static_cast<Calc>(a.omega), alpha,
                                                static_cast<Calc>(a.tau * a.g[0]),
                                                static_cast<Calc>(a.tau * a.g[1]),
                                                static_cast<Calc>(a.tau * a.g[2])));
    } else {
        finite = finite_calc(collide_shifted<L, Calc, false, kRcp<M>, IO::SX>(f, static_cast<Calc>(a.omega), Calc(0),
                                                                      Calc(0), Calc(0), Calc(0)));
    }
    static_for<0, Q>([&](auto qc) {
        constexpr int q = decltype(qc)::value;
        store_stream1<T>(&plane_ref<B32>(dst, q, Ly.plane, e), IO::template store<q>(f[q]));
    });
    return finite;
}

// Blocks of 256 per CU the cell-tile kernel is compiled for: 6 (80 VGPRs)
// with 32- and 16-bit storage up to D3Q19; no bound with 64-bit arithmetic,
// whose values take twice the registers, or with D3Q27, where 80 VGPRs
// spill.
template <class L, Mode M>
constexpr int cells_min_blocks() {
    return (M == Mode::EXACT64 || M == Mode::F64S || M == Mode::DUAL64 || L::Q > 19) ? 1 : 6;
}

// Cell-tile kernel: wavefront w of the grid takes the 64 cells
// 64 (w mod tpp) .. + 63 (x fastest) of the z plane 1 + w / tpp.
template <class L, Mode M, bool FS = true, bool B32 = false>
__global__ void __launch_bounds__(256, (cells_min_blocks<L, M>())) k_cells(StepArgs a, int tpp, double rxl) {
    const int64_t w = static_cast<int64_t>(blockIdx.x) * (blockDim.x >> 6) + (threadIdx.x >> 6);
    const int z = 1 + static_cast<int>(w / tpp);
    if (z > a.L.zl) return;
    const int64_t c = (w % tpp) * 64 + (threadIdx.x & 63);
    const int64_t xl = a.L.xl;
    if (c >= xl * a.L.yl) return;
    // row = c / xl: a double estimate is off by at most one, then corrected
    int64_t y = static_cast<int64_t>(static_cast<double>(c) * rxl);
    int64_t x = c - y * xl;
    if (x < 0) {
        --y;
        x += xl;
    } else if (x >= xl) {
        ++y;
        x -= xl;
    }
    any_cell<L, M, FS, B32>(a, 1 + static_cast<int>(x), 1 + static_cast<int>(y), z);
}

// Near-boundary kernel: one thread per listed NEAR cell.
template <class L, Mode M>
__global__ void __launch_bounds__(128) k_near(StepArgs a) {
    const int64_t i = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
    if (i >= a.nb) return;
    const CellCoord c = a.blist[i];
    near_cell<L, M>(a, c.x, c.y, c.z);
}

}  // namespace lbm_amd
