// Per-field kernels of one mode: host-format conversion (get/set of the
// whole ghosted field), moments of interior cells, fp64 conservation sums,
// equilibrium initialisation and materialisation of boundary-cell values.
#pragma once

#include "collide.hpp"
#include "domain_view.hpp"
#include "kernels.hpp"
#include "lbm_hip/lattice.hpp"
#include "storage.hpp"

namespace lbm_amd {

struct GhostDims {
    int nx, ny, nz;
};

// Ghosted box -> reference AoS doubles, idx(x,y,z)*Q + q (include/domain.hpp:61-64).
template <class L, Mode M>
__global__ void k_get_pdf(FieldLayout Ly, const void* src_, double* out, GhostDims g) {
    using IO = Io<M, L>;
    using T = typename IO::T;
    const int64_t n = static_cast<int64_t>(g.nx) * g.ny * g.nz;
    const int64_t i = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
    if (i >= n) return;
    const int x = static_cast<int>(i % g.nx);
    const int y = static_cast<int>((i / g.nx) % g.ny);
    const int z = static_cast<int>(i / (static_cast<int64_t>(g.nx) * g.ny));
    const int64_t e = cell_elem(Ly, x, y, z);
    const T* src = static_cast<const T*>(src_);
    static_for<0, L::Q>([&](auto qc) {
        constexpr int q = decltype(qc)::value;
        double v = static_cast<double>(IO::template to_calc<q>(src[q * Ly.plane + e]));
        if constexpr (StorageOf<M>::shifted) v += weight<L>(q);
        out[i * L::Q + q] = v;
    });
}

template <class L, Mode M>
__global__ void k_set_pdf(FieldLayout Ly, void* dst_, const double* in, GhostDims g) {
    using IO = Io<M, L>;
    using T = typename IO::T;
    using Calc = typename IO::Calc;
    const int64_t n = static_cast<int64_t>(g.nx) * g.ny * g.nz;
    const int64_t i = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
    if (i >= n) return;
    const int x = static_cast<int>(i % g.nx);
    const int y = static_cast<int>((i / g.nx) % g.ny);
    const int z = static_cast<int>(i / (static_cast<int64_t>(g.nx) * g.ny));
    const int64_t e = cell_elem(Ly, x, y, z);
    T* dst = static_cast<T*>(dst_);
    static_for<0, L::Q>([&](auto qc) {
        constexpr int q = decltype(qc)::value;
        double v = in[i * L::Q + q];
        if constexpr (StorageOf<M>::shifted) v -= weight<L>(q);
        dst[q * Ly.plane + e] = IO::template from_calc<q>(static_cast<Calc>(v));
    });
}

// Stored values of one cell as reference pdfs in double.
template <class L, Mode M>
__device__ __forceinline__ void load_cell_f64(const FieldLayout& Ly, const void* src_, int64_t e,
                                              double (&f)[L::Q]) {
    using IO = Io<M, L>;
    using T = typename IO::T;
    const T* src = static_cast<const T*>(src_);
    static_for<0, L::Q>([&](auto qc) {
        constexpr int q = decltype(qc)::value;
        f[q] = static_cast<double>(IO::template to_calc<q>(src[q * Ly.plane + e]));
    });
}

// density() then velocity(density) of interior cells, x fastest
// (include/io/vtk.hpp:62-73). EXACT64 follows the reference sums; the shifted
// modes return rho = 1 + sum g and u = sum c g / rho evaluated in double.
template <class L, Mode M>
__global__ void k_moments(FieldLayout Ly, const void* src, double* rho, double* u) {
    const int64_t n = static_cast<int64_t>(Ly.xl) * Ly.yl * Ly.zl;
    const int64_t i = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
    if (i >= n) return;
    const int x = 1 + static_cast<int>(i % Ly.xl);
    const int y = 1 + static_cast<int>((i / Ly.xl) % Ly.yl);
    const int z = 1 + static_cast<int>(i / (static_cast<int64_t>(Ly.xl) * Ly.yl));
    double f[L::Q];
    load_cell_f64<L, M>(Ly, src, cell_elem(Ly, x, y, z), f);
    double r, v[3];
    if constexpr (M == Mode::EXACT64) {
        moments_exact<L>(f, r, v);
    } else {
        double d = 0.0, jx = 0.0, jy = 0.0, jz = 0.0;
        static_for<0, L::Q>([&](auto qc) {
            constexpr int q = decltype(qc)::value;
            d += f[q];
            jx += L::cx(q) * f[q];
            jy += L::cy(q) * f[q];
            jz += L::cz(q) * f[q];
        });
        r = 1.0 + d;
        v[0] = jx / r;
        v[1] = jy / r;
        v[2] = jz / r;
    }
    if (rho) rho[i] = r;
    if (u) {
        u[3 * i + 0] = v[0];
        u[3 * i + 1] = v[1];
        u[3 * i + 2] = v[2];
    }
}

// Block sum of 4 doubles over a 256-thread block; result valid in thread 0.
__device__ __forceinline__ void block_sum4(double (&s)[4]) {
    __shared__ double part[4][4];  // [wave][component], 256 threads = 4 waves
#pragma unroll
    for (int k = 0; k < 4; ++k) {
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) s[k] += __shfl_down(s[k], off, 64);
    }
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (lane == 0)
#pragma unroll
        for (int k = 0; k < 4; ++k) part[wave][k] = s[k];
    __syncthreads();
    if (threadIdx.x == 0)
#pragma unroll
        for (int k = 0; k < 4; ++k) s[k] = part[0][k] + part[1][k] + part[2][k] + part[3][k];
}

// fp64 mass and momentum over interior fluid cells: per-block partials, then
// one block adds the partials in a fixed order (deterministic).
template <class L, Mode M>
__global__ void __launch_bounds__(256) k_conservation(FieldLayout Ly, const void* src,
                                                      const uint8_t* kind, double* partial) {
    const int64_t n = static_cast<int64_t>(Ly.xl) * Ly.yl * Ly.zl;
    double s[4] = {0.0, 0.0, 0.0, 0.0};
    for (int64_t i = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; i < n;
         i += static_cast<int64_t>(gridDim.x) * blockDim.x) {
        const int x = 1 + static_cast<int>(i % Ly.xl);
        const int y = 1 + static_cast<int>((i / Ly.xl) % Ly.yl);
        const int z = 1 + static_cast<int>(i / (static_cast<int64_t>(Ly.xl) * Ly.yl));
        const int64_t e = cell_elem(Ly, x, y, z);
        if (kind[e] >= KIND_COND0) continue;
        double f[L::Q];
        load_cell_f64<L, M>(Ly, src, e, f);
        static_for<0, L::Q>([&](auto qc) {
            constexpr int q = decltype(qc)::value;
            double v = f[q];
            if constexpr (StorageOf<M>::shifted) v += weight<L>(q);
            s[0] += v;
            s[1] += L::cx(q) * v;
            s[2] += L::cy(q) * v;
            s[3] += L::cz(q) * v;
        });
    }
    block_sum4(s);
    if (threadIdx.x == 0)
#pragma unroll
        for (int k = 0; k < 4; ++k) partial[4 * blockIdx.x + k] = s[k];
}

static __global__ void __launch_bounds__(256) k_sum_partials(const double* partial, int nblocks,
                                                      double* out) {
    double s[4] = {0.0, 0.0, 0.0, 0.0};
    for (int b = threadIdx.x; b < nblocks; b += blockDim.x)
#pragma unroll
        for (int k = 0; k < 4; ++k) s[k] += partial[4 * b + k];
    block_sum4(s);
    if (threadIdx.x == 0)
#pragma unroll
        for (int k = 0; k < 4; ++k) out[k] = s[k];
}

// Interior fluid cells <- feq(rho, u) from x-fastest interior arrays.
template <class L, Mode M>
__global__ void k_init_eq(FieldLayout Ly, void* dst_, const uint8_t* kind, const double* rho,
                          const double* u) {
    using IO = Io<M, L>;
    using T = typename IO::T;
    using Calc = typename IO::Calc;
    const int64_t n = static_cast<int64_t>(Ly.xl) * Ly.yl * Ly.zl;
    const int64_t i = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
    if (i >= n) return;
    const int x = 1 + static_cast<int>(i % Ly.xl);
    const int y = 1 + static_cast<int>((i / Ly.xl) % Ly.yl);
    const int z = 1 + static_cast<int>(i / (static_cast<int64_t>(Ly.xl) * Ly.yl));
    const int64_t e = cell_elem(Ly, x, y, z);
    if (kind[e] >= KIND_COND0) return;
    T* dst = static_cast<T*>(dst_);
    const double r = rho[i];
    const double v[3] = {u[3 * i + 0], u[3 * i + 1], u[3 * i + 2]};
    static_for<0, L::Q>([&](auto qc) {
        constexpr int q = decltype(qc)::value;
        double val;
        if constexpr (M == Mode::EXACT64) {
            val = feq_exact<L, q>(r, v);
        } else {
            const double cu = L::cx(q) * v[0] + L::cy(q) * v[1] + L::cz(q) * v[2];
            const double uu = v[0] * v[0] + v[1] * v[1] + v[2] * v[2];
            val = weight<L>(q) * ((r - 1.0) + r * (3.0 * cu + 4.5 * cu * cu - 1.5 * uu));
        }
        dst[q * Ly.plane + e] = IO::template from_calc<q>(static_cast<Calc>(val));
    });
}

// Writes into every non-fluid cell b the rule values the reference's collide
// pass 2 would have left there (only directions whose b + c_q is an interior
// fluid cell; other directions keep their stored value). Reads and writes the
// same buffer: threads write only non-fluid cells and read only fluid ones.
template <class L, Mode M>
__global__ void k_materialise(StepArgs a, GhostDims g) {
    using IO = Io<M, L>;
    using T = typename IO::T;
    using Calc = typename IO::Calc;
    constexpr int Q = L::Q;
    const FieldLayout& Ly = a.L;
    const int64_t n = static_cast<int64_t>(g.nx) * g.ny * g.nz;
    const int64_t i = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
    if (i >= n) return;
    const int bx = static_cast<int>(i % g.nx);
    const int by = static_cast<int>((i / g.nx) % g.ny);
    const int bz = static_cast<int>(i / (static_cast<int64_t>(g.nx) * g.ny));
    // cells at a ghost coordinate of a periodic axis are never read
    if ((Ly.per[0] && (bx == 0 || bx == g.nx - 1)) || (Ly.per[1] && (by == 0 || by == g.ny - 1)) ||
        (Ly.per[2] && (bz == 0 || bz == g.nz - 1)))
        return;
    const int64_t be = cell_elem(Ly, bx, by, bz);
    const uint8_t k = a.kind[be];
    if (k < KIND_COND0 || k >= KIND_INERT) return;
    const GCond& cd = cond_at(a.conds, k - KIND_COND0);
    const T* src = static_cast<const T*>(a.src);
    T* dst = static_cast<T*>(a.dst);
    static_for<0, Q>([&](auto qc) {
        constexpr int q = decltype(qc)::value;
        const int x = wrap_if(bx + L::cx(q), Ly.xl, Ly.per[0]);
        const int y = wrap_if(by + L::cy(q), Ly.yl, Ly.per[1]);
        const int z = wrap_if(bz + L::cz(q), Ly.zl, Ly.per[2]);
        if (x < 1 || x > Ly.xl || y < 1 || y > Ly.yl || z < 1 || z > Ly.zl) return;
        const int64_t xe = cell_elem(Ly, x, y, z);
        if (a.kind[xe] >= KIND_COND0) return;
        Calc own[Q];
        static_for<0, Q>([&](auto kc) {
            constexpr int kk = decltype(kc)::value;
            own[kk] = IO::template to_calc<kk>(src[kk * Ly.plane + xe]);
        });
        bool stale;
        const Calc v = pull_nonfluid<L, M, q>(a, cd, own, x, y, z, bx, by, bz, be, stale);
        if (!stale) dst[q * Ly.plane + be] = IO::template from_calc<q>(v);
    });
}

}  // namespace lbm_amd
