// Batched adjoint kernels (D3Q19; F32 and F64S storage).
#include "batch_adjoint.hpp"
#include "kernels_adjoint.hpp"

namespace lbm_amd {

namespace {

// Cell (x, y, z) of lane `lane` in cell tile `local` of a domain; false past
// the plane's last cell. The row is a double estimate of c / xl (off by at
// most one) corrected in integers.
__device__ __forceinline__ bool tile_cell(const BatchAdjView& v, int local, int lane, int& x, int& y, int& z) {
    z = 1 + local / v.tpp;
    const int c = (local % v.tpp) * 64 + lane;
    const int xl = v.L.xl;
    if (c >= xl * v.L.yl) return false;
    int yy = static_cast<int>(static_cast<double>(c) * v.rxl);
    int xx = c - yy * xl;
    if (xx < 0) {
        --yy;
        xx += xl;
    } else if (xx >= xl) {
        ++yy;
        xx -= xl;
    }
    x = 1 + xx;
    y = 1 + yy;
    return true;
}

// Slot i of a domain's chunk. The address is the same for the whole
// wavefront; readfirstlane keeps it (and every address derived from it) in
// scalar registers instead of a 64-bit product per lane.
template <class T>
__device__ __forceinline__ T* slot(const BatchAdjView& v, int i) {
    const uint64_t p = reinterpret_cast<uint64_t>(static_cast<T*>(v.chunk) + i * v.slot_elems);
    const uint32_t lo = static_cast<uint32_t>(__builtin_amdgcn_readfirstlane(static_cast<int>(p & 0xffffffffu)));
    const uint32_t hi = static_cast<uint32_t>(__builtin_amdgcn_readfirstlane(static_cast<int>(p >> 32)));
    return reinterpret_cast<T*>((static_cast<uint64_t>(hi) << 32) | lo);
}

// One forward step of every domain from slot src_slot to slot dst_slot.
template <class L, Mode M, bool B32>
__global__ void __launch_bounds__(256) k_badj_recompute(BatchAdjArgs b) {
    using T = typename StorageOf<M>::T;
    const int lane = threadIdx.x & 63;
    const TileWalk w = tile_walk(b.ntiles, b.nxcd);
    for (int t = w.t; t < w.end; t += w.stride) {
        const TileRef tr = b.tiles[t];
        const BatchAdjView& v = b.views[tr.dom];
        int x, y, z;
        if (!tile_cell(v, tr.local, lane, x, y, z)) continue;
        StepArgs a{};
        a.L = v.L;
        a.src = slot<T>(v, b.src_slot);
        a.dst = slot<T>(v, b.dst_slot);
        a.kind = v.kind;
        a.lmask = v.lmask;
        a.lcond = v.lcond;
        a.conds = v.conds;
        a.materialised = b.materialised;
        a.tau = v.tau;
        a.omega = v.omega;
        any_cell<L, M, false, B32>(a, x, y, z);  // the adjoint takes no FreeSlip condition
    }
}

// One adjoint step of every domain: g^{t-1} in gprev_slot, mu^{t+1} in
// mu_in_slot, mu^t into mu_out_slot; kinetic-energy objective.
template <class L, Mode M, bool B32>
__global__ void __launch_bounds__(256) k_badj_backward(BatchAdjArgs b) {
    using T = typename StorageOf<M>::T;
    static_assert(sizeof(T) == sizeof(typename StorageOf<M>::Calc), "mu slots hold the arithmetic type");
    const int lane = threadIdx.x & 63;
    const TileWalk w = tile_walk(b.ntiles, b.nxcd);
    for (int t = w.t; t < w.end; t += w.stride) {
        const TileRef tr = b.tiles[t];
        const BatchAdjView& v = b.views[tr.dom];
        int x, y, z;
        if (!tile_cell(v, tr.local, lane, x, y, z)) continue;
        const auto e = elem_at<B32>(v.L, x, y, z);
        const uint8_t k = kind_at(v.kind, e);
        if (k >= KIND_COND0) continue;
        AdjArgs a{};
        a.L = v.L;
        a.gprev = slot<T>(v, b.gprev_slot);
        a.gcur = v.gfinal;
        a.mu_in = slot<T>(v, b.mu_in_slot);
        a.mu_out = slot<T>(v, b.mu_out_slot);
        a.kind = v.kind;
        a.conds = v.conds;
        a.materialised = b.materialised;
        a.terminal = b.terminal;
        a.tau = v.tau;
        a.omega = v.omega;
        a.objective = OBJ_KINETIC;
        adj_any_cell<L, M, false, B32>(a, x, y, z, e, k == KIND_NEAR, v.gw + 3 * static_cast<size_t>(e));
    }
}

// Per-domain sums of the gradient accumulators: block d sums domain d's
// interior elements with a fixed assignment and a fixed tree, so the result
// does not depend on scheduling.
__global__ void __launch_bounds__(256) k_badj_reduce(const BatchAdjView* views, double* out) {
    const BatchAdjView& v = views[blockIdx.x];
    const int64_t n = static_cast<int64_t>(v.L.xl) * v.L.yl * v.L.zl;
    double s0 = 0.0, s1 = 0.0, s2 = 0.0;
    for (int64_t i = threadIdx.x; i < n; i += blockDim.x) {
        const int x = 1 + static_cast<int>(i % v.L.xl);
        const int y = 1 + static_cast<int>((i / v.L.xl) % v.L.yl);
        const int z = 1 + static_cast<int>(i / (static_cast<int64_t>(v.L.xl) * v.L.yl));
        const double* g = v.gw + 3 * cell_elem(v.L, x, y, z);
        s0 += g[0];
        s1 += g[1];
        s2 += g[2];
    }
    __shared__ double sh[3][256];
    sh[0][threadIdx.x] = s0;
    sh[1][threadIdx.x] = s1;
    sh[2][threadIdx.x] = s2;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (static_cast<int>(threadIdx.x) < w)
            for (int k = 0; k < 3; ++k) sh[k][threadIdx.x] += sh[k][threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x < 3) out[3 * blockIdx.x + threadIdx.x] = sh[threadIdx.x][0];
}

template <class L, Mode M>
struct BatchAdjImpl {
    static hipError_t recompute(const BatchAdjArgs& a, int grid, hipStream_t s) {
        if (a.idx32)
            hipLaunchKernelGGL((k_badj_recompute<L, M, true>), dim3(grid), dim3(256), 0, s, a);
        else
            hipLaunchKernelGGL((k_badj_recompute<L, M, false>), dim3(grid), dim3(256), 0, s, a);
        return hipGetLastError();
    }
    static hipError_t backward(const BatchAdjArgs& a, int grid, hipStream_t s) {
        if (a.idx32)
            hipLaunchKernelGGL((k_badj_backward<L, M, true>), dim3(grid), dim3(256), 0, s, a);
        else
            hipLaunchKernelGGL((k_badj_backward<L, M, false>), dim3(grid), dim3(256), 0, s, a);
        return hipGetLastError();
    }
    static hipError_t reduce(const BatchAdjView* views, int ndom, double* out, hipStream_t s) {
        hipLaunchKernelGGL(k_badj_reduce, dim3(ndom), dim3(256), 0, s, views, out);
        return hipGetLastError();
    }
    static BatchAdjOps make() {
        BatchAdjOps o{};
        o.recompute = &recompute;
        o.backward = &backward;
        o.reduce = &reduce;
        return o;
    }
};

const BatchAdjOps kF32 = BatchAdjImpl<D3Q19, Mode::F32>::make();
const BatchAdjOps kF64S = BatchAdjImpl<D3Q19, Mode::F64S>::make();

}  // namespace

const BatchAdjOps* batch_adj_ops(int q, Mode m) {
    if (q != 19) return nullptr;
    if (m == Mode::F32) return &kF32;
    if (m == Mode::F64S) return &kF64S;
    return nullptr;
}

}  // namespace lbm_amd
