// Batched kernels (D3Q19; F32, F16, BF16, F64S storage).
#include "batch.hpp"
#include "kernels.hpp"

namespace lbm_amd {

namespace {

__device__ __forceinline__ StepArgs view_args(const BatchView& v, int parity, int mat) {
    StepArgs a{};
    a.L = v.L;
    a.src = v.buf[parity];
    a.dst = v.buf[parity ^ 1];
    a.kind = v.kind;
    a.conds = v.conds;
    a.materialised = mat;
    a.tau = v.tau;
    a.omega = v.omega;
    return a;
}

// Persistent: each block walks tiles blockIdx.x, blockIdx.x + gridDim.x, ...
template <class L, Mode M, int V>
__global__ void __launch_bounds__(256) k_batch_bulk(BatchArgs b) {
    for (int64_t t = blockIdx.x; t < b.ntiles; t += gridDim.x) {
        const TileRef tr = b.tiles[t];
        const BatchView& v = b.views[tr.dom];
        const int bxl = v.bx_log2;
        const int tx = threadIdx.x & ((1 << bxl) - 1);
        const int ty = threadIdx.x >> bxl;
        const int by_n = 256 >> bxl;
        const int bx = tr.local % v.nbx;
        const int r = tr.local / v.nbx;
        const int by = r % v.nby;
        const int z = 1 + r / v.nby;
        const StepArgs a = view_args(v, b.parity, b.materialised);
        bulk_cells<L, M, V, true, false>(a, 1 + ((bx << bxl) + tx) * V, 1 + by * by_n + ty, z);
    }
}

template <class L, Mode M>
__global__ void __launch_bounds__(128) k_batch_near(BatchArgs b) {
    for (int64_t i = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; i < b.nnear;
         i += static_cast<int64_t>(gridDim.x) * blockDim.x) {
        const NearRef n = b.near[i];
        const StepArgs a = view_args(b.views[n.dom], b.parity, b.materialised);
        near_cell<L, M>(a, n.x, n.y, n.z);
    }
}

template <class L, Mode M, int... Vs>
struct BatchImpl {
    template <int V>
    static bool try_v(const BatchArgs& a, int V_, int grid, hipStream_t s) {
        if (V_ != V) return false;
        hipLaunchKernelGGL((k_batch_bulk<L, M, V>), dim3(grid), dim3(256), 0, s, a);
        return true;
    }
    static hipError_t launch(const BatchArgs& a, int V, int grid_bulk, int grid_near, hipStream_t s) {
        if (a.ntiles > 0 && !(try_v<Vs>(a, V, grid_bulk, s) || ...)) return hipErrorInvalidValue;
        if (a.nnear > 0) hipLaunchKernelGGL((k_batch_near<L, M>), dim3(grid_near), dim3(128), 0, s, a);
        return hipGetLastError();
    }
    static BatchOps make() {
        BatchOps o{};
        o.q = L::Q;
        o.mode = M;
        const int v[] = {Vs...};
        for (int i = 0; i < 4; ++i) o.vecs[i] = i < static_cast<int>(sizeof...(Vs)) ? v[i] : 0;
        o.launch = &launch;
        return o;
    }
};

const BatchOps kF32 = BatchImpl<D3Q19, Mode::F32, 1, 2, 4>::make();
const BatchOps kF16 = BatchImpl<D3Q19, Mode::F16, 2, 4>::make();
const BatchOps kBF16 = BatchImpl<D3Q19, Mode::BF16, 2, 4>::make();
const BatchOps kF64S = BatchImpl<D3Q19, Mode::F64S, 1, 2>::make();

}  // namespace

const BatchOps* batch_ops(int q, Mode m) {
    if (q != 19) return nullptr;
    switch (m) {
        case Mode::F32: return &kF32;
        case Mode::F16: return &kF16;
        case Mode::BF16: return &kBF16;
        case Mode::F64S: return &kF64S;
        default: return nullptr;
    }
}

}  // namespace lbm_amd
