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
    a.lmask = v.lmask;
    a.lcond = v.lcond;
    a.conds = v.conds;
    a.materialised = mat;
    a.tau = v.tau;
    a.omega = v.omega;
    return a;
}

// Persistent: each block walks tiles blockIdx.x, blockIdx.x + gridDim.x, ...
// With FUSED the tile also steps its NEAR cells (boundary rules) right after
// its bulk cells, while their rows are still in cache; without it they are
// left to k_batch_near.
template <class L, Mode M, int V, bool FUSED>
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
        const int x0 = 1 + ((bx << bxl) + tx) * V, y = 1 + by * by_n + ty;
        bool finite;
        if constexpr (FUSED) {
            bool nl[V];
#pragma unroll
            for (int i = 0; i < V; ++i) nl[i] = false;
            finite = bulk_cells<L, M, V, true, false>(a, x0, y, z, nl);
            for (int i = 0; i < V; ++i)
                if (nl[i]) finite &= near_cell<L, M>(a, x0 + i, y, z);
        } else {
            finite = bulk_cells<L, M, V, true, false>(a, x0, y, z);
        }
        // every lane that saw a non-finite value writes the same flag
        if (!finite) b.status[tr.dom] = 1;
    }
}

// Cell tiles: each wavefront walks its tiles (tile_walk, one L2 domain's
// range of the list); each lane takes one cell of its tile.
// The tile index is wavefront-uniform; readfirstlane states it, so the tile
// and the domain view are read into scalar registers once per wavefront.
template <class L, Mode M, bool FS, bool B32>
__global__ void __launch_bounds__(256) k_batch_cells(BatchArgs b) {
    const int lane = threadIdx.x & 63;
    const TileWalk w = tile_walk(b.ntiles, b.nxcd);
    for (int t = w.t; t < w.end; t += w.stride) {
        const TileRef tr = b.tiles[t];
        const BatchView& v = b.views[tr.dom];
        const int z = 1 + tr.local / v.tpp;
        const int c = (tr.local % v.tpp) * 64 + lane;
        if (c < v.L.xl * v.L.yl) {
            // row = c / xl: float estimate (off by at most one below 2^22), then corrected
            int y = static_cast<int>(static_cast<float>(c) * v.rxl);
            int x = c - y * v.L.xl;
            if (x < 0) {
                --y;
                x += v.L.xl;
            } else if (x >= v.L.xl) {
                ++y;
                x -= v.L.xl;
            }
            const StepArgs a = view_args(v, b.parity, b.materialised);
            if (!any_cell<L, M, FS, B32>(a, 1 + x, 1 + y, z)) b.status[tr.dom] = 1;
        }
    }
}

template <class L, Mode M>
__global__ void __launch_bounds__(128) k_batch_near(BatchArgs b) {
    for (int64_t i = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; i < b.nnear;
         i += static_cast<int64_t>(gridDim.x) * blockDim.x) {
        const NearRef n = b.near[i];
        const StepArgs a = view_args(b.views[n.dom], b.parity, b.materialised);
        if (!near_cell<L, M>(a, n.x, n.y, n.z)) b.status[n.dom] = 1;
    }
}

template <class L, Mode M, int... Vs>
struct BatchImpl {
    template <int V>
    static bool try_v(const BatchArgs& a, int V_, bool fused, int grid, hipStream_t s) {
        if (V_ != V) return false;
        if (fused)
            hipLaunchKernelGGL((k_batch_bulk<L, M, V, true>), dim3(grid), dim3(256), 0, s, a);
        else
            hipLaunchKernelGGL((k_batch_bulk<L, M, V, false>), dim3(grid), dim3(256), 0, s, a);
        return true;
    }
    static hipError_t launch(const BatchArgs& a, int V, int kernel, int grid_bulk, int grid_near,
                             hipStream_t s) {
        if (kernel == BK_CELLS) {
            if (a.ntiles > 0) {
                if (a.freeslip) {
                    if (a.idx32)
                        hipLaunchKernelGGL((k_batch_cells<L, M, true, true>), dim3(grid_bulk), dim3(256), 0, s, a);
                    else
                        hipLaunchKernelGGL((k_batch_cells<L, M, true, false>), dim3(grid_bulk), dim3(256), 0, s, a);
                } else {
                    if (a.idx32)
                        hipLaunchKernelGGL((k_batch_cells<L, M, false, true>), dim3(grid_bulk), dim3(256), 0, s, a);
                    else
                        hipLaunchKernelGGL((k_batch_cells<L, M, false, false>), dim3(grid_bulk), dim3(256), 0, s, a);
                }
            }
            return hipGetLastError();
        }
        const bool fused = kernel == BK_ROWS_FUSED;
        if (a.ntiles > 0 && !(try_v<Vs>(a, V, fused, grid_bulk, s) || ...)) return hipErrorInvalidValue;
        if (!fused && a.nnear > 0)
            hipLaunchKernelGGL((k_batch_near<L, M>), dim3(grid_near), dim3(128), 0, s, a);
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
