// Batched time step over many independent domains of uneven size in one
// launch: a work list of tiles (domain, block of cells) walked by a
// persistent grid, plus a combined list of near-boundary cells. Every domain
// keeps its own buffers, cell kinds, boundary table and relaxation time; the
// per-cell arithmetic is the single-domain kernels' (bulk_cells, near_cell),
// so a batched step is bit-identical to stepping each domain alone.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "domain_view.hpp"
#include "storage.hpp"

namespace lbm_amd {

struct BatchView {
    FieldLayout L;
    void* buf[2];
    const uint8_t* kind;
    const uint32_t* lmask;
    const uint8_t* lcond;
    const Cond* conds;
    double tau, omega;
    int bx_log2;  // row tiles: threads along x per tile = 1 << bx_log2 (of 256)
    int nbx, nby;
    int tpp;      // cell tiles: 64-cell tiles per z plane = ceil(xl * yl / 64)
    float rxl;    // 1 / xl (cell tiles: row of a plane index)
};

// Row tile: a block of BX x BY threads over (x, y) at one z, local =
// bx + nbx * (by + nby * (z - 1)). Cell tile: 64 consecutive interior cells
// (x fastest) of one z plane for one wavefront, local = i + tpp * (z - 1),
// cells 64 i .. 64 i + 63 of the plane.
struct TileRef {
    int dom;
    int local;
};

// Kernel forms of a batched step.
enum BatchKernel : int {
    BK_ROWS_SEPARATE = 0,  // row tiles for bulk cells + a launch over the NEAR list
    BK_ROWS_FUSED = 1,     // row tiles stepping their NEAR cells too
    BK_CELLS = 2,          // 64-cell tiles per wavefront, every cell kind in one pass
};

struct NearRef {
    int dom;
    int x, y, z;
};

struct BatchArgs {
    const BatchView* views;
    const TileRef* tiles;
    int64_t ntiles;
    const NearRef* near;
    int64_t nnear;
    int parity;        // src = buf[parity], dst = buf[parity ^ 1]
    int materialised;  // the step pulls stored boundary values
    int freeslip;      // some domain has a FreeSlip condition (cell tiles built with its rule)
    int idx32;         // every domain's plane is under 4 GiB (cell tiles use 32-bit byte offsets)
    int nxcd;          // L2 domains the launch's blocks are dealt to (cell tiles: tile_walk)
    // one flag per domain, set to 1 by a step in which a fluid cell of the
    // domain read a NaN or infinite value (never cleared by the kernels)
    int* status;
};

struct BatchOps {
    int q;
    Mode mode;
    int vecs[4];
    hipError_t (*launch)(const BatchArgs& a, int V, int kernel, int grid_bulk, int grid_near, hipStream_t s);
};

const BatchOps* batch_ops(int q, Mode m);

// Tiles of one wavefront in the cell-tile kernels: t, t + stride, ... below
// end. The hardware deals the blocks of a launch to its nxcd L2 domains in
// turn (block b to domain b mod nxcd). Each domain's wavefronts walk one
// contiguous range of the tile list, sized by its share of the wavefronts,
// so tiles that share halo rows and planes (neighbours in the list, of one
// batched domain) are fetched into one L2. The ranges partition the list,
// so every tile is stepped exactly once. nxcd <= 1: the plain walk first,
// first + (wavefronts in the grid), ... Requires gridDim.x * blockDim.x / 64
// * ntiles < 2^63 and ntiles < 2^30.
struct TileWalk {
    int t, end, stride;
};

__device__ __forceinline__ TileWalk tile_walk(int64_t ntiles, int nxcd) {
    const int wpb = static_cast<int>(blockDim.x) >> 6;
    const int g = static_cast<int>(gridDim.x), b = static_cast<int>(blockIdx.x);
    int slot = b, nslots = g, before = 0;  // before: blocks dealt to the domains ahead of this one
    if (nxcd > 1) {
        const int x = b % nxcd, full = g / nxcd, rem = g % nxcd;
        slot = b / nxcd;
        nslots = full + (x < rem ? 1 : 0);
        before = x * full + (x < rem ? x : rem);
    }
    const int64_t waves = static_cast<int64_t>(g) * wpb;
    const int64_t lo = ntiles * (static_cast<int64_t>(before) * wpb) / waves;
    const int64_t hi = ntiles * (static_cast<int64_t>(before + nslots) * wpb) / waves;
    TileWalk w;
    w.t = __builtin_amdgcn_readfirstlane(static_cast<int>(lo) + slot * wpb + (static_cast<int>(threadIdx.x) >> 6));
    w.end = static_cast<int>(hi);
    w.stride = nslots * wpb;
    return w;
}

}  // namespace lbm_amd
