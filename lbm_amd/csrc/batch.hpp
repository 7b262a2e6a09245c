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

}  // namespace lbm_amd
