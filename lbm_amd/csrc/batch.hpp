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
    const Cond* conds;
    double tau, omega;
    int bx_log2;  // threads along x per tile = 1 << bx_log2 (of 256)
    int nbx, nby;
    int pad;
};

struct TileRef {
    int dom;
    int local;  // bx + nbx * (by + nby * (z - 1))
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
};

struct BatchOps {
    int q;
    Mode mode;
    int vecs[4];
    hipError_t (*launch)(const BatchArgs& a, int V, int grid_bulk, int grid_near, hipStream_t s);
};

const BatchOps* batch_ops(int q, Mode m);

}  // namespace lbm_amd
