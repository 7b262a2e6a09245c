// Batched adjoint over the domains of a batch (D3Q19, F32 / F64S).
//
// Every domain owns one device chunk of field-sized slots: checkpoints of the
// forward run, the recomputed states of the current segment and two adjoint
// (mu) buffers. The kernels walk the batch's 64-cell tiles; a launch names
// its slots by index, the same for every domain, so one launch advances all
// domains by one recompute or one adjoint step. The per-cell arithmetic is
// the single-domain code (any_cell forward, adj_cell backward).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "batch.hpp"
#include "domain_view.hpp"
#include "storage.hpp"

namespace lbm_amd {

struct BatchAdjView {
    FieldLayout L;
    void* chunk;          // slots of slot_elems storage elements each
    int64_t slot_elems;   // Q * plane
    const void* gfinal;   // the forward run's final field (the domain's buffer)
    const uint8_t* kind;
    const uint32_t* lmask;
    const uint8_t* lcond;
    const Cond* conds;
    double tau, omega;
    double* gw;           // 3 doubles per element: dJ/du_wall accumulated per cell
    int tpp;              // 64-cell tiles per z plane
    double rxl;           // 1 / xl
};

struct BatchAdjArgs {
    const BatchAdjView* views;
    const TileRef* tiles;  // cell tiles: local = i + tpp * (z - 1)
    int64_t ntiles;
    int src_slot, dst_slot;                  // recompute step
    int gprev_slot, mu_in_slot, mu_out_slot; // adjoint step
    int terminal;
    int materialised;
    int idx32;  // every domain's plane is under 4 GiB (32-bit byte offsets)
    int nxcd;   // L2 domains the launch's blocks are dealt to (tile_walk)
};

struct BatchAdjOps {
    hipError_t (*recompute)(const BatchAdjArgs& a, int grid, hipStream_t s);
    hipError_t (*backward)(const BatchAdjArgs& a, int grid, hipStream_t s);
    // out[3 d + k] = sum over the elements of domain d of gw (fixed order)
    hipError_t (*reduce)(const BatchAdjView* views, int ndom, double* out, hipStream_t s);
};

const BatchAdjOps* batch_adj_ops(int q, Mode m);

}  // namespace lbm_amd
