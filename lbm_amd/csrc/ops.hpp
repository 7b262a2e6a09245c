// Host-side launch table of one (lattice, storage mode) pair. Each mode
// translation unit fills one ModeOps per lattice; the C-ABI picks the table
// when a domain is created and calls through it.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "domain_view.hpp"
#include "storage.hpp"

namespace lbm_amd {

struct GhostDims;

// Tile of the two-step pass (k_bulk2): cells along x and y of one block.
constexpr int TB_TX = 64;
constexpr int TB_TY = 8;

// Events recorded after kernel launches (lbm_time_steps): mark() records
// the next event and tags the launch it ends (0: bulk, cell-tile or two-step
// pass; 1: near-boundary kernel). One record per launch keeps the timing's
// own cost out of the stream as far as it can.
struct EventLog {
    hipEvent_t* ev;
    int* kind;
    size_t n, cap;
    void mark(hipStream_t s, int k) {
        if (n < cap) {
            (void)hipEventRecord(ev[n], s);
            kind[n] = k;
            ++n;
        }
    }
};

struct StepPlan {
    int V;          // cells per thread along x in k_bulk
    int BX, BY;     // k_bulk block shape (BX * BY == 256)
    int nbx, nby;   // k_bulk blocks along x-vectors and rows
    int64_t nblocks_bulk;
    bool flags;     // domain has non-bulk interior cells: read the kind array
    bool brink;     // Brinkman field present
    bool has_bulk;  // at least one BULK cell
    // Cell tiles (k_cells) instead of k_bulk + k_near: 64 consecutive cells of
    // a z plane per wavefront, every cell kind in one pass.
    bool cells;
    bool idx32;            // cell-tile and in-place kernels index a plane with 32-bit byte offsets (plane < 4 GiB)
    bool freeslip;         // a FreeSlip condition exists (cell tiles built with its rule)
    int tpp;               // 64-cell tiles per z plane
    int64_t nblocks_cells; // blocks of 4 wavefronts
    double rxl;            // 1 / xl
    // Two steps per pass (k_bulk2): 64 x 8 cell tiles in x-y, tb_zc planes
    // per block along z.
    bool pairs;
    // Two steps per pass for domains with boundary cells (k_emit_rules +
    // k_wall2), same tiles and chunks; the first step after set-up stays a
    // single step. wrules: the boundary rule families built in (RULE_*).
    bool wpairs;
    int wrules;
    int tb_ntx, tb_nty, tb_zc;
    int tb_nxcd;  // L2 domains the blocks are grouped for (1: identity order)
    int64_t tb_nblocks;
};

struct ModeOps {
    Mode mode;
    int q;
    int default_vec;
    int default_bx;  // threads along x of a k_bulk block when not tuned (0: row width)
    // Supported values of StepPlan::V, 0-terminated.
    int vecs[4];
    // bulk + near for one step (src/dst in args); log (may be null) marks
    // the end of each launch
    hipError_t (*step)(const StepArgs& a, const StepPlan& p, hipStream_t s, EventLog* log);
    // two steps in one pass for bulk-only domains (null where not built)
    hipError_t (*step2)(const StepArgs& a, const StepPlan& p, hipStream_t s, EventLog* log);
    // blocks of the two-step pass resident per CU (0 where not built)
    int (*step2_blocks_per_cu)(bool brink);
    // two steps in one pass for domains with boundary cells (null where not built)
    hipError_t (*step2w)(const StepArgs& a, const StepPlan& p, hipStream_t s, EventLog* log);
    int (*step2w_blocks_per_cu)(int rules);
    hipError_t (*get_pdf)(const FieldLayout& Ly, const void* src, double* out, int nx, int ny,
                          int nz, hipStream_t s);
    hipError_t (*set_pdf)(const FieldLayout& Ly, void* dst, const double* in, int nx, int ny,
                          int nz, hipStream_t s);
    hipError_t (*moments)(const FieldLayout& Ly, const void* src, double* rho, double* u,
                          hipStream_t s);
    // partial must hold 4 * nblocks doubles; out 4 doubles.
    hipError_t (*conservation)(const FieldLayout& Ly, const void* src, const uint8_t* kind,
                               double* partial, int nblocks, double* out, hipStream_t s);
    hipError_t (*init_eq)(const FieldLayout& Ly, void* dst, const uint8_t* kind,
                          const double* rho, const double* u, hipStream_t s);
    hipError_t (*materialise)(const StepArgs& a, int nx, int ny, int nz, hipStream_t s);
    // in-place (AA) layout: one step from representation phase (0: natural,
    // 1: swapped) to the other; and the swapped -> natural conversion
    hipError_t (*aa_step)(const StepArgs& a, const StepPlan& p, int phase, hipStream_t s);
    hipError_t (*aa_swap)(const StepArgs& a, const StepPlan& p, hipStream_t s);
};

// Defined in the mode translation units.
const ModeOps* ops_exact(int q);
const ModeOps* ops_fast_q15(Mode m);
const ModeOps* ops_fast_q19(Mode m);
const ModeOps* ops_fast_q27(Mode m);

inline const ModeOps* mode_ops(int q, Mode m) {
    if (m == Mode::EXACT64) return ops_exact(q);
    switch (q) {
        case 15: return ops_fast_q15(m);
        case 19: return ops_fast_q19(m);
        case 27: return ops_fast_q27(m);
        default: return nullptr;
    }
}

}  // namespace lbm_amd
