// Device-visible description of one domain: layout of the distribution
// fields in HBM, cell kinds, boundary-condition table, relaxation and forcing.
//
// Layout. The reference indexes the ghosted box (xl+2)(yl+2)(zl+2) with
// x fastest (include/domain.hpp:61-64) and stores each cell's Q pdfs together.
// Here every direction q has its own plane (structure of arrays) and rows are
// padded so that interior x = 1 of every row starts on a 128-byte line:
//
//   elem(q, x, y, z) = q * plane + off + x + px * (y + ny * z),
//   off = A - 1, px = roundup(xl + 2, A), A = 128 / sizeof(storage element).
//
// The same element index (without the q term) addresses the uint8 cell-kind
// array. Each plane ends with a few spare elements so that vector loads of the
// last row stay inside the allocation.
#pragma once

#include <cstdint>

namespace lbm_amd {

// Cell kinds (uint8 per ghosted cell).
enum : uint8_t {
    KIND_BULK = 0,      // interior fluid cell whose Q sources are all fluid
    KIND_NEAR = 1,      // interior fluid cell with at least one non-fluid source
    KIND_COND0 = 2,     // non-fluid cell with boundary condition (kind - 2)
    KIND_INERT = 254,   // never read (unassigned ghost, unreachable solid)
};
constexpr int kMaxConds = KIND_INERT - KIND_COND0;  // 252 distinct conditions
constexpr uint8_t LCOND_MIXED = 0xFF;  // links of a NEAR cell under more than one condition

// Boundary-condition kinds, numbered as in include/lbm_hip.h.
enum : int {
    BC_NOSLIP = 1,
    BC_MOVINGWALL = 2,
    BC_FREESLIP = 3,
    BC_OUTFLOW = 4,
    BC_INFLOW = 5,
    BC_PRESSURE = 6,
};

// Rule families a kernel is built with (RULES): NoSlip is always there; a
// condition of a family left out would be evaluated as NoSlip, so the host
// leaves a family out only when the domain has no condition of it.
enum : int { RULE_MOVINGWALL = 1, RULE_INFLOW = 2, RULE_OUTFLOW = 4, RULES_ALL = 7 };

struct Cond {
    int kind;
    int pad;
    double vel[3];
    double rho;
    // Inflow only: the reference equilibrium feq(rho_ref, u_in) in fp64,
    // and the same shifted by the weights (feq_q - w_q) for the shifted modes.
    double feq[27];
    double geq[27];
    // MovingWall: c_q . u_w evaluated in fp64 as c_qx u_x + c_qy u_y + c_qz u_z
    // (left to right), and that value rounded to fp32, for the shifted modes.
    double cu64[27];
    float cu32[27];
};

struct CellCoord {
    int x, y, z, pad;
};

struct FieldLayout {
    int64_t off;    // element offset of (0,0,0)
    int64_t px;     // row pitch (elements)
    int64_t pxy;    // px * ny
    int64_t plane;  // elements per direction plane
    int xl, yl, zl;
    int per[3];     // periodic axes
};

struct StepArgs {
    FieldLayout L;
    const void* src;
    void* dst;
    const uint8_t* kind;
    // per element: bit q set when the cell is fluid and its source x - c_q is
    // not (the links a near-boundary cell takes from boundary rules)
    const uint32_t* lmask;
    // per element: for a NEAR cell, the condition index shared by every link
    // of its mask, or LCOND_MIXED where they differ (edges and corners);
    // may be null (every cell then takes the per-link path)
    const uint8_t* lcond;
    const CellCoord* blist;  // cells of kind NEAR
    int64_t nb;
    const Cond* conds;
    int materialised;   // 1: non-fluid sources hold values to be read as stored
    double tau;         // relaxation time
    double omega;       // 1 / tau
    // Brinkman extension: per-cell porosity in the Calc type of the mode (same
    // element index as kind) and body acceleration; alpha == nullptr disables it.
    const void* alpha;
    double g[3];
};

__host__ __device__ inline int64_t cell_elem(const FieldLayout& L, int x, int y, int z) {
    return L.off + x + L.px * y + L.pxy * z;
}

}  // namespace lbm_amd
