// Launch table of the adjoint / tangent kernels of one (lattice, mode).
#pragma once

#include <hip/hip_runtime.h>

#include "ops.hpp"

namespace lbm_amd {

enum : int { OBJ_PLANE_MISFIT = 1, OBJ_KINETIC = 2 };

struct AdjArgs {
    FieldLayout L;
    const void* gprev;   // g^{t-1}, storage type
    const void* gcur;    // g^T (terminal step only)
    const void* mu_in;   // mu^{t+1}, Calc type, same element layout
    void* mu_out;        // mu^t
    const uint8_t* kind;
    const CellCoord* blist;
    int64_t nb;
    const Cond* conds;
    int materialised;    // forward step t pulled stored boundary values
    int terminal;        // t == T: lambda from the objective
    double tau, omega;
    const void* alpha;   // Calc type per element, or null
    double g[3];
    double* grad_alpha;  // fp64 per element, accumulated
    double* grad_wall;   // fp64 3 per NEAR cell, accumulated
    int objective;
    int plane_x;
    const double* target;  // plane misfit: 3 doubles per (y, z), y fastest
};

struct AdjOps {
    // one backward step: k_adj_bulk over the V = 1 plan p, k_adj_near over the NEAR list
    hipError_t (*backward_step)(const AdjArgs& a, const StepPlan& p, bool brink, hipStream_t s);
    // objective of a field of this mode / of a dual field: out[0] = J, out[1] = dJ (dual)
    hipError_t (*objective)(const FieldLayout& Ly, const void* src, const uint8_t* kind, int obj,
                            int plane, const double* target, double* partial, int nblocks,
                            double* out, hipStream_t s);
    hipError_t (*objective_dual)(const FieldLayout& Ly, const void* src, const uint8_t* kind, int obj,
                                 int plane, const double* target, double* partial, int nblocks,
                                 double* out, hipStream_t s);
    // field of this mode -> dual field (value, tangent 0)
    hipError_t (*to_dual)(const FieldLayout& Ly, const void* src, void* dst, hipStream_t s);
    // one forward step in dual numbers (plan with V = 1)
    hipError_t (*dual_step)(const StepArgs& a, const StepPlan& p, hipStream_t s, EventLog* log);
    size_t calc_bytes;  // bytes of the mode's arithmetic type (mu buffers)
};

const AdjOps* adj_ops(int q, Mode m);

}  // namespace lbm_amd
