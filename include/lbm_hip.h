/*
 * lbm_hip.h -- C ABI of the MI355X collide-stream time step.
 *
 * One handle (lbm_domain) is one single-device domain: the ghosted lattice of
 * the reference's Domain<M> (include/domain.h:10-56) with its boundary
 * handlers, held in HBM of one GPU and advanced by hand-written HIP kernels.
 * Every entry point returns 0 on success or a negative LBM_ERR_* code; the
 * message of the last failure on the calling thread is lbm_last_error().
 * Calls on one handle must come from one host thread at a time. Work is
 * enqueued on the handle's own HIP stream; functions that return data to the
 * host synchronise that stream. No pointer passed in is retained after a call
 * returns.
 *
 * Coordinates are the reference's ghosted coordinates: interior cells are
 * 1..xl, 1..yl, 1..zl; 0 and l+1 are the ghost layer (include/domain.hpp:67-70).
 */
#ifndef LBM_HIP_H
#define LBM_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define LBM_HIP_ABI_VERSION 5

/* Status codes. */
#define LBM_OK 0
#define LBM_ERR_ARG (-1)         /* invalid argument */
#define LBM_ERR_STATE (-2)       /* call not valid in the handle's state */
#define LBM_ERR_HIP (-3)         /* HIP runtime error */
#define LBM_ERR_ALLOC (-4)       /* device or host allocation failed */
#define LBM_ERR_UNSUPPORTED (-5) /* combination not implemented */
#define LBM_ERR_GEOMETRY (-6)    /* scenario geometry rejected at finalize */

/* Storage of the distribution fields.
 * F64  : the reference's fp64 pdfs and operation order (parity mode).
 * F32, F16, BF16 : shifted values f - w stored in that format, fp32 math.
 * F64S : shifted values in fp64 with fp64 math (reference for the others). */
enum lbm_storage {
    LBM_STORAGE_F64 = 0,
    LBM_STORAGE_F32 = 1,
    LBM_STORAGE_F16 = 2,
    LBM_STORAGE_BF16 = 3,
    LBM_STORAGE_F64S = 4
};

/* Boundary conditions: the reference's NonFluidCollision subclasses
 * (include/boundary.h:11-70). params[4] = {vx, vy, vz, rho}:
 *   NOSLIP      -                          NoSlipBoundary(domain)
 *   MOVINGWALL  wall velocity vx,vy,vz      MovingWallBoundary(domain, u_w)
 *   FREESLIP    -                          FreeSlipBoundary(domain)
 *   OUTFLOW     rho = reference density     OutflowBoundary(domain, rho_ref)
 *   INFLOW      vx,vy,vz, rho = ref density InflowBoundary(domain, u_in, rho_ref)
 *   PRESSURE    rho = input density         PressureBoundary(domain, rho_in) */
enum lbm_bc {
    LBM_BC_NOSLIP = 1,
    LBM_BC_MOVINGWALL = 2,
    LBM_BC_FREESLIP = 3,
    LBM_BC_OUTFLOW = 4,
    LBM_BC_INFLOW = 5,
    LBM_BC_PRESSURE = 6
};

/* Distribution storage: two buffers stepped A -> B -> A (the reference's
 * two fields), or one buffer updated in place with the AA pattern (half the
 * memory; NoSlip, MovingWall, Inflow, Outflow and Pressure boundaries; no
 * adjoint or batching). Both give the same bits in fp64 and fp32 storage. */
enum lbm_layout {
    LBM_LAYOUT_AB = 0,
    LBM_LAYOUT_AA = 1
};

typedef struct lbm_desc {
    int q;             /* 15, 19 or 27: lbm::model::d3q15/d3q19/d3q27 (include/model.h) */
    int storage;       /* enum lbm_storage */
    int device;        /* HIP device ordinal */
    int periodic[3];   /* extension: periodic x, y, z (the reference has none) */
    uint64_t xl, yl, zl; /* interior lengths, Domain(xl, yl, zl, ...) */
    double tau;        /* BGKCollision(tau), include/collision.hpp:55-58 */
    double origin[3];  /* Domain origin (VTK output only) */
    double spacing[3]; /* Domain spacing (VTK output only) */
    int layout;        /* enum lbm_layout */
} lbm_desc;

typedef struct lbm_domain lbm_domain;

/* Message of the last failed call on this thread ("" if none). */
const char* lbm_last_error(void);
int lbm_abi_version(void);

/* Lattice tables of lbm::model::d3q15/d3q19/d3q27 (include/model.h:13-139)
 * and lbm::C_S (include/lbmdefinitions.h:47) as this library uses them:
 * c has 3*q ints (x, y, z per direction), w has q doubles; any output may be
 * NULL. Needs no GPU. */
int lbm_lattice(int q, int* c, double* w, double* cs);

/* Fills defaults: q=19, F64, device 0, no periodic axes, tau=1, origin 0,
 * spacing 1, lengths 0. */
int lbm_desc_init(lbm_desc* desc);

/* Replaces Domain<M>(xl, yl, zl, collision, origin, spacing)
 * (include/domain.hpp:87-98) plus BGKCollision<M>(tau) (src/main.cpp:36):
 * every cell of the ghosted box starts as a BGK fluid cell holding the
 * lattice weights (include/cell.hpp:9-15). */
int lbm_domain_create(const lbm_desc* desc, lbm_domain** out);
int lbm_domain_destroy(lbm_domain* d);

/* Replaces Domain::setBoundaryCondition(BoundaryKeeper::get_collision<K>(...),
 * x0, xE, y0, yE, z0, zE) (include/domain.hpp:175-194,
 * include/io/scenario.h:99-127): inclusive box in ghosted coordinates, later
 * calls override earlier ones cell by cell. */
int lbm_set_boundary_box(lbm_domain* d, int kind, const double params[4], uint64_t x0,
                         uint64_t xe, uint64_t y0, uint64_t ye, uint64_t z0, uint64_t ze);

/* Replaces the mask part of io::read_vtk_point_file (include/io/vtk.hpp:137-149):
 * fluid_mask has xl*yl*zl bytes, x fastest; interior cells with 0 become
 * NoSlip cells. */
int lbm_set_solid_mask(lbm_domain* d, const uint8_t* fluid_mask);

/* Extension: Brinkman porosity alpha in [0,1] per interior cell (xl*yl*zl,
 * x fastest, NULL removes it) and body acceleration g[3]. The equilibrium
 * velocity becomes (1 - alpha) * (u + tau * g). Shifted storage modes only. */
int lbm_set_brinkman(lbm_domain* d, const double* alpha, const double g[3]);

/* Replaces Domain::set_nonfluid_cells_nullcollide (include/domain.hpp:101-113,
 * called at src/main.cpp:44): freezes the geometry, builds the device cell
 * kinds and the near-boundary cell list. Rejects (LBM_ERR_GEOMETRY) a domain
 * in which an interior fluid cell would pull from a ghost cell that no
 * boundary covers. Must precede lbm_step. */
int lbm_finalize(lbm_domain* d);

/* Replaces nsteps iterations of { stream(); swap(); collide(); }
 * (src/main.cpp:50-52, include/domain.hpp:116-172). Asynchronous. */
int lbm_step(lbm_domain* d, uint64_t nsteps);
/* Waits for the handle's stream. */
int lbm_sync(lbm_domain* d);

/* Replaces reads of cell(x,y,z)[q] over the whole ghosted collide_field
 * (include/cell.hpp:24-33): out has (xl+2)(yl+2)(zl+2)*q doubles in the
 * reference order idx(x,y,z)*q + q. Non-fluid cells hold the values the
 * reference's boundary pass leaves there. */
int lbm_get_pdf(lbm_domain* d, double* out);
/* Replaces writes of cell(x,y,z)[q] over the whole ghosted collide_field;
 * the next step pulls the written values from non-fluid cells as stored. */
int lbm_set_pdf(lbm_domain* d, const double* in);
/* Extension: every interior fluid cell <- feq(rho, u) from x-fastest
 * interior arrays rho[xl*yl*zl], u[3*xl*yl*zl] (host pointers). */
int lbm_init_equilibrium(lbm_domain* d, const double* rho, const double* u);

/* Replaces Cell::density() / Cell::velocity(density) of every interior cell
 * as the VTK writer evaluates them (include/io/vtk.hpp:62-73): rho has
 * xl*yl*zl doubles, u 3*xl*yl*zl, x fastest. Either may be NULL. */
int lbm_moments(lbm_domain* d, double* rho, double* u);
/* lbm_moments for callers that cannot allocate device memory (code under
 * graph capture, tight serving loops, harnesses that forbid hipMalloc): the
 * same outputs, bit for bit, with the device staging of the moments in a
 * fixed per-GPU array of 2^20 cells (32 MiB) instead of a per-call
 * allocation. Domains with more interior cells are rejected
 * (LBM_ERR_UNSUPPORTED). Calls on one GPU share that array, so calls that
 * overlap on one GPU race. No counterpart in the reference. */
int lbm_moments_noalloc(lbm_domain* d, double* rho, double* u);
/* Extension: fp64 sums over interior fluid cells of rho and of the momentum
 * (x, y, z), reduced on the device. */
int lbm_conservation(lbm_domain* d, double out[4]);

/* Timing: runs nsteps steps with a HIP event on the handle's stream before
 * the first launch and after every launch. out[0] = total ms, out[1] = ms in
 * bulk launches (row, cell-tile, in-place or two-step kernels), out[2] = ms
 * in near-boundary launches, out[3] = number of steps. A launch's time runs
 * from the previous event, so it includes the gap before the launch. */
int lbm_time_steps(lbm_domain* d, uint64_t nsteps, double out[4]);

/* Extension: bandwidth of a plain device copy (16-byte vector loads and
 * stores) of `bytes` bytes on `device`, repeated `reps` times after one
 * warm-up copy. out[0] = ms per copy, out[1] = GB/s counting read + write.
 * Used to calibrate HBM traffic counters and as the achievable-bandwidth
 * reference next to the roofline. */
int lbm_copy_bandwidth(int device, uint64_t bytes, int reps, double out[2]);
/* Extension: bandwidth of a q-plane fp32 copy (q <= 27 planes of plane_elems
 * elements, even; each thread loads its two elements of every plane, then
 * stores them non-temporally): the q read + q write streams of one step
 * without its arithmetic. out as lbm_copy_bandwidth. */
int lbm_copy_planes_bandwidth(int device, int q, uint64_t plane_elems, int reps, double out[2]);

/* ---------------------------------------------------------------------------
 * Differentiation (extension; the reference has none). A tape records one
 * forward run of `nsteps` steps from the domain's state at lbm_adj_create
 * and returns the objective; lbm_adj_backward runs the hand-written adjoint
 * step backwards through it (sqrt(N) checkpoints of the distribution field,
 * segments recomputed on the device) and returns fp64 gradients with respect
 * to the Brinkman porosity of every interior cell and to the velocity of the
 * moving walls (summed over all MovingWall conditions). The pair has the
 * shape of a torch.autograd.Function: forward(params) -> J,
 * backward(dJ) -> dJ * grad. Shifted storage modes (F32, F64S), D3Q19, with
 * NoSlip, MovingWall and Inflow boundaries. The domain must be finalized.
 */
enum lbm_objective {
    LBM_OBJ_PLANE_VELOCITY_MISFIT = 1, /* sum over cells with x == plane_x of |u - u_target|^2 */
    LBM_OBJ_KINETIC_ENERGY = 2         /* 1/2 sum over fluid cells of |u|^2 */
};

typedef struct lbm_adj_desc {
    int objective;          /* enum lbm_objective */
    int plane_x;            /* interior x of the misfit plane (1..xl) */
    const double* target_u; /* yl*zl*3 target velocities of the plane, y fastest (copied) */
    uint64_t nsteps;        /* steps of the forward run */
    uint64_t checkpoint_every; /* 0: ceil(sqrt(nsteps)) */
} lbm_adj_desc;

typedef struct lbm_tape lbm_tape;

int lbm_adj_create(lbm_domain* d, const lbm_adj_desc* desc, lbm_tape** out);
int lbm_adj_destroy(lbm_tape* t);
/* Forward: alpha (xl*yl*zl, x fastest) replaces the domain's porosity when
 * not NULL (the domain needs a Brinkman field then); J receives the objective.
 * Leaves the final state in the domain. */
int lbm_adj_forward(lbm_tape* t, const double* alpha, double* J);
/* Backward after a forward: grad_alpha (xl*yl*zl, or NULL) and grad_wall[3]
 * (or NULL) receive dJ times the gradients. */
int lbm_adj_backward(lbm_tape* t, double dJ, double* grad_alpha, double grad_wall[3]);
/* lbm_adj_forward / lbm_adj_backward with the porosity and its gradient in
 * device memory of the domain's GPU (xl*yl*zl doubles, x fastest), for
 * callers whose fields already live there (torch tensors): no host copy.
 * The caller has finished writing d_alpha before the call; the call returns
 * with d_grad_alpha written. No reference counterpart. */
int lbm_adj_forward_device(lbm_tape* t, const double* d_alpha, double* J);
int lbm_adj_backward_device(lbm_tape* t, double dJ, double* d_grad_alpha, double grad_wall[3]);
/* fp64 forward-mode directional derivative along dalpha (dual numbers through
 * the same step kernels): J and dJ/ds of J(alpha + s dalpha) at s = 0. */
int lbm_jvp(lbm_tape* t, const double* alpha, const double* dalpha, double* J, double* dJ);
/* Timing of the last backward: out[0] ms in recompute steps, out[1] ms in
 * adjoint steps, out[2] forward ms of the last lbm_adj_forward, out[3] steps. */
int lbm_adj_timing(lbm_tape* t, double out[4]);

/* ---------------------------------------------------------------------------
 * Batching (extension; the reference runs one domain per process). A batch
 * steps many finalized domains of any sizes and relaxation times together:
 * one persistent launch walks a tile work list over all domains and steps
 * their bulk and near-boundary cells (or, as the tuner decides, a second
 * launch covers the near-boundary cells; see lbm_batch_kernel);
 * lbm_batch_step captures the whole
 * run of nsteps as a hipGraph and launches it once. Results are bit-identical
 * to stepping each domain alone. All domains share the lattice and storage
 * (D3Q19; F32, F16, BF16 or F64S), the device, and their step parity; the
 * domains stay owned by the caller and must outlive the batch. Fields and
 * moments are read through the individual domain handles after lbm_batch_sync.
 */
typedef struct lbm_batch lbm_batch;

int lbm_batch_create(lbm_domain* const* domains, int n, lbm_batch** out);
int lbm_batch_destroy(lbm_batch* b);
/* Autotuner: times `trial_steps` steps of every candidate launch shape
 * (kernel form, cells per thread, threads along x per tile, persistent
 * blocks per CU) and
 * keeps the fastest; best_ms receives its time per step. */
int lbm_batch_tune(lbm_batch* b, int trial_steps, double* best_ms);
int lbm_batch_step(lbm_batch* b, uint64_t nsteps);
int lbm_batch_sync(lbm_batch* b);
/* Times one graph launch of nsteps with HIP events: out[0] = ms,
 * out[1] = interior cells of all domains (for MLUPS). */
int lbm_batch_time_steps(lbm_batch* b, uint64_t nsteps, double out[2]);

typedef struct lbm_batch_info {
    uint64_t ndomains;
    uint64_t ncells;   /* interior cells of all domains */
    uint64_t ntiles;
    uint64_t nnear;
    int vec;           /* cells per thread */
    int blocks_per_cu; /* persistent blocks per CU (0: one block per tile) */
    int bx_policy;     /* 0: smallest power of two covering a row, 1: half of it */
    int ncu;
    int kernel;        /* LBM_BATCH_ROWS_SEPARATE, LBM_BATCH_ROWS_FUSED or LBM_BATCH_CELLS */
} lbm_batch_info;
int lbm_batch_get_info(lbm_batch* b, lbm_batch_info* info);
/* Divergence flags (extension; the reference has one domain and no such
 * check): nonfinite (or NULL) receives one int per domain, in creation
 * order, 1 if a batched step since creation or the last reset read a NaN or
 * infinite value in one of that domain's fluid cells (the collision's
 * density and velocity of the incoming values are tested), else 0. Flags are
 * set inside the step kernels and only ever raised there; reset != 0 clears
 * them after they are read. Waits for the batch's stream. */
int lbm_batch_status(lbm_batch* b, int* nonfinite, int reset);
/* Kernel forms of a batched step. ROWS_*: blocks of threads over (x, y) rows
 * of one z plane, V cells per thread, near-boundary cells in a second launch
 * over their list (SEPARATE) or in the same pass (FUSED). CELLS: 64
 * consecutive cells of a z plane per wavefront, one cell per lane, every cell
 * kind in one pass. */
enum lbm_batch_kernel {
    LBM_BATCH_ROWS_SEPARATE = 0,
    LBM_BATCH_ROWS_FUSED = 1,
    LBM_BATCH_CELLS = 2
};
/* Sets a launch configuration directly (what lbm_batch_tune chooses):
 * kernel form, cells per thread (ROWS_*: one the storage mode supports),
 * bx_policy 0 or 1 (ROWS_*), persistent blocks per CU (0: one block per
 * tile or per four cell tiles). */
int lbm_batch_set_config(lbm_batch* b, int kernel, int vec, int bx_policy, int blocks_per_cu);

/* Batched adjoint (extension): every domain of a batch differentiated in
 * lockstep, objective J_d = kinetic energy of domain d after nsteps steps
 * from the domains' state at lbm_batch_adj_create, gradient of J_d with
 * respect to domain d's moving-wall velocity (summed over its MovingWall
 * conditions). Checkpoints every checkpoint_every steps (k); each domain
 * holds N/k + k + 1 field-sized slots on the device. Backward recomputes
 * each segment and steps the adjoint of all domains with one launch per
 * step. checkpoint_every = 0: k = ceil(sqrt(N)) where those slots fit in
 * free device memory, else two levels as lbm_batch_adj_create2 with
 * k = ceil(N^(2/3)) and sub_every = ceil(sqrt(k)). D3Q19, F32 or F64S,
 * NoSlip / MovingWall / Inflow. */
typedef struct lbm_batch_tape lbm_batch_tape;
int lbm_batch_adj_create(lbm_batch* b, uint64_t nsteps, uint64_t checkpoint_every, lbm_batch_tape** out);
/* Two checkpoint levels: checkpoints every checkpoint_every steps (k) in the
 * forward; the backward recomputes each segment once to lay down a
 * checkpoint every sub_every steps (k2) inside it and differentiates the
 * sub-segments from those. Slots per domain: N/k + k/k2 + k2 + 3, at the
 * cost of one more recompute of every step. Gradients are the same bits as
 * with one level. */
int lbm_batch_adj_create2(lbm_batch* b, uint64_t nsteps, uint64_t checkpoint_every, uint64_t sub_every,
                          lbm_batch_tape** out);
/* out[0] = k, out[1] = k2 (0: one level), out[2] = slots per domain. */
int lbm_batch_adj_levels(lbm_batch_tape* t, uint64_t out[3]);
int lbm_batch_adj_destroy(lbm_batch_tape* t);
/* J receives one objective value per domain (or NULL); leaves the final state
 * in the domains. */
int lbm_batch_adj_forward(lbm_batch_tape* t, double* J);
/* grad_wall receives 3 values per domain (x, y, z of dJ_d/du_wall). */
int lbm_batch_adj_backward(lbm_batch_tape* t, double* grad_wall);
/* out[0] ms in recompute launches, out[1] ms in adjoint launches, out[2]
 * forward ms, out[3] steps. */
int lbm_batch_adj_timing(lbm_batch_tape* t, double out[4]);

/* Layout and size information. */
typedef struct lbm_info {
    uint64_t ncells_interior;  /* xl*yl*zl */
    uint64_t ncells_fluid;     /* interior fluid cells */
    uint64_t ncells_near;      /* fluid cells with a non-fluid source */
    uint64_t bytes_per_value;  /* storage element size */
    uint64_t field_bytes;      /* bytes of one distribution buffer */
    uint64_t device_bytes;     /* all device memory of the handle */
    int vec;                   /* cells per thread of the bulk kernel */
    int block_x, block_y;      /* bulk kernel block shape */
    int nconds;                /* distinct boundary conditions */
    int cell_tiles;            /* 1: steps run as 64-cell tiles per wavefront (k_cells) */
    int layout;                /* enum lbm_layout */
    int steps_per_pass;        /* 2: lbm_step runs pairs of steps as one pass (k_bulk2) */
    int z_chunk;               /* planes per block of the two-step pass */
} lbm_info;
int lbm_get_info(lbm_domain* d, lbm_info* info);

/* Tuning. vec = 0 (default): domains with boundary cells step as 64-cell
 * tiles per wavefront, every cell kind in one kernel; bulk-only domains use
 * the row kernel with the mode's default cells per thread. vec > 0: the row
 * kernel with vec cells per thread (one the mode supports) plus a kernel over
 * the near-boundary cells. block_x: threads along x per row-kernel block (a
 * divisor of 256; 0 = automatic). Results are the same bits either way. */
int lbm_set_tuning(lbm_domain* d, int vec, int block_x);

/* Index arithmetic of the cell-tile and in-place kernels: 32 (default) uses
 * a uniform per-direction base plus a 32-bit byte offset wherever a
 * direction plane holds less than 4 GiB; 64 forces 64-bit element indices.
 * Results are the same bits either way (the switch exists to test both
 * builds). No reference counterpart. */
int lbm_set_index_width(lbm_domain* d, int bits);

/* Time steps per pass over the field. With planes under 4 GiB, lbm_step
 * runs two steps in one pass: the field is read once and written once for
 * both, with the intermediate state held on chip. This applies to fp32,
 * fp16 and bf16 storage (16-bit modes: x rows longer than 32 cells), on
 * bulk-only domains (periodic along all three axes) and on domains with
 * NoSlip, MovingWall, Inflow, Outflow and Pressure boundaries and solid
 * cells, where the first step after initialisation or lbm_set_pdf stays a
 * single step. A domain with a FreeSlip condition, a Brinkman field, the
 * in-place layout or a vector width set by lbm_set_tuning steps singly.
 * Results on fluid cells are the same bits as single steps. steps_per_pass:
 * 0 (default, automatic): two steps per pass where that measured faster,
 * which leaves domains with boundary cells in 16-bit storage stepping
 * singly; 2: two steps per pass wherever it applies; 1: single steps. z_chunk: planes per
 * block of the pass (0 = automatic). No reference counterpart: a scheduling
 * choice of this implementation. */
int lbm_set_time_blocking(lbm_domain* d, int steps_per_pass, int z_chunk);

#ifdef __cplusplus
}
#endif
#endif /* LBM_HIP_H */
