// C ABI of batched stepping (include/lbm_hip.h, lbm_batch_*).
//
// The batch borrows finalized domains: their buffers, cell kinds, near lists
// and boundary tables are described to the device by one BatchView each. A
// launch configuration (cells per thread V, threads along x per tile, blocks
// per CU of the persistent grid) determines the tile list; lbm_batch_tune
// times the candidates on the real batch and keeps the fastest. A run of
// nsteps is captured once into a hipGraph (2 kernel nodes per step, no
// memset nodes) and replayed as one launch.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <string>
#include <vector>

#include "batch.hpp"
#include "batch_adjoint.hpp"
#include "domain_impl.hpp"
#include "lbm_hip.h"
#include "ops_adjoint.hpp"

using namespace lbm_amd;

struct lbm_batch {
    std::vector<lbm_domain*> doms;
    const BatchOps* ops = nullptr;
    int device = 0;
    int ncu = 0;
    hipStream_t stream = nullptr;
    int V = 0, bpc = 8, bx_policy = 0;
    int kernel = BK_CELLS;      // enum BatchKernel
    bool any_compact = false;   // some domain packs its rows: row kernels take V = 1 only
    bool any_freeslip = false;  // some domain has a FreeSlip condition
    bool all_idx32 = true;      // every domain's plane is under 4 GiB
    int64_t near_cells = 0;     // NEAR cells of all domains
    BatchView* d_views = nullptr;
    TileRef* d_tiles = nullptr;
    NearRef* d_near = nullptr;
    int* d_status = nullptr;    // per-domain non-finite flags (BatchArgs::status)
    int64_t ntiles = 0, nnear = 0;
    uint64_t ncells = 0;
    hipGraphExec_t exec = nullptr;
    hipGraph_t graph = nullptr;
    uint64_t exec_steps = 0;
    int exec_parity = -1, exec_mat = -1;
};

namespace {

void drop_graph(lbm_batch* b) {
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    b->exec = nullptr;
    b->graph = nullptr;
    b->exec_steps = 0;
}

void free_lists(lbm_batch* b) {
    (void)hipFree(b->d_views);
    (void)hipFree(b->d_tiles);
    (void)hipFree(b->d_near);
    b->d_views = nullptr;
    b->d_tiles = nullptr;
    b->d_near = nullptr;
}

// Builds views, tiles and (for BK_ROWS_SEPARATE) the near list and uploads
// them. Row tiles depend on (V, bx_policy); with the NEAR cells fused into
// the tile kernel every domain gets tiles over all its rows. Cell tiles cover
// every plane of every domain in 64-cell pieces.
int configure(lbm_batch* b, int V, int bx_policy, int bpc, int kernel) {
    if (kernel == BK_CELLS) V = 1;
    if (V > 1)
        for (const lbm_domain* d : b->doms)
            if (d->compact) return fail(LBM_ERR_ARG, "a domain with packed rows takes 1 cell per thread");
    drop_graph(b);
    std::vector<BatchView> views;
    std::vector<TileRef> tiles;
    std::vector<NearRef> near;
    for (size_t di = 0; di < b->doms.size(); ++di) {
        const lbm_domain* d = b->doms[di];
        BatchView v{};
        v.L = d->L;
        v.buf[0] = d->buf[0];
        v.buf[1] = d->buf[1];
        v.kind = d->d_kind;
        v.lmask = d->d_lmask;
        v.lcond = d->d_lcond;
        v.conds = d->d_conds;
        v.tau = d->desc.tau;
        v.omega = 1.0 / d->desc.tau;
        const int nvx = (d->L.xl + V - 1) / V;
        int bx = 1, l2 = 0;
        while (bx < nvx && bx < 256) {
            bx *= 2;
            ++l2;
        }
        if (bx_policy == 1 && l2 > 0) {
            bx /= 2;
            --l2;
        }
        v.bx_log2 = l2;
        v.nbx = (nvx + bx - 1) / bx;
        v.nby = (d->L.yl + (256 / bx) - 1) / (256 / bx);
        v.tpp = static_cast<int>((static_cast<int64_t>(d->L.xl) * d->L.yl + 63) / 64);
        v.rxl = 1.0f / static_cast<float>(d->L.xl);
        const int64_t nt = kernel == BK_CELLS ? static_cast<int64_t>(v.tpp) * d->L.zl
                                              : static_cast<int64_t>(v.nbx) * v.nby * d->L.zl;
        // cell tiles split a plane index with a float reciprocal (exact below 2^22)
        if (nt > 0x3fffffff || static_cast<int64_t>(d->L.xl) * d->L.yl > (int64_t{1} << 22))
            return fail(LBM_ERR_UNSUPPORTED, "domain too large for a batch tile index");
        if (kernel != BK_ROWS_SEPARATE || d->plan.has_bulk)
            for (int64_t t = 0; t < nt; ++t) tiles.push_back(TileRef{static_cast<int>(di), static_cast<int>(t)});
        views.push_back(v);
    }
    // near cells: the domains' own lists, gathered
    b->near_cells = 0;
    for (size_t di = 0; di < b->doms.size(); ++di) {
        const lbm_domain* d = b->doms[di];
        b->near_cells += d->nb;
        if (!d->nb || kernel != BK_ROWS_SEPARATE) continue;
        std::vector<CellCoord> bl(static_cast<size_t>(d->nb));
        HIPCHK(hipMemcpy(bl.data(), d->d_blist, bl.size() * sizeof(CellCoord), hipMemcpyDeviceToHost));
        for (const CellCoord& c : bl) near.push_back(NearRef{static_cast<int>(di), c.x, c.y, c.z});
    }
    free_lists(b);
    HIPCHK(hipMalloc(&b->d_views, views.size() * sizeof(BatchView)));
    HIPCHK(hipMemcpy(b->d_views, views.data(), views.size() * sizeof(BatchView), hipMemcpyHostToDevice));
    if (!tiles.empty()) {
        HIPCHK(hipMalloc(&b->d_tiles, tiles.size() * sizeof(TileRef)));
        HIPCHK(hipMemcpy(b->d_tiles, tiles.data(), tiles.size() * sizeof(TileRef), hipMemcpyHostToDevice));
    }
    if (!near.empty()) {
        HIPCHK(hipMalloc(&b->d_near, near.size() * sizeof(NearRef)));
        HIPCHK(hipMemcpy(b->d_near, near.data(), near.size() * sizeof(NearRef), hipMemcpyHostToDevice));
    }
    if (tiles.size() > 0x3fffffffu) return fail(LBM_ERR_UNSUPPORTED, "too many batch tiles");
    b->ntiles = static_cast<int64_t>(tiles.size());
    b->nnear = static_cast<int64_t>(near.size());
    b->V = V;
    b->bx_policy = bx_policy;
    b->bpc = bpc;
    b->kernel = kernel;
    return LBM_OK;
}

int check_parity(lbm_batch* b, int& parity, int& mat) {
    parity = b->doms[0]->cur;
    mat = b->doms[0]->materialised ? 1 : 0;
    for (const lbm_domain* d : b->doms)
        if (d->cur != parity || (d->materialised ? 1 : 0) != mat)
            return fail(LBM_ERR_STATE, "batched domains must be at the same step parity");
    return LBM_OK;
}

// Enqueues nsteps on the batch stream (used both eagerly and under capture).
hipError_t enqueue(lbm_batch* b, uint64_t nsteps, int parity, int mat) {
    BatchArgs a{};
    a.views = b->d_views;
    a.tiles = b->d_tiles;
    a.ntiles = b->ntiles;
    a.near = b->d_near;
    a.nnear = b->nnear;
    a.freeslip = b->any_freeslip ? 1 : 0;
    a.idx32 = b->all_idx32 ? 1 : 0;
    // the forward step keeps the plain walk: grouping tiles by L2 domain
    // measured slower there (512 cavities of 16-96 cells a side: 2.43
    // against 2.10 ms per step), while it pays in the adjoint's kernels
    a.nxcd = 1;
    a.status = b->d_status;
    // blocks that give every tile its own block (cell tiles: 4 wavefronts a block)
    const int64_t full = std::max<int64_t>(1, b->kernel == BK_CELLS ? (b->ntiles + 3) / 4 : b->ntiles);
    const int grid_bulk = static_cast<int>(b->bpc ? std::min<int64_t>(full, static_cast<int64_t>(b->ncu) * b->bpc)
                                                  : std::min<int64_t>(full, 0x7fffffff));
    const int grid_near = static_cast<int>(
        std::max<int64_t>(1, std::min<int64_t>((b->nnear + 127) / 128, static_cast<int64_t>(b->ncu) * 16)));
    for (uint64_t s = 0; s < nsteps; ++s) {
        a.parity = parity;
        a.materialised = (s == 0) ? mat : 0;
        const hipError_t e = b->ops->launch(a, b->V, b->kernel, grid_bulk, grid_near, b->stream);
        if (e != hipSuccess) return e;
        parity ^= 1;
    }
    return hipSuccess;
}

void advance_domains(lbm_batch* b, uint64_t nsteps) {
    for (lbm_domain* d : b->doms) {
        d->cur ^= static_cast<int>(nsteps & 1);
        if (nsteps) d->materialised = false;
    }
}

int sync_domains(lbm_batch* b) {
    for (lbm_domain* d : b->doms) HIPCHK(hipStreamSynchronize(d->stream));
    return LBM_OK;
}

}  // namespace

extern "C" {

int lbm_batch_create(lbm_domain* const* domains, int n, lbm_batch** out) {
    if (!domains || n < 1 || !out) return fail(LBM_ERR_ARG, "need at least one domain");
    *out = nullptr;
    lbm_domain* d0 = domains[0];
    if (int rc = check_handle(d0)) return rc;
    for (int i = 0; i < n; ++i) {
        const lbm_domain* d = domains[i];
        if (!d) return fail(LBM_ERR_ARG, "null domain in batch");
        if (!d->finalized) return fail(LBM_ERR_STATE, "batched domains must be finalized");
        if (d->Q != d0->Q || d->mode != d0->mode || d->desc.device != d0->desc.device)
            return fail(LBM_ERR_ARG, "batched domains must share lattice, storage and device");
        if (d->d_alpha) return fail(LBM_ERR_UNSUPPORTED, "batched Brinkman domains are not supported");
        if (d->aa) return fail(LBM_ERR_UNSUPPORTED, "batching needs the two-buffer layout");
    }
    const BatchOps* ops = batch_ops(d0->Q, d0->mode);
    if (!ops) return fail(LBM_ERR_UNSUPPORTED, "batching needs D3Q19 with F32, F16, BF16 or F64S storage");
    lbm_batch* b = new lbm_batch();
    b->doms.assign(domains, domains + n);
    b->ops = ops;
    b->device = d0->desc.device;
    (void)hipDeviceGetAttribute(&b->ncu, hipDeviceAttributeMultiprocessorCount, b->device);
    if (b->ncu <= 0) b->ncu = 256;
    for (const lbm_domain* d : b->doms) {
        b->ncells += static_cast<uint64_t>(d->L.xl) * d->L.yl * d->L.zl;
        b->any_compact |= d->compact;
        b->any_freeslip |= d->plan.freeslip;
        b->all_idx32 &= d->plan.idx32;
    }
    if (hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess) {
        delete b;
        return fail(LBM_ERR_HIP, "hipStreamCreate failed");
    }
    if (hipMalloc(&b->d_status, sizeof(int) * static_cast<size_t>(n)) != hipSuccess ||
        hipMemset(b->d_status, 0, sizeof(int) * static_cast<size_t>(n)) != hipSuccess) {
        (void)hipFree(b->d_status);
        (void)hipStreamDestroy(b->stream);
        delete b;
        return fail(LBM_ERR_ALLOC, "batch status flags");
    }
    // start from the domains' own cells-per-thread setting
    int V0 = ops->vecs[0];
    for (int i = 0; i < 4 && ops->vecs[i]; ++i)
        if (ops->vecs[i] == d0->plan.V) V0 = d0->plan.V;
    int rc = configure(b, V0, 0, 8, BK_CELLS);
    if (rc) {
        free_lists(b);
        (void)hipFree(b->d_status);
        (void)hipStreamDestroy(b->stream);
        delete b;
        return rc;
    }
    *out = b;
    return LBM_OK;
}

int lbm_batch_destroy(lbm_batch* b) {
    if (!b) return LBM_OK;
    (void)hipSetDevice(b->device);
    (void)hipStreamSynchronize(b->stream);
    drop_graph(b);
    free_lists(b);
    (void)hipFree(b->d_status);
    (void)hipStreamDestroy(b->stream);
    delete b;
    return LBM_OK;
}

}  // extern "C"

namespace {

// Captures (once per (nsteps, parity, materialised)) and instantiates the run.
int ensure_graph(lbm_batch* b, uint64_t nsteps, int parity, int mat) {
    if (!(b->exec && b->exec_steps == nsteps && b->exec_parity == parity && b->exec_mat == mat)) {
        drop_graph(b);
        HIPCHK(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
        const hipError_t e = enqueue(b, nsteps, parity, mat);
        hipGraph_t g = nullptr;
        const hipError_t e2 = hipStreamEndCapture(b->stream, &g);
        if (e != hipSuccess || e2 != hipSuccess) {
            if (g) (void)hipGraphDestroy(g);
            return fail(LBM_ERR_HIP, std::string("graph capture: ") + hipGetErrorString(e != hipSuccess ? e : e2));
        }
        b->graph = g;
        HIPCHK(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
        b->exec_steps = nsteps;
        b->exec_parity = parity;
        b->exec_mat = mat;
    }
    return LBM_OK;
}

}  // namespace

extern "C" {

int lbm_batch_step(lbm_batch* b, uint64_t nsteps) {
    if (!b) return fail(LBM_ERR_ARG, "null batch");
    HIPCHK(hipSetDevice(b->device));
    if (nsteps == 0) return LBM_OK;
    int parity, mat;
    if (int rc = check_parity(b, parity, mat)) return rc;
    if (int rc = sync_domains(b)) return rc;  // pending per-domain work (init, set_pdf)
    if (int rc = ensure_graph(b, nsteps, parity, mat)) return rc;
    HIPCHK(hipGraphLaunch(b->exec, b->stream));
    advance_domains(b, nsteps);
    return LBM_OK;
}

int lbm_batch_sync(lbm_batch* b) {
    if (!b) return fail(LBM_ERR_ARG, "null batch");
    HIPCHK(hipSetDevice(b->device));
    HIPCHK(hipStreamSynchronize(b->stream));
    return LBM_OK;
}

int lbm_batch_time_steps(lbm_batch* b, uint64_t nsteps, double out[2]) {
    if (!b || !out || nsteps == 0) return fail(LBM_ERR_ARG, "bad arguments");
    HIPCHK(hipSetDevice(b->device));
    Events ev(2);
    HIPCHK(ev.create());
    // build (capture + instantiate) outside the timed region, then time one launch
    int parity, mat;
    if (int rc = check_parity(b, parity, mat)) return rc;
    if (int rc = sync_domains(b)) return rc;
    if (int rc = ensure_graph(b, nsteps, parity, mat)) return rc;
    HIPCHK(hipEventRecord(ev[0], b->stream));
    HIPCHK(hipGraphLaunch(b->exec, b->stream));
    HIPCHK(hipEventRecord(ev[1], b->stream));
    advance_domains(b, nsteps);
    HIPCHK(hipEventSynchronize(ev[1]));
    float ms = 0;
    HIPCHK(hipEventElapsedTime(&ms, ev[0], ev[1]));
    out[0] = ms;
    out[1] = static_cast<double>(b->ncells);
    return LBM_OK;
}

int lbm_batch_tune(lbm_batch* b, int trial_steps, double* best_ms) {
    if (!b || trial_steps < 1) return fail(LBM_ERR_ARG, "bad arguments");
    HIPCHK(hipSetDevice(b->device));
    int parity, mat;
    if (int rc = check_parity(b, parity, mat)) return rc;
    if (int rc = sync_domains(b)) return rc;
    double best = 1e300;
    const int v0 = b->V, p0 = b->bx_policy, b0 = b->bpc, k0 = b->kernel;
    int bv = v0, bp = p0, bb = b0, bk = k0;
    // The trials step the real fields; save every domain's current field and
    // restore it afterwards, on every exit, so tuning leaves the state as it
    // found it. (Trial steps write only fluid cells, which the next step
    // overwrites in the other buffer.)
    const size_t nd = b->doms.size();
    std::vector<void*> saved(nd, nullptr);
    std::vector<int> saved_cur(nd);
    std::vector<bool> saved_mat(nd);
    for (size_t i = 0; i < nd; ++i) {
        saved_cur[i] = b->doms[i]->cur;
        saved_mat[i] = b->doms[i]->materialised;
    }
    // the trials may also raise non-finite flags: keep the ones the batch had
    std::vector<int> saved_status(nd);
    HIPCHK(hipMemcpy(saved_status.data(), b->d_status, sizeof(int) * nd, hipMemcpyDeviceToHost));
    Events ev(2);
    // Kernels are compiled without FMA contraction, so a cell's arithmetic is
    // the same for every V, tile shape and grid size: the search changes
    // speed only, never results.
    auto search = [&]() -> int {
        for (size_t i = 0; i < nd; ++i) {
            lbm_domain* d = b->doms[i];
            if (hipMalloc(&saved[i], d->field_bytes) != hipSuccess) {
                saved[i] = nullptr;
                return fail(LBM_ERR_ALLOC, "tuning needs a copy of every field");
            }
            HIPCHK(hipMemcpyAsync(saved[i], d->buf[d->cur], d->field_bytes, hipMemcpyDeviceToDevice, b->stream));
        }
        HIPCHK(ev.create());
        for (int kern : {BK_CELLS, BK_ROWS_FUSED, BK_ROWS_SEPARATE}) {
            for (int i = 0; i < 4 && b->ops->vecs[i]; ++i)
                for (int pol = 0; pol < 2; ++pol)
                    for (int bpc : {0, 4, 8, 16}) {
                        // cell tiles have one shape: try it once per grid size
                        if (kern == BK_CELLS && (i > 0 || pol > 0)) continue;
                        if (kern != BK_CELLS && b->ops->vecs[i] > 1 && b->any_compact) continue;
                        if (int rc = configure(b, b->ops->vecs[i], pol, bpc, kern)) return rc;
                        // trial runs advance the fields; the saved copies restore them
                        int p, m;
                        check_parity(b, p, m);
                        HIPCHK(enqueue(b, 1, p, m));  // warm-up
                        advance_domains(b, 1);
                        check_parity(b, p, m);
                        HIPCHK(hipEventRecord(ev[0], b->stream));
                        HIPCHK(enqueue(b, static_cast<uint64_t>(trial_steps), p, m));
                        HIPCHK(hipEventRecord(ev[1], b->stream));
                        HIPCHK(hipEventSynchronize(ev[1]));
                        advance_domains(b, static_cast<uint64_t>(trial_steps));
                        float ms = 0;
                        HIPCHK(hipEventElapsedTime(&ms, ev[0], ev[1]));
                        const double per = ms / trial_steps;
                        if (per < best) {
                            best = per;
                            bv = b->V;
                            bp = pol;
                            bb = bpc;
                            bk = kern;
                        }
                    }
        }
        return LBM_OK;
    };
    int rc = search();
    const std::string err = rc ? last_error() : std::string();
    // Restore every domain (the fields only if every copy was taken) and free
    // what the search took, whether it finished or not.
    bool all_saved = true;
    for (void* p : saved) all_saved &= (p != nullptr);
    for (size_t i = 0; i < nd; ++i) {
        lbm_domain* d = b->doms[i];
        d->cur = saved_cur[i];
        d->materialised = saved_mat[i];
        if (all_saved) {
            const hipError_t e = hipMemcpyAsync(d->buf[d->cur], saved[i], d->field_bytes, hipMemcpyDeviceToDevice,
                                                b->stream);
            if (e != hipSuccess && !rc) rc = fail(LBM_ERR_HIP, std::string("restoring a field: ") + hipGetErrorString(e));
        }
    }
    {
        const hipError_t e = hipStreamSynchronize(b->stream);
        if (e != hipSuccess && !rc) rc = fail(LBM_ERR_HIP, std::string("hipStreamSynchronize: ") + hipGetErrorString(e));
        const hipError_t e2 = hipMemcpy(b->d_status, saved_status.data(), sizeof(int) * nd, hipMemcpyHostToDevice);
        if (e2 != hipSuccess && !rc) rc = fail(LBM_ERR_HIP, std::string("restoring status flags: ") + hipGetErrorString(e2));
    }
    for (void* p : saved)
        if (p) (void)hipFree(p);
    if (rc) {
        // keep the configuration the batch had, and the first error's message
        (void)configure(b, v0, p0, b0, k0);
        return fail(rc, err.empty() ? last_error() : err);
    }
    if (int c = configure(b, bv, bp, bb, bk)) return c;
    if (best_ms) *best_ms = best;
    return LBM_OK;
}

int lbm_batch_set_config(lbm_batch* b, int kernel, int vec, int bx_policy, int blocks_per_cu) {
    if (!b) return fail(LBM_ERR_ARG, "null batch");
    bool ok = false;
    for (int i = 0; i < 4 && b->ops->vecs[i]; ++i) ok |= (b->ops->vecs[i] == vec);
    if (!ok && kernel != BK_CELLS) return fail(LBM_ERR_ARG, "cells per thread not built for this storage mode");
    if (bx_policy < 0 || bx_policy > 1 || blocks_per_cu < 0 || blocks_per_cu > 64 || kernel < BK_ROWS_SEPARATE ||
        kernel > BK_CELLS)
        return fail(LBM_ERR_ARG, "bad batch configuration");
    HIPCHK(hipSetDevice(b->device));
    HIPCHK(hipStreamSynchronize(b->stream));
    drop_graph(b);
    return configure(b, vec, bx_policy, blocks_per_cu, kernel);
}

int lbm_batch_status(lbm_batch* b, int* nonfinite, int reset) {
    if (!b) return fail(LBM_ERR_ARG, "null batch");
    HIPCHK(hipSetDevice(b->device));
    HIPCHK(hipStreamSynchronize(b->stream));
    const size_t n = b->doms.size();
    if (nonfinite) HIPCHK(hipMemcpy(nonfinite, b->d_status, sizeof(int) * n, hipMemcpyDeviceToHost));
    if (reset) {
        // outside any captured graph: the replayed steps only ever set flags
        HIPCHK(hipMemsetAsync(b->d_status, 0, sizeof(int) * n, b->stream));
        HIPCHK(hipStreamSynchronize(b->stream));
    }
    return LBM_OK;
}

int lbm_batch_get_info(lbm_batch* b, lbm_batch_info* info) {
    if (!b || !info) return fail(LBM_ERR_ARG, "null argument");
    info->ndomains = b->doms.size();
    info->ncells = b->ncells;
    info->ntiles = static_cast<uint64_t>(b->ntiles);
    info->nnear = static_cast<uint64_t>(b->near_cells);
    info->vec = b->V;
    info->blocks_per_cu = b->bpc;
    info->bx_policy = b->bx_policy;
    info->ncu = b->ncu;
    info->kernel = b->kernel;
    return LBM_OK;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Batched adjoint: all domains of a batch differentiated together, kinetic
// energy objective per domain, gradients with respect to each domain's
// moving-wall velocity. Forward: the batch's graph-captured steps with a
// checkpoint of every domain every k steps. Backward: per segment, one
// recompute launch and one adjoint launch per step for all domains.
//
// Two levels (k2 > 0): the forward keeps a checkpoint every k steps; the
// backward recomputes each such segment once to lay down a checkpoint
// every k2 steps inside it, then treats each sub-segment as above. Slots
// per domain: N/k + k/k2 + k2 + 3 instead of N/k + k + 1, for one more
// recompute of every step.

struct lbm_batch_tape {
    lbm_batch* b = nullptr;
    const BatchAdjOps* ops = nullptr;
    const AdjOps* aops = nullptr;
    uint64_t T = 0, k = 0, nck = 0;
    uint64_t k2 = 0, n2 = 0;  // second level: sub-segment length (0: one level) and sub-segments per segment
    int nslots = 0;
    std::vector<void*> chunk;
    double* d_gw = nullptr;
    std::vector<BatchAdjView> views;
    BatchAdjView* d_views = nullptr;
    TileRef* d_tiles = nullptr;
    int64_t ntiles = 0;
    double* d_out = nullptr;
    bool start_mat = false;
    bool forward_done = false;
    double timing[4] = {0, 0, 0, 0};
};

namespace {

void btape_free(lbm_batch_tape* t) {
    for (void* p : t->chunk) (void)hipFree(p);
    t->chunk.clear();
    (void)hipFree(t->d_gw);
    (void)hipFree(t->d_views);
    (void)hipFree(t->d_tiles);
    (void)hipFree(t->d_out);
    t->d_gw = nullptr;
    t->d_views = nullptr;
    t->d_tiles = nullptr;
    t->d_out = nullptr;
}

// Slots: checkpoints 0 .. nck-1; with two levels, sub-segment checkpoints
// j = 1 .. n2-1 of the segment being differentiated; states j = 1 .. kin-1
// of the (sub-)segment, kin = k2 or k; two mu buffers; with two levels, two
// scratch states for laying down the sub-segment checkpoints.
uint64_t inner_len(const lbm_batch_tape* t) { return t->k2 ? t->k2 : t->k; }
int sub_slot(const lbm_batch_tape* t, uint64_t j) { return static_cast<int>(t->nck + j - 1); }
int seg_slot(const lbm_batch_tape* t, uint64_t j) {
    return static_cast<int>(t->nck + (t->k2 ? t->n2 - 1 : 0) + j - 1);
}
int mu_slot(const lbm_batch_tape* t, int i) { return seg_slot(t, inner_len(t)) + i; }
int tmp_slot(const lbm_batch_tape* t, int i) { return mu_slot(t, 2) + i; }

int grid_for(const lbm_batch_tape* t) {
    return static_cast<int>(std::max<int64_t>(1, std::min<int64_t>((t->ntiles + 3) / 4, 0x7fffffff)));
}

}  // namespace

namespace {

int batch_adj_create(lbm_batch* b, uint64_t nsteps, uint64_t k1, uint64_t k2, bool automatic, lbm_batch_tape** out) {
    if (!b || !out) return fail(LBM_ERR_ARG, "null argument");
    *out = nullptr;
    HIPCHK(hipSetDevice(b->device));
    lbm_domain* d0 = b->doms[0];
    const BatchAdjOps* ops = batch_adj_ops(d0->Q, d0->mode);
    const AdjOps* aops = adj_ops(d0->Q, d0->mode);
    if (!ops || !aops) return fail(LBM_ERR_UNSUPPORTED, "batched adjoint needs D3Q19 with F32 or F64S storage");
    for (const lbm_domain* d : b->doms)
        for (const HostCond& c : d->conds)
            if (c.kind != LBM_BC_NOSLIP && c.kind != LBM_BC_MOVINGWALL && c.kind != LBM_BC_INFLOW)
                return fail(LBM_ERR_UNSUPPORTED, "adjoint supports NoSlip, MovingWall and Inflow boundaries");
    if (nsteps < 1 || nsteps > 10000000) return fail(LBM_ERR_ARG, "nsteps in [1, 1e7]");
    int parity, mat;
    if (int rc = check_parity(b, parity, mat)) return rc;
    if (int rc = sync_domains(b)) return rc;

    lbm_batch_tape* t = new lbm_batch_tape();
    t->b = b;
    t->ops = ops;
    t->aops = aops;
    t->T = nsteps;
    auto plan = [&](uint64_t ka, uint64_t kb) {
        t->k = std::max<uint64_t>(1, std::min(ka, t->T));
        t->nck = (t->T + t->k - 1) / t->k;
        t->k2 = kb ? std::max<uint64_t>(1, std::min(kb, t->k)) : 0;
        t->n2 = t->k2 ? (t->k + t->k2 - 1) / t->k2 : 0;
        const uint64_t n = t->k2 ? t->nck + (t->n2 - 1) + (t->k2 - 1) + 2 + 2 : t->nck + (t->k - 1) + 2;
        t->nslots = static_cast<int>(std::min<uint64_t>(n, 0x7fffffff));
        return n;
    };
    size_t field_total = 0;
    for (const lbm_domain* d : b->doms) field_total += d->field_bytes;
    if (automatic) {
        // sqrt(N) checkpoints where their slots fit in free memory (leaving a
        // tenth of it), else two levels of about N^(2/3) and N^(1/3) steps
        const uint64_t ks = static_cast<uint64_t>(std::ceil(std::sqrt(static_cast<double>(nsteps))));
        size_t free_b = 0, total_b = 0;
        HIPCHK(hipMemGetInfo(&free_b, &total_b));
        const uint64_t n1 = plan(ks, 0);
        if (static_cast<double>(n1) * static_cast<double>(field_total) > 0.9 * static_cast<double>(free_b)) {
            const uint64_t ka = static_cast<uint64_t>(std::ceil(std::pow(static_cast<double>(nsteps), 2.0 / 3.0)));
            plan(ka, static_cast<uint64_t>(std::ceil(std::sqrt(static_cast<double>(ka)))));
        }
    } else {
        plan(k1, k2);
    }
    t->start_mat = mat != 0;
    auto bail = [&](int code, const std::string& m) {
        btape_free(t);
        delete t;
        return fail(code, m);
    };
    if (t->nslots >= 0x7fffffff) return bail(LBM_ERR_ARG, "too many checkpoint slots");

    std::vector<TileRef> tiles;
    int64_t gw_elems = 0;
    std::vector<int64_t> gw_off;
    for (size_t di = 0; di < b->doms.size(); ++di) {
        lbm_domain* d = b->doms[di];
        void* c = nullptr;
        if (hipMalloc(&c, static_cast<size_t>(t->nslots) * d->field_bytes) != hipSuccess)
            return bail(LBM_ERR_ALLOC,
                        "checkpoint slots do not fit in device memory (lower nsteps, or use two checkpoint levels)");
        t->chunk.push_back(c);
        BatchAdjView v{};
        v.L = d->L;
        v.chunk = c;
        v.slot_elems = static_cast<int64_t>(d->field_bytes / d->esize);
        v.gfinal = d->buf[d->cur];
        v.kind = d->d_kind;
        v.lmask = d->d_lmask;
        v.lcond = d->d_lcond;
        v.conds = d->d_conds;
        v.tau = d->desc.tau;
        v.omega = 1.0 / d->desc.tau;
        v.tpp = static_cast<int>((static_cast<int64_t>(d->L.xl) * d->L.yl + 63) / 64);
        v.rxl = 1.0 / static_cast<double>(d->L.xl);
        t->views.push_back(v);
        gw_off.push_back(gw_elems);
        gw_elems += 3 * d->L.plane;
        const int64_t nt = static_cast<int64_t>(v.tpp) * d->L.zl;
        if (static_cast<int64_t>(tiles.size()) + nt > 0x3fffffff) return bail(LBM_ERR_UNSUPPORTED, "too many tiles");
        for (int64_t i = 0; i < nt; ++i) tiles.push_back(TileRef{static_cast<int>(di), static_cast<int>(i)});
        // checkpoint 0: the domain's current state
        if (hipMemcpyAsync(c, d->buf[d->cur], d->field_bytes, hipMemcpyDeviceToDevice, b->stream) != hipSuccess)
            return bail(LBM_ERR_HIP, "checkpoint copy failed");
    }
    if (hipMalloc(&t->d_gw, sizeof(double) * gw_elems) != hipSuccess) return bail(LBM_ERR_ALLOC, "gradient buffers");
    for (size_t di = 0; di < t->views.size(); ++di) t->views[di].gw = t->d_gw + gw_off[di];
    t->ntiles = static_cast<int64_t>(tiles.size());
    if (hipMalloc(&t->d_views, sizeof(BatchAdjView) * t->views.size()) != hipSuccess ||
        hipMalloc(&t->d_tiles, sizeof(TileRef) * tiles.size()) != hipSuccess ||
        hipMalloc(&t->d_out, sizeof(double) * 3 * t->views.size()) != hipSuccess)
        return bail(LBM_ERR_ALLOC, "tape tables");
    if (hipMemcpy(t->d_tiles, tiles.data(), sizeof(TileRef) * tiles.size(), hipMemcpyHostToDevice) != hipSuccess ||
        hipStreamSynchronize(b->stream) != hipSuccess)
        return bail(LBM_ERR_HIP, "tape setup");
    *out = t;
    return LBM_OK;
}

}  // namespace

extern "C" {

int lbm_batch_adj_create(lbm_batch* b, uint64_t nsteps, uint64_t checkpoint_every, lbm_batch_tape** out) {
    return batch_adj_create(b, nsteps, checkpoint_every, 0, checkpoint_every == 0, out);
}

int lbm_batch_adj_create2(lbm_batch* b, uint64_t nsteps, uint64_t checkpoint_every, uint64_t sub_every,
                          lbm_batch_tape** out) {
    if (checkpoint_every == 0 || sub_every == 0) return fail(LBM_ERR_ARG, "checkpoint_every and sub_every must be > 0");
    return batch_adj_create(b, nsteps, checkpoint_every, sub_every, false, out);
}

int lbm_batch_adj_levels(lbm_batch_tape* t, uint64_t out[3]) {
    if (!t || !out) return fail(LBM_ERR_ARG, "null argument");
    out[0] = t->k;
    out[1] = t->k2;
    out[2] = static_cast<uint64_t>(t->nslots);
    return LBM_OK;
}

int lbm_batch_adj_destroy(lbm_batch_tape* t) {
    if (!t) return LBM_OK;
    (void)hipSetDevice(t->b->device);
    (void)hipStreamSynchronize(t->b->stream);
    btape_free(t);
    delete t;
    return LBM_OK;
}

int lbm_batch_adj_forward(lbm_batch_tape* t, double* J) {
    if (!t) return fail(LBM_ERR_ARG, "null tape");
    lbm_batch* b = t->b;
    HIPCHK(hipSetDevice(b->device));
    if (int rc = sync_domains(b)) return rc;
    for (size_t di = 0; di < b->doms.size(); ++di) {
        lbm_domain* d = b->doms[di];
        HIPCHK(hipMemcpyAsync(d->buf[0], t->chunk[di], d->field_bytes, hipMemcpyDeviceToDevice, b->stream));
        d->cur = 0;
        d->materialised = t->start_mat;
    }
    Events ev(2);
    HIPCHK(ev.create());
    HIPCHK(hipEventRecord(ev[0], b->stream));
    for (uint64_t si = 0; si < t->nck; ++si) {
        const uint64_t steps = std::min(t->k, t->T - si * t->k);
        int parity, mat;
        if (int rc = check_parity(b, parity, mat)) return rc;
        if (int rc = ensure_graph(b, steps, parity, mat)) return rc;
        HIPCHK(hipGraphLaunch(b->exec, b->stream));
        advance_domains(b, steps);
        if (si + 1 < t->nck)
            for (size_t di = 0; di < b->doms.size(); ++di) {
                lbm_domain* d = b->doms[di];
                char* dst = static_cast<char*>(t->chunk[di]) + (si + 1) * d->field_bytes;
                HIPCHK(hipMemcpyAsync(dst, d->buf[d->cur], d->field_bytes, hipMemcpyDeviceToDevice, b->stream));
            }
    }
    HIPCHK(hipEventRecord(ev[1], b->stream));
    for (size_t di = 0; di < b->doms.size(); ++di) t->views[di].gfinal = b->doms[di]->buf[b->doms[di]->cur];
    HIPCHK(hipMemcpyAsync(t->d_views, t->views.data(), sizeof(BatchAdjView) * t->views.size(), hipMemcpyHostToDevice,
                          b->stream));
    for (size_t di = 0; di < b->doms.size(); ++di) {
        lbm_domain* d = b->doms[di];
        HIPCHK(t->aops->objective(d->L, d->buf[d->cur], d->d_kind, OBJ_KINETIC, 0, nullptr, d->d_partial, d->nred,
                                  d->d_out4, b->stream));
    }
    HIPCHK(hipStreamSynchronize(b->stream));
    float ms = 0;
    HIPCHK(hipEventElapsedTime(&ms, ev[0], ev[1]));
    t->timing[2] = ms;
    t->timing[3] = static_cast<double>(t->T);
    if (J)
        for (size_t di = 0; di < b->doms.size(); ++di) {
            double o[4];
            HIPCHK(hipMemcpy(o, b->doms[di]->d_out4, sizeof(o), hipMemcpyDeviceToHost));
            J[di] = o[0];
        }
    t->forward_done = true;
    return LBM_OK;
}

int lbm_batch_adj_backward(lbm_batch_tape* t, double* grad_wall) {
    if (!t) return fail(LBM_ERR_ARG, "null tape");
    if (!t->forward_done) return fail(LBM_ERR_STATE, "lbm_batch_adj_backward needs a preceding lbm_batch_adj_forward");
    lbm_batch* b = t->b;
    HIPCHK(hipSetDevice(b->device));
    int64_t gw_elems = 0;
    for (const lbm_domain* d : b->doms) gw_elems += 3 * d->L.plane;
    HIPCHK(hipMemsetAsync(t->d_gw, 0, sizeof(double) * gw_elems, b->stream));
    BatchAdjArgs a{};
    a.views = t->d_views;
    a.tiles = t->d_tiles;
    a.ntiles = t->ntiles;
    a.idx32 = b->all_idx32 ? 1 : 0;
    // tiles grouped by L2 domain: the recompute and adjoint steps move more
    // data per cell than the forward step (slots, mu, gradient accumulators);
    // measured on 512 cavities over 2000 steps, recompute 7.1 against 8.9 s
    // and adjoint 10.7 against 11.3 s
    a.nxcd = xcd_count(b->ncu);
    const int grid = grid_for(t);
    Events ev(2);
    HIPCHK(ev.create());
    double rec = 0, adj = 0;
    // elapsed time between two points of the stream, into acc
    auto timed = [&](double& acc, auto&& body) -> int {
        HIPCHK(hipEventRecord(ev[0], b->stream));
        if (int rc = body()) return rc;
        HIPCHK(hipEventRecord(ev[1], b->stream));
        HIPCHK(hipEventSynchronize(ev[1]));
        float ms = 0;
        HIPCHK(hipEventElapsedTime(&ms, ev[0], ev[1]));
        acc += ms;
        return LBM_OK;
    };
    // recompute states start+1 .. end-1 from slot first (state `start`) into
    // the segment slots, then the adjoint steps end .. start+1
    auto segment = [&](uint64_t start, uint64_t end, int first) -> int {
        auto state = [&](uint64_t s) { return s == start ? first : seg_slot(t, s - start); };
        if (int rc = timed(rec, [&]() -> int {
                for (uint64_t s = start + 1; s < end; ++s) {
                    a.src_slot = state(s - 1);
                    a.dst_slot = state(s);
                    a.materialised = (s == 1 && t->start_mat) ? 1 : 0;
                    HIPCHK(t->ops->recompute(a, grid, b->stream));
                }
                return LBM_OK;
            }))
            return rc;
        return timed(adj, [&]() -> int {
            for (uint64_t s = end; s > start; --s) {
                a.gprev_slot = state(s - 1);
                a.mu_in_slot = mu_slot(t, static_cast<int>((s + 1) & 1));
                a.mu_out_slot = mu_slot(t, static_cast<int>(s & 1));
                a.terminal = (s == t->T) ? 1 : 0;
                a.materialised = (s == 1 && t->start_mat) ? 1 : 0;
                HIPCHK(t->ops->backward(a, grid, b->stream));
            }
            return LBM_OK;
        });
    };
    for (uint64_t si = t->nck; si-- > 0;) {
        const uint64_t start = si * t->k;
        const uint64_t end = std::min(start + t->k, t->T);
        if (!t->k2) {
            if (int rc = segment(start, end, static_cast<int>(si))) return rc;
            continue;
        }
        // second level: one recompute of the segment lays down the checkpoints
        // of its sub-segments (states start + j k2), passing through two
        // scratch slots in between
        const uint64_t nsub = (end - start + t->k2 - 1) / t->k2;
        if (int rc = timed(rec, [&]() -> int {
                int src = static_cast<int>(si);
                const uint64_t last = start + (nsub - 1) * t->k2;
                for (uint64_t s = start + 1; s <= last; ++s) {
                    const int dst = (s - start) % t->k2 == 0 ? sub_slot(t, (s - start) / t->k2)
                                                              : tmp_slot(t, static_cast<int>(s & 1));
                    a.src_slot = src;
                    a.dst_slot = dst;
                    a.materialised = (s == 1 && t->start_mat) ? 1 : 0;
                    HIPCHK(t->ops->recompute(a, grid, b->stream));
                    src = dst;
                }
                return LBM_OK;
            }))
            return rc;
        for (uint64_t sj = nsub; sj-- > 0;) {
            const uint64_t s0 = start + sj * t->k2;
            const uint64_t s1 = std::min(s0 + t->k2, end);
            if (int rc = segment(s0, s1, sj == 0 ? static_cast<int>(si) : sub_slot(t, sj))) return rc;
        }
    }
    HIPCHK(t->ops->reduce(t->d_views, static_cast<int>(t->views.size()), t->d_out, b->stream));
    HIPCHK(hipStreamSynchronize(b->stream));
    t->timing[0] = rec;
    t->timing[1] = adj;
    if (grad_wall)
        HIPCHK(hipMemcpy(grad_wall, t->d_out, sizeof(double) * 3 * t->views.size(), hipMemcpyDeviceToHost));
    return LBM_OK;
}

int lbm_batch_adj_timing(lbm_batch_tape* t, double out[4]) {
    if (!t || !out) return fail(LBM_ERR_ARG, "null argument");
    for (int i = 0; i < 4; ++i) out[i] = t->timing[i];
    return LBM_OK;
}

}  // extern "C"
