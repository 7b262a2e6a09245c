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
#include <string>
#include <vector>

#include "batch.hpp"
#include "domain_impl.hpp"
#include "lbm_hip.h"

using namespace lbm_amd;

struct lbm_batch {
    std::vector<lbm_domain*> doms;
    const BatchOps* ops = nullptr;
    int device = 0;
    int ncu = 0;
    hipStream_t stream = nullptr;
    int V = 0, bpc = 8, bx_policy = 0;
    int kernel = BK_CELLS;      // enum BatchKernel
    int64_t near_cells = 0;     // NEAR cells of all domains
    BatchView* d_views = nullptr;
    TileRef* d_tiles = nullptr;
    NearRef* d_near = nullptr;
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
    }
    const BatchOps* ops = batch_ops(d0->Q, d0->mode);
    if (!ops) return fail(LBM_ERR_UNSUPPORTED, "batching needs D3Q19 with F32, F16, BF16 or F64S storage");
    lbm_batch* b = new lbm_batch();
    b->doms.assign(domains, domains + n);
    b->ops = ops;
    b->device = d0->desc.device;
    (void)hipDeviceGetAttribute(&b->ncu, hipDeviceAttributeMultiprocessorCount, b->device);
    if (b->ncu <= 0) b->ncu = 256;
    for (const lbm_domain* d : b->doms)
        b->ncells += static_cast<uint64_t>(d->L.xl) * d->L.yl * d->L.zl;
    if (hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess) {
        delete b;
        return fail(LBM_ERR_HIP, "hipStreamCreate failed");
    }
    // start from the domains' own cells-per-thread setting
    int V0 = ops->vecs[0];
    for (int i = 0; i < 4 && ops->vecs[i]; ++i)
        if (ops->vecs[i] == d0->plan.V) V0 = d0->plan.V;
    int rc = configure(b, V0, 0, 8, BK_CELLS);
    if (rc) {
        free_lists(b);
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
    hipEvent_t e0, e1;
    HIPCHK(hipEventCreate(&e0));
    HIPCHK(hipEventCreate(&e1));
    // build (capture + instantiate) outside the timed region, then time one launch
    int parity, mat;
    if (int rc = check_parity(b, parity, mat)) return rc;
    if (int rc = sync_domains(b)) return rc;
    if (int rc = ensure_graph(b, nsteps, parity, mat)) return rc;
    HIPCHK(hipEventRecord(e0, b->stream));
    HIPCHK(hipGraphLaunch(b->exec, b->stream));
    HIPCHK(hipEventRecord(e1, b->stream));
    advance_domains(b, nsteps);
    HIPCHK(hipEventSynchronize(e1));
    float ms = 0;
    HIPCHK(hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
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
    int bv = b->V, bp = b->bx_policy, bb = b->bpc, bk = b->kernel;
    // The trials step the real fields; save every domain's current field and
    // restore it afterwards, so tuning leaves the state as it found it.
    // (Trial steps write only fluid cells, which the next step overwrites in
    // the other buffer.)
    std::vector<void*> saved(b->doms.size(), nullptr);
    std::vector<int> saved_cur(b->doms.size());
    std::vector<bool> saved_mat(b->doms.size());
    auto release = [&] {
        for (void* p : saved) (void)hipFree(p);
    };
    for (size_t i = 0; i < b->doms.size(); ++i) {
        lbm_domain* d = b->doms[i];
        if (hipMalloc(&saved[i], d->field_bytes) != hipSuccess) {
            release();
            return fail(LBM_ERR_ALLOC, "tuning needs a copy of every field");
        }
        HIPCHK(hipMemcpyAsync(saved[i], d->buf[d->cur], d->field_bytes, hipMemcpyDeviceToDevice, b->stream));
        saved_cur[i] = d->cur;
        saved_mat[i] = d->materialised;
    }
    hipEvent_t e0, e1;
    HIPCHK(hipEventCreate(&e0));
    HIPCHK(hipEventCreate(&e1));
    // Kernels are compiled without FMA contraction, so a cell's arithmetic is
    // the same for every V, tile shape and grid size: the search changes
    // speed only, never results.
    for (int kern : {BK_CELLS, BK_ROWS_FUSED, BK_ROWS_SEPARATE}) {
        for (int i = 0; i < 4 && b->ops->vecs[i]; ++i)
            for (int pol = 0; pol < 2; ++pol)
                for (int bpc : {0, 4, 8, 16}) {
                    // cell tiles have one shape: try it once per grid size
                    if (kern == BK_CELLS && (i > 0 || pol > 0)) continue;
                    if (int rc = configure(b, b->ops->vecs[i], pol, bpc, kern)) return rc;
                    // trial runs advance the fields; tuning happens before the timed run
                    int p, m;
                    check_parity(b, p, m);
                    HIPCHK(enqueue(b, 1, p, m));  // warm-up
                    advance_domains(b, 1);
                    check_parity(b, p, m);
                    HIPCHK(hipEventRecord(e0, b->stream));
                    HIPCHK(enqueue(b, static_cast<uint64_t>(trial_steps), p, m));
                    HIPCHK(hipEventRecord(e1, b->stream));
                    HIPCHK(hipEventSynchronize(e1));
                    advance_domains(b, static_cast<uint64_t>(trial_steps));
                    float ms = 0;
                    HIPCHK(hipEventElapsedTime(&ms, e0, e1));
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
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    for (size_t i = 0; i < b->doms.size(); ++i) {
        lbm_domain* d = b->doms[i];
        d->cur = saved_cur[i];
        d->materialised = saved_mat[i];
        HIPCHK(hipMemcpyAsync(d->buf[d->cur], saved[i], d->field_bytes, hipMemcpyDeviceToDevice, b->stream));
    }
    HIPCHK(hipStreamSynchronize(b->stream));
    release();
    if (int rc = configure(b, bv, bp, bb, bk)) return rc;
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
