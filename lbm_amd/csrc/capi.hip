// C ABI of the MI355X collide-stream step (include/lbm_hip.h).
//
// Host side of one domain: the handler map of the reference's ghosted box
// (one boundary-condition object per setBoundaryCondition call, last call
// wins per cell), its translation into device cell kinds at finalize, the
// two distribution buffers and the per-step launch of the kernels of the
// domain's (lattice, storage) pair.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <string>
#include <vector>

#include "domain_impl.hpp"
#include "lbm_hip.h"
#include "domain_view.hpp"
#include "lbm_hip/lattice.hpp"
#include "ops.hpp"
#include "storage.hpp"

using namespace lbm_amd;


namespace lbm_amd {
std::string& last_error() {
    thread_local std::string e;
    return e;
}
}  // namespace lbm_amd

namespace {

// Reference equilibrium (include/collision.hpp:34-51) on the host, for the
// constant Inflow table.
void host_feq(const HostLattice& lat, double density, const double* v, double* feq) {
    // separate multiplies and adds, as the kernels round them
#pragma clang fp contract(off)
    const double cs = kCS;
    for (int q = 0; q < lat.Q; ++q) {
        double c_dot_u = 0.0;
        double u_dot_u = 0.0;
        for (int d = 0; d < 3; ++d) {
            c_dot_u += static_cast<double>(lat.c[q][d]) * v[d];
            u_dot_u += v[d] * v[d];
        }
        feq[q] = lat.w[q] * density *
                 (1 + c_dot_u / (cs * cs) + (c_dot_u) * (c_dot_u) / (2 * cs * cs * cs * cs) -
                  u_dot_u / (2 * cs * cs));
    }
}

__global__ void __launch_bounds__(256) k_copy16(const uint4* __restrict__ src, uint4* __restrict__ dst,
                                                int64_t n) {
    const int64_t i = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
    if (i < n) dst[i] = src[i];
}

// q-plane copy: thread i copies elements 2i, 2i + 1 of every plane (all
// loads, then all non-temporal stores): the data movement of one pull step
// without its arithmetic, as a bandwidth reference for q read and q write
// streams.
template <int Q>
__global__ void __launch_bounds__(256) k_copy_planes(const float2* __restrict__ src, float2* __restrict__ dst,
                                                     int64_t plane2) {
    const int64_t i = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
    if (i >= plane2) return;
    float2 v[Q];
#pragma unroll
    for (int k = 0; k < Q; ++k) v[k] = src[k * plane2 + i];
#pragma unroll
    for (int k = 0; k < Q; ++k) {
        uint32_t __attribute__((ext_vector_type(2))) r;
        __builtin_memcpy(&r, &v[k], sizeof(r));
        __builtin_nontemporal_store(r, reinterpret_cast<uint32_t __attribute__((ext_vector_type(2)))*>(dst + k * plane2 + i));
    }
}

template <class T>
__global__ void k_fill_planes(T* buf, int64_t plane, int q, const double* values) {
    const int64_t n = plane * q;
    for (int64_t i = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; i < n;
         i += static_cast<int64_t>(gridDim.x) * blockDim.x)
        buf[i] = static_cast<T>(values[i / plane]);
}

}  // namespace

namespace {

int make_plan(lbm_domain* d) {
    const ModeOps* o = d->ops;
    if (d->compact && d->tune_vec > 1)
        return fail(LBM_ERR_ARG, "this domain packs its rows (short x rows): only 1 cell per thread");
    int V = d->tune_vec ? d->tune_vec : (d->compact ? o->vecs[0] : o->default_vec);
    bool ok = false;
    for (int i = 0; i < 4 && o->vecs[i]; ++i) ok |= (o->vecs[i] == V);
    if (!ok) return fail(LBM_ERR_ARG, "vector width " + std::to_string(V) + " not built for this mode");
    StepPlan& p = d->plan;
    p.V = V;
    const int nvx = (d->L.xl + V - 1) / V;
    int bx = d->tune_bx;
    if (!bx) {
        bx = 1;
        while (bx < nvx && bx < 256) bx *= 2;
        // the mode's preferred width, where rows are at least that long
        if (o->default_bx && d->tune_vec == 0 && bx > o->default_bx) bx = o->default_bx;
    }
    if (bx < 1 || bx > 256 || (256 % bx)) return fail(LBM_ERR_ARG, "block_x must divide 256");
    p.BX = bx;
    p.BY = 256 / bx;
    p.nbx = (nvx + p.BX - 1) / p.BX;
    p.nby = (d->L.yl + p.BY - 1) / p.BY;
    p.nblocks_bulk = static_cast<int64_t>(p.nbx) * p.nby * d->L.zl;
    if (p.nblocks_bulk > 0x7fffffffLL) return fail(LBM_ERR_UNSUPPORTED, "grid too large");
    // Domains with boundary cells step in cell tiles unless a vector width was
    // asked for; bulk-only (periodic) domains keep the vectorised row kernel.
    const int64_t plane_cells = static_cast<int64_t>(d->L.xl) * d->L.yl;
    p.tpp = static_cast<int>((plane_cells + 63) / 64);
    p.nblocks_cells = (static_cast<int64_t>(p.tpp) * d->L.zl + 3) / 4;
    p.rxl = 1.0 / static_cast<double>(d->L.xl);
    p.freeslip = false;
    for (const HostCond& c : d->conds) p.freeslip |= c.kind == LBM_BC_FREESLIP;
    p.idx32 = !d->force64 && d->L.plane * static_cast<int64_t>(d->esize) < (int64_t{1} << 32);
    p.cells = d->tune_vec == 0 && (d->nb > 0 || d->compact) &&
              p.nblocks_cells <= 0x7fffffffLL && (plane_cells + 63) / 64 <= 0x7fffffffLL;
    if (!p.cells && d->compact && V > 1)
        return fail(LBM_ERR_UNSUPPORTED, "packed rows need the cell-tile kernel or 1 cell per thread");
    // Two steps per pass where every interior cell is bulk: no cell pulls from
    // a non-fluid source, so all sources are interior cells (periodic images).
    p.tb_ntx = (d->L.xl + TB_TX - 1) / TB_TX;
    p.tb_nty = (d->L.yl + TB_TY - 1) / TB_TY;
    const int64_t tiles = static_cast<int64_t>(p.tb_ntx) * p.tb_nty;
    int ncu = 0;
    if ((o->step2 || o->step2w) &&
        hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, d->desc.device) != hipSuccess)
        ncu = 0;
    // Which pass would run: k_bulk2 where every interior cell is bulk, k_wall2
    // where there are boundary cells (not FreeSlip, whose rule reads other
    // cells than the one it serves; no Brinkman field).
    const bool walled = d->nb > 0 || p.flags;
    const bool pass_ok = d->tune_pass != 1 && !d->aa &&
                         d->L.plane * static_cast<int64_t>(d->esize) < (int64_t{1} << 32) &&
                         (d->esize >= 4 || d->L.xl > 32);
    // Default z chunks: a block marches through its zc planes plus one more
    // at each end, so the pass takes about rounds x (zc + 2) plane times,
    // with rounds = ceil(blocks / resident blocks). The chunk count that
    // minimises it is taken, among chunks of at most 32 planes (longer ones
    // measured slower on planes of many tiles).
    int zc = d->tune_zc;
    const bool use_w = walled && o->step2w != nullptr;
    // rule families the walled pass is built with: NoSlip (+ MovingWall) or all
    p.wrules = RULE_MOVINGWALL;
    for (const HostCond& c : d->conds)
        if (c.kind != LBM_BC_NOSLIP && c.kind != LBM_BC_MOVINGWALL) p.wrules = RULES_ALL;
    if (!zc && (use_w ? o->step2w != nullptr : o->step2 != nullptr)) {
        const int per_cu = use_w ? o->step2w_blocks_per_cu(p.wrules) : o->step2_blocks_per_cu(p.brink);
        const int64_t resident = std::max<int64_t>(1, static_cast<int64_t>(ncu) * per_cu);
        const int64_t zl = d->L.zl;
        int64_t best = -1, best_cost = 0;
        for (int64_t c = (zl + 31) / 32; c <= zl; ++c) {
            const int64_t z = (zl + c - 1) / c;
            if (c > 1 && (zl + c - 2) / (c - 1) == z) continue;  // same chunk length as c - 1
            const int64_t cost = ((c * tiles + resident - 1) / resident) * (z + 2);
            if (best < 0 || cost < best_cost) {
                best = c;
                best_cost = cost;
            }
        }
        zc = static_cast<int>((zl + best - 1) / best);
    }
    p.tb_zc = std::max(1, std::min(zc, d->L.zl));
    // gfx950 deals blocks in turn to its XCDs (32 CUs and one L2 each);
    // grouping neighbouring tiles on one XCD lets them share halo lines in L2
    p.tb_nxcd = xcd_count(ncu);
    p.tb_nblocks = tiles * ((d->L.zl + p.tb_zc - 1) / p.tb_zc);
    // (byte offsets within a plane are 32-bit in the pass.) The 16-bit pass
    // is bound by instruction issue: on rows of 32 cells or fewer most lanes
    // of its 64-cell tile rows idle, and single steps (cell tiles over packed
    // rows) measured faster there.
    p.pairs = pass_ok && o->step2 != nullptr && !walled && p.has_bulk && p.tb_nblocks <= 0x7fffffffLL;
    // Automatic choice (tune_pass 0): the walled pass with fp32 storage (256^3
    // cavity D3Q19: 0.36 against 0.50 ms per step; 384x192x192 channel around
    // spheres D3Q27: 0.80 against 0.82). With 16-bit storage it is bound by
    // instruction issue and measured slower than single steps (the cavity
    // in fp16: 0.39 against 0.28 ms per step).
    p.wpairs = pass_ok && o->step2w != nullptr && walled && d->tune_vec == 0 && !p.freeslip && !p.brink &&
               d->n_fluid > 0 && (d->tune_pass == 2 || d->esize == 4) &&
               p.tb_nblocks <= 0x7fffffffLL;
    return LBM_OK;
}

void free_device(lbm_domain* d) {
    (void)hipFree(d->buf[0]);
    (void)hipFree(d->buf[1]);
    (void)hipFree(d->d_kind);
    (void)hipFree(d->d_lmask);
    (void)hipFree(d->d_blist);
    (void)hipFree(d->d_conds);
    (void)hipFree(d->d_alpha);
    (void)hipFree(d->d_partial);
    (void)hipFree(d->d_out4);
    d->buf[0] = d->buf[1] = nullptr;
    d->d_kind = nullptr;
    d->d_lmask = nullptr;
    d->d_lcond = nullptr;
    d->d_blist = nullptr;
    d->d_conds = nullptr;
    d->d_alpha = nullptr;
    d->d_partial = nullptr;
    d->d_out4 = nullptr;
}

int fill_initial(lbm_domain* d, void* buf) {
    // EXACT64 holds the weights; shifted modes hold w - w = 0.
    if (d->mode != Mode::EXACT64) {
        HIPCHK(hipMemsetAsync(buf, 0, d->field_bytes, d->stream));
        return LBM_OK;
    }
    DeviceTemp dv;
    HIPCHK(hipMalloc(&dv.p, sizeof(double) * d->Q));
    HIPCHK(hipMemcpyAsync(dv.p, d->lat.w, sizeof(double) * d->Q, hipMemcpyHostToDevice, d->stream));
    hipLaunchKernelGGL(k_fill_planes<double>, dim3(4096), dim3(256), 0, d->stream,
                       static_cast<double*>(buf), d->L.plane, d->Q, static_cast<const double*>(dv.p));
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(d->stream));
    return LBM_OK;
}

}  // namespace

extern "C" {

const char* lbm_last_error(void) { return lbm_amd::last_error().c_str(); }
int lbm_abi_version(void) { return LBM_HIP_ABI_VERSION; }

int lbm_lattice(int q, int* c, double* w, double* cs) {
    if (q != 15 && q != 19 && q != 27) return fail(LBM_ERR_ARG, "q must be 15, 19 or 27");
    const HostLattice lat = lattice_for(q);
    for (int i = 0; i < q; ++i) {
        if (c)
            for (int k = 0; k < 3; ++k) c[3 * i + k] = lat.c[i][k];
        if (w) w[i] = lat.w[i];
    }
    if (cs) *cs = kCS;
    return LBM_OK;
}

int lbm_desc_init(lbm_desc* desc) {
    if (!desc) return fail(LBM_ERR_ARG, "null desc");
    std::memset(desc, 0, sizeof(*desc));
    desc->q = 19;
    desc->storage = LBM_STORAGE_F64;
    desc->tau = 1.0;
    desc->spacing[0] = desc->spacing[1] = desc->spacing[2] = 1.0;
    return LBM_OK;
}

int lbm_domain_create(const lbm_desc* desc, lbm_domain** out) {
    if (!desc || !out) return fail(LBM_ERR_ARG, "null argument");
    *out = nullptr;
    if (desc->q != 15 && desc->q != 19 && desc->q != 27)
        return fail(LBM_ERR_ARG, "q must be 15, 19 or 27");
    if (desc->storage < 0 || desc->storage > 4) return fail(LBM_ERR_ARG, "unknown storage mode");
    if (desc->layout != LBM_LAYOUT_AB && desc->layout != LBM_LAYOUT_AA) return fail(LBM_ERR_ARG, "unknown layout");
    const uint64_t maxl = 1u << 20;
    if (desc->xl < 1 || desc->yl < 1 || desc->zl < 1 || desc->xl > maxl || desc->yl > maxl ||
        desc->zl > maxl)
        return fail(LBM_ERR_ARG, "lengths must be in [1, 2^20]");
    if (!(desc->tau > 0.5) || !std::isfinite(desc->tau))
        return fail(LBM_ERR_ARG, "tau must be > 0.5");
    hipError_t e = hipSetDevice(desc->device);
    if (e != hipSuccess) return fail(LBM_ERR_HIP, std::string("hipSetDevice: ") + hipGetErrorString(e));

    lbm_domain* d = new lbm_domain();
    d->desc = *desc;
    d->mode = static_cast<Mode>(desc->storage);
    d->ops = mode_ops(desc->q, d->mode);
    if (!d->ops) {
        delete d;
        return fail(LBM_ERR_UNSUPPORTED, "no kernels for this lattice/storage");
    }
    d->lat = lattice_for(desc->q);
    d->Q = desc->q;
    d->nx = static_cast<int>(desc->xl + 2);
    d->ny = static_cast<int>(desc->yl + 2);
    d->nz = static_cast<int>(desc->zl + 2);
    d->ncell = static_cast<int64_t>(d->nx) * d->ny * d->nz;
    d->esize = elem_size(d->mode);
    const int64_t A = 128 / static_cast<int64_t>(d->esize);
    FieldLayout& L = d->L;
    // Rows start on a 128-byte line (x = 1 aligned for the vector row
    // kernel) unless a row is shorter than two lines: then padding would be a
    // large share of every line read, and rows are packed at pitch xl + 2
    // for the cell-tile kernel, which needs no alignment.
    d->compact = d->nx < 2 * A;
    L.off = d->compact ? 0 : A - 1;
    L.px = d->compact ? d->nx : roundup(d->nx, A);
    L.pxy = L.px * d->ny;
    // Planes start on a 128-byte line; 16 spare elements at the end keep the
    // vector kernel's last row reads inside the plane.
    L.plane = roundup(L.off + L.pxy * d->nz + 16, A);
    L.xl = static_cast<int>(desc->xl);
    L.yl = static_cast<int>(desc->yl);
    L.zl = static_cast<int>(desc->zl);
    for (int k = 0; k < 3; ++k) L.per[k] = desc->periodic[k] ? 1 : 0;
    d->field_bytes = static_cast<size_t>(L.plane) * d->Q * d->esize;
    d->handler.assign(static_cast<size_t>(d->ncell), 0);

    auto bail = [&](int code, const std::string& msg) {
        free_device(d);
        if (d->stream) (void)hipStreamDestroy(d->stream);
        delete d;
        return fail(code, msg);
    };
    if (hipStreamCreateWithFlags(&d->stream, hipStreamNonBlocking) != hipSuccess)
        return bail(LBM_ERR_HIP, "hipStreamCreate failed");
    d->aa = desc->layout == LBM_LAYOUT_AA;
    const int nbuf = d->aa ? 1 : 2;
    for (int k = 0; k < nbuf; ++k)
        if (hipMalloc(&d->buf[k], d->field_bytes) != hipSuccess)
            return bail(LBM_ERR_ALLOC, "hipMalloc of a distribution buffer failed");
    if (hipMalloc(&d->d_kind, static_cast<size_t>(L.plane)) != hipSuccess)
        return bail(LBM_ERR_ALLOC, "hipMalloc of the kind array failed");
    d->device_bytes = nbuf * d->field_bytes + static_cast<size_t>(L.plane);
    for (int k = 0; k < nbuf; ++k) {
        int rc = fill_initial(d, d->buf[k]);
        if (rc) return bail(rc, last_error());
    }
    if (hipStreamSynchronize(d->stream) != hipSuccess) return bail(LBM_ERR_HIP, "initial fill failed");
    *out = d;
    return LBM_OK;
}

int lbm_domain_destroy(lbm_domain* d) {
    if (!d) return LBM_OK;
    (void)hipSetDevice(d->desc.device);
    if (d->stream) (void)hipStreamSynchronize(d->stream);
    free_device(d);
    if (d->stream) (void)hipStreamDestroy(d->stream);
    delete d;
    return LBM_OK;
}

int lbm_set_boundary_box(lbm_domain* d, int kind, const double params[4], uint64_t x0,
                         uint64_t xe, uint64_t y0, uint64_t ye, uint64_t z0, uint64_t ze) {
    if (!d) return fail(LBM_ERR_ARG, "null domain handle");
    if (d->finalized) return fail(LBM_ERR_STATE, "geometry is frozen after lbm_finalize");
    if (kind < LBM_BC_NOSLIP || kind > LBM_BC_PRESSURE)
        return fail(LBM_ERR_ARG, "unknown boundary kind " + std::to_string(kind));
    // include/domain.hpp:180-181
    if (xe < x0 || ye < y0 || ze < z0) return fail(LBM_ERR_ARG, "box end before box start");
    if (xe > static_cast<uint64_t>(d->nx - 1) || ye > static_cast<uint64_t>(d->ny - 1) ||
        ze > static_cast<uint64_t>(d->nz - 1))
        return fail(LBM_ERR_ARG, "box exceeds the ghosted domain");
    if (d->conds.size() >= 0xFFFE) return fail(LBM_ERR_ARG, "too many boundary conditions");
    HostCond c{};
    c.kind = kind;
    c.rho = 1.0;
    if (params) {
        c.vel[0] = params[0];
        c.vel[1] = params[1];
        c.vel[2] = params[2];
        c.rho = params[3];
    }
    d->conds.push_back(c);
    const uint16_t h = static_cast<uint16_t>(d->conds.size());
    for (uint64_t z = z0; z <= ze; ++z)
        for (uint64_t y = y0; y <= ye; ++y)
            for (uint64_t x = x0; x <= xe; ++x) d->handler[d->hidx(x, y, z)] = h;
    return LBM_OK;
}

int lbm_set_solid_mask(lbm_domain* d, const uint8_t* fluid_mask) {
    if (!d || !fluid_mask) return fail(LBM_ERR_ARG, "null argument");
    if (d->finalized) return fail(LBM_ERR_STATE, "geometry is frozen after lbm_finalize");
    if (d->mask_cond < 0) {
        HostCond c{};
        c.kind = LBM_BC_NOSLIP;
        c.rho = 1.0;
        d->conds.push_back(c);
        d->mask_cond = static_cast<int>(d->conds.size()) - 1;
    }
    const uint16_t h = static_cast<uint16_t>(d->mask_cond + 1);
    int64_t i = 0;
    for (int z = 1; z <= d->L.zl; ++z)
        for (int y = 1; y <= d->L.yl; ++y)
            for (int x = 1; x <= d->L.xl; ++x, ++i)
                if (!fluid_mask[i]) d->handler[d->hidx(x, y, z)] = h;
    return LBM_OK;
}

int lbm_set_brinkman(lbm_domain* d, const double* alpha, const double g[3]) {
    if (!d) return fail(LBM_ERR_ARG, "null domain handle");
    if (d->finalized) return fail(LBM_ERR_STATE, "call lbm_set_brinkman before lbm_finalize");
    if (d->mode == Mode::EXACT64)
        return fail(LBM_ERR_UNSUPPORTED, "Brinkman forcing needs a shifted storage mode");
    const int64_t n = static_cast<int64_t>(d->L.xl) * d->L.yl * d->L.zl;
    if (alpha)
        d->h_alpha.assign(alpha, alpha + n);
    else
        d->h_alpha.clear();
    for (int k = 0; k < 3; ++k) d->g[k] = g ? g[k] : 0.0;
    return LBM_OK;
}

int lbm_finalize(lbm_domain* d) {
    if (int rc = check_handle(d)) return rc;
    if (d->finalized) return fail(LBM_ERR_STATE, "already finalized");
    if (d->aa)
        for (const HostCond& c : d->conds)
            if (c.kind == LBM_BC_FREESLIP)
                return fail(LBM_ERR_UNSUPPORTED, "the in-place layout has no FreeSlip boundary (use the two-buffer layout)");
    const HostLattice& lat = d->lat;
    const int Q = d->Q;
    const FieldLayout& L = d->L;
    auto is_fluid = [&](int64_t x, int64_t y, int64_t z) { return d->handler[d->hidx(x, y, z)] == 0; };

    // Reject pulls from unassigned ghost cells, and FreeSlip chains that would
    // consult one (the reference treats such a ghost as a BGK cell that is
    // collided but never streamed into).
    for (int z = 1; z <= L.zl; ++z)
        for (int y = 1; y <= L.yl; ++y)
            for (int x = 1; x <= L.xl; ++x) {
                if (!is_fluid(x, y, z)) continue;
                for (int q = 0; q < Q; ++q) {
                    const int64_t sx = d->wrapc(x - lat.c[q][0], 0);
                    const int64_t sy = d->wrapc(y - lat.c[q][1], 1);
                    const int64_t sz = d->wrapc(z - lat.c[q][2], 2);
                    const uint16_t h = d->handler[d->hidx(sx, sy, sz)];
                    if (!d->interior(sx, sy, sz) && h == 0)
                        return fail(LBM_ERR_GEOMETRY,
                                    "fluid cell (" + std::to_string(x) + "," + std::to_string(y) +
                                        "," + std::to_string(z) +
                                        ") pulls from a ghost cell that no boundary covers");
                    if (h != 0 && d->conds[h - 1].kind == LBM_BC_FREESLIP) {
                        const int64_t cand[6][3] = {{x, sy, sz}, {sx, y, sz}, {sx, sy, z},
                                                    {sx, y, z},  {x, sy, z},  {x, y, sz}};
                        for (int j = 0; j < 6; ++j)
                            if (!d->interior(cand[j][0], cand[j][1], cand[j][2]) &&
                                d->handler[d->hidx(cand[j][0], cand[j][1], cand[j][2])] == 0)
                                return fail(LBM_ERR_GEOMETRY,
                                            "free-slip rule would read a ghost cell that no "
                                            "boundary covers");
                    }
                }
            }

    // Distinct conditions (the rules depend only on kind and parameters).
    std::vector<int> cond_map(d->conds.size(), -1);
    std::vector<HostCond> uniq;
    for (size_t i = 0; i < d->conds.size(); ++i) {
        const HostCond& c = d->conds[i];
        for (size_t j = 0; j < uniq.size(); ++j)
            if (uniq[j].kind == c.kind && std::memcmp(uniq[j].vel, c.vel, sizeof(c.vel)) == 0 &&
                std::memcmp(&uniq[j].rho, &c.rho, sizeof(double)) == 0) {
                cond_map[i] = static_cast<int>(j);
                break;
            }
        if (cond_map[i] < 0) {
            cond_map[i] = static_cast<int>(uniq.size());
            uniq.push_back(c);
        }
    }
    if (static_cast<int>(uniq.size()) > kMaxConds)
        return fail(LBM_ERR_UNSUPPORTED, "more than 252 distinct boundary conditions");
    std::vector<Cond> table(uniq.size());
    for (size_t j = 0; j < uniq.size(); ++j) {
        Cond& c = table[j];
        std::memset(&c, 0, sizeof(c));
        c.kind = uniq[j].kind;
        for (int k = 0; k < 3; ++k) c.vel[k] = uniq[j].vel[k];
        c.rho = uniq[j].rho;
        for (int q = 0; q < Q; ++q) {
            c.cu64[q] = lat.c[q][0] * c.vel[0] + lat.c[q][1] * c.vel[1] + lat.c[q][2] * c.vel[2];
            c.cu32[q] = static_cast<float>(c.cu64[q]);
        }
        if (c.kind == LBM_BC_INFLOW) {
            host_feq(lat, c.rho, c.vel, c.feq);
            for (int q = 0; q < Q; ++q) {
                const double cu = lat.c[q][0] * c.vel[0] + lat.c[q][1] * c.vel[1] + lat.c[q][2] * c.vel[2];
                const double uu = c.vel[0] * c.vel[0] + c.vel[1] * c.vel[1] + c.vel[2] * c.vel[2];
                c.geq[q] = lat.w[q] * ((c.rho - 1.0) + c.rho * (3.0 * cu + 4.5 * cu * cu - 1.5 * uu));
            }
        }
    }

    // Cell kinds in the padded device layout, the NEAR list, the link masks
    // of NEAR cells (bit q: the source x - c_q is not fluid) and the
    // condition their links share (LCOND_MIXED where they differ).
    std::vector<uint8_t> kind(static_cast<size_t>(L.plane), KIND_INERT);
    std::vector<uint32_t> lmask(static_cast<size_t>(L.plane), 0u);
    std::vector<uint8_t> lcond(static_cast<size_t>(L.plane), LCOND_MIXED);
    std::vector<CellCoord> blist;
    uint64_t nfluid = 0, nbulk = 0, nonbulk_interior = 0;
    for (int z = 0; z < d->nz; ++z)
        for (int y = 0; y < d->ny; ++y)
            for (int x = 0; x < d->nx; ++x) {
                const uint16_t h = d->handler[d->hidx(x, y, z)];
                const int64_t e = cell_elem(L, x, y, z);
                const bool in = d->interior(x, y, z);
                if (h != 0) {
                    kind[e] = static_cast<uint8_t>(KIND_COND0 + cond_map[h - 1]);
                    if (in) ++nonbulk_interior;
                    continue;
                }
                if (!in) continue;  // unassigned ghost: INERT
                ++nfluid;
                uint32_t m = 0;
                int shared = -1;  // condition of the links so far (-2: more than one)
                for (int q = 0; q < Q; ++q) {
                    const int64_t sx = d->wrapc(x - lat.c[q][0], 0);
                    const int64_t sy = d->wrapc(y - lat.c[q][1], 1);
                    const int64_t sz = d->wrapc(z - lat.c[q][2], 2);
                    const uint16_t hs = d->handler[d->hidx(sx, sy, sz)];
                    if (hs != 0) {
                        m |= 1u << q;
                        const int ci = cond_map[hs - 1];
                        shared = shared == -1 ? ci : (shared == ci ? ci : -2);
                    }
                }
                if (m) {
                    kind[e] = KIND_NEAR;
                    lmask[e] = m;
                    if (shared >= 0) lcond[e] = static_cast<uint8_t>(shared);
                    blist.push_back(CellCoord{x, y, z, 0});
                    ++nonbulk_interior;
                } else {
                    kind[e] = KIND_BULK;
                    ++nbulk;
                }
            }

    HIPCHK(hipMemcpyAsync(d->d_kind, kind.data(), kind.size(), hipMemcpyHostToDevice, d->stream));
    if (!d->d_lmask) {
        const size_t bytes = lmask.size() * (sizeof(uint32_t) + sizeof(uint8_t));
        HIPCHK(hipMalloc(&d->d_lmask, bytes));
        d->d_lcond = reinterpret_cast<uint8_t*>(d->d_lmask + lmask.size());
        d->device_bytes += bytes;
    }
    HIPCHK(hipMemcpyAsync(d->d_lmask, lmask.data(), lmask.size() * sizeof(uint32_t), hipMemcpyHostToDevice,
                          d->stream));
    HIPCHK(hipMemcpyAsync(d->d_lcond, lcond.data(), lcond.size(), hipMemcpyHostToDevice, d->stream));
    d->nb = static_cast<int64_t>(blist.size());
    if (d->nb) {
        HIPCHK(hipMalloc(&d->d_blist, blist.size() * sizeof(CellCoord)));
        HIPCHK(hipMemcpyAsync(d->d_blist, blist.data(), blist.size() * sizeof(CellCoord),
                              hipMemcpyHostToDevice, d->stream));
        d->device_bytes += blist.size() * sizeof(CellCoord);
    }
    if (!table.empty()) {
        HIPCHK(hipMalloc(&d->d_conds, table.size() * sizeof(Cond)));
        HIPCHK(hipMemcpyAsync(d->d_conds, table.data(), table.size() * sizeof(Cond),
                              hipMemcpyHostToDevice, d->stream));
    }
    d->nconds_dev = static_cast<int>(table.size());
    if (!d->h_alpha.empty()) {
        const std::vector<char> a = alpha_in_layout(d, d->h_alpha.data());
        HIPCHK(hipMalloc(&d->d_alpha, a.size()));
        HIPCHK(hipMemcpyAsync(d->d_alpha, a.data(), a.size(), hipMemcpyHostToDevice, d->stream));
        d->device_bytes += a.size();
    }
    d->nred = static_cast<int>(std::min<int64_t>(1024, (static_cast<int64_t>(L.xl) * L.yl * L.zl + 255) / 256));
    HIPCHK(hipMalloc(&d->d_partial, sizeof(double) * 4 * d->nred));
    HIPCHK(hipMalloc(&d->d_out4, sizeof(double) * 4));
    HIPCHK(hipStreamSynchronize(d->stream));

    d->n_fluid = nfluid;
    d->plan.flags = nonbulk_interior > 0;
    d->plan.brink = !d->h_alpha.empty();
    d->plan.has_bulk = nbulk > 0;
    if (int rc = make_plan(d)) return rc;
    d->finalized = true;
    return LBM_OK;
}

int lbm_set_tuning(lbm_domain* d, int vec, int block_x) {
    if (int rc = check_handle(d)) return rc;  // (planning queries the domain's device)
    if (d->aa && vec > 0) return fail(LBM_ERR_UNSUPPORTED, "the in-place layout steps with cell tiles only");
    const int ov = d->tune_vec, ob = d->tune_bx;
    d->tune_vec = vec;
    d->tune_bx = block_x;
    if (d->finalized) {
        if (int rc = make_plan(d)) {
            d->tune_vec = ov;
            d->tune_bx = ob;
            make_plan(d);
            return rc;
        }
    }
    return LBM_OK;
}

static int run_steps(lbm_domain* d, uint64_t nsteps, EventLog* log) {
    if (d->aa) {
        for (uint64_t t = 0; t < nsteps; ++t) {
            const StepArgs a = d->args(d->buf[0], d->buf[0], d->materialised ? 1 : 0);
            hipError_t e = d->ops->aa_step(a, d->plan, d->aa_phase, d->stream);
            if (log) log->mark(d->stream, 0);
            if (e != hipSuccess) return fail(LBM_ERR_HIP, std::string("step launch: ") + hipGetErrorString(e));
            d->aa_phase ^= 1;
            d->materialised = false;
        }
        return LBM_OK;
    }
    uint64_t t = 0;
    if (d->plan.wpairs) {
        // the first step after set-up pulls stored boundary values: single
        if (d->materialised && t < nsteps) {
            const StepArgs a = d->args(d->buf[d->cur], d->buf[d->cur ^ 1], 1);
            hipError_t e = d->ops->step(a, d->plan, d->stream, log);
            if (e != hipSuccess) return fail(LBM_ERR_HIP, std::string("step launch: ") + hipGetErrorString(e));
            d->cur ^= 1;
            d->materialised = false;
            ++t;
        }
        for (; t + 2 <= nsteps; t += 2) {
            const StepArgs a = d->args(d->buf[d->cur], d->buf[d->cur ^ 1], 0);
            hipError_t e = d->ops->step2w(a, d->plan, d->stream, log);
            if (e != hipSuccess) return fail(LBM_ERR_HIP, std::string("two-step launch: ") + hipGetErrorString(e));
            d->cur ^= 1;
        }
    }
    if (d->plan.pairs)
        for (; t + 2 <= nsteps; t += 2) {
            // (every source is an interior cell, so the flag has no effect here;
            // it is passed as a single step would take it)
            const StepArgs a = d->args(d->buf[d->cur], d->buf[d->cur ^ 1], d->materialised ? 1 : 0);
            hipError_t e = d->ops->step2(a, d->plan, d->stream, log);
            if (e != hipSuccess) return fail(LBM_ERR_HIP, std::string("two-step launch: ") + hipGetErrorString(e));
            d->cur ^= 1;
            d->materialised = false;
        }
    for (; t < nsteps; ++t) {
        const StepArgs a = d->args(d->buf[d->cur], d->buf[d->cur ^ 1], d->materialised ? 1 : 0);
        hipError_t e = d->ops->step(a, d->plan, d->stream, log);
        if (e != hipSuccess) return fail(LBM_ERR_HIP, std::string("step launch: ") + hipGetErrorString(e));
        d->cur ^= 1;
        d->materialised = false;
    }
    return LBM_OK;
}

}  // extern "C"

namespace lbm_amd {

// Boundary cells of the current field <- the values the reference's boundary
// pass leaves there (include/domain.hpp:157-165): the rule value of every
// link that points into a fluid cell, from that cell's stored values. The
// steps compute those values where they are pulled and do not store them;
// readers of whole cells (lbm_get_pdf, lbm_moments) see the reference's
// field. In the in-place layout the boundary cells already hold them.
int materialise_boundaries(lbm_domain* d) {
    if (d->finalized && !d->materialised && !d->aa) {
        const StepArgs a = d->args(d->buf[d->cur], d->buf[d->cur], 0);
        HIPCHK(d->ops->materialise(a, d->nx, d->ny, d->nz, d->stream));
    }
    return LBM_OK;
}

// In-place layout: bring the field to the natural representation (the
// readers' and writers' layout) if the last step left it swapped.
int aa_settle(lbm_domain* d) {
    if (!d->aa || d->aa_phase == 0) return LBM_OK;
    const StepArgs a = d->args(d->buf[0], d->buf[0], 0);
    const hipError_t e = d->ops->aa_swap(a, d->plan, d->stream);
    if (e != hipSuccess) return fail(LBM_ERR_HIP, std::string("in-place layout swap: ") + hipGetErrorString(e));
    d->aa_phase = 0;
    return LBM_OK;
}

}  // namespace lbm_amd

extern "C" {

int lbm_step(lbm_domain* d, uint64_t nsteps) {
    if (int rc = check_handle(d)) return rc;
    if (!d->finalized) return fail(LBM_ERR_STATE, "call lbm_finalize before lbm_step");
    return run_steps(d, nsteps, nullptr);
}

int lbm_sync(lbm_domain* d) {
    if (int rc = check_handle(d)) return rc;
    HIPCHK(hipStreamSynchronize(d->stream));
    return LBM_OK;
}

int lbm_time_steps(lbm_domain* d, uint64_t nsteps, double out[4]) {
    if (int rc = check_handle(d)) return rc;
    if (!d->finalized) return fail(LBM_ERR_STATE, "call lbm_finalize first");
    if (!out || nsteps == 0 || nsteps > 100000) return fail(LBM_ERR_ARG, "nsteps in [1, 1e5]");
    // one event before the first launch and one after each launch; a launch's
    // time runs from the previous event (so it includes its launch gap)
    Events ev(1 + 2 * nsteps);
    HIPCHK(ev.create());
    std::vector<int> kind(2 * nsteps, 0);
    EventLog log{ev.data() + 1, kind.data(), 0, 2 * nsteps};
    HIPCHK(hipEventRecord(ev[0], d->stream));
    if (int rc = run_steps(d, nsteps, &log)) return rc;
    HIPCHK(hipStreamSynchronize(d->stream));
    float ms = 0;
    double bulk = 0, near = 0;
    for (size_t i = 0; i < log.n; ++i) {
        HIPCHK(hipEventElapsedTime(&ms, ev[i], ev[i + 1]));
        (kind[i] ? near : bulk) += ms;
    }
    HIPCHK(hipEventElapsedTime(&ms, ev[0], ev[log.n]));
    out[0] = ms;
    out[1] = bulk;
    out[2] = near;
    out[3] = static_cast<double>(nsteps);
    return LBM_OK;
}

int lbm_get_pdf(lbm_domain* d, double* out) {
    if (int rc = check_handle(d)) return rc;
    if (!out) return fail(LBM_ERR_ARG, "null output");
    if (!d->finalized && !d->materialised) return fail(LBM_ERR_STATE, "inconsistent state");
    if (int rc = aa_settle(d)) return rc;
    if (int rc = materialise_boundaries(d)) return rc;
    const size_t bytes = static_cast<size_t>(d->ncell) * d->Q * sizeof(double);
    double* dev = nullptr;
    HIPCHK(hipMalloc(&dev, bytes));
    hipError_t e = d->ops->get_pdf(d->L, d->buf[d->cur], dev, d->nx, d->ny, d->nz, d->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(out, dev, bytes, hipMemcpyDeviceToHost, d->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(d->stream);
    (void)hipFree(dev);
    if (e != hipSuccess) return fail(LBM_ERR_HIP, std::string("get_pdf: ") + hipGetErrorString(e));
    return LBM_OK;
}

int lbm_set_pdf(lbm_domain* d, const double* in) {
    if (int rc = check_handle(d)) return rc;
    if (!in) return fail(LBM_ERR_ARG, "null input");
    const size_t bytes = static_cast<size_t>(d->ncell) * d->Q * sizeof(double);
    double* dev = nullptr;
    HIPCHK(hipMalloc(&dev, bytes));
    hipError_t e = hipMemcpyAsync(dev, in, bytes, hipMemcpyHostToDevice, d->stream);
    if (e == hipSuccess) e = d->ops->set_pdf(d->L, d->buf[d->cur], dev, d->nx, d->ny, d->nz, d->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(d->stream);
    (void)hipFree(dev);
    if (e != hipSuccess) return fail(LBM_ERR_HIP, std::string("set_pdf: ") + hipGetErrorString(e));
    d->materialised = true;
    d->aa_phase = 0;
    return LBM_OK;
}

int lbm_init_equilibrium(lbm_domain* d, const double* rho, const double* u) {
    if (int rc = check_handle(d)) return rc;
    if (!d->finalized) return fail(LBM_ERR_STATE, "call lbm_finalize first (fluid cells are needed)");
    if (!rho || !u) return fail(LBM_ERR_ARG, "null input");
    const size_t n = static_cast<size_t>(d->L.xl) * d->L.yl * d->L.zl;
    if (d->aa) {
        // boundary cells back to the weights, as the two-buffer layout holds them
        if (int rc = fill_initial(d, d->buf[0])) return rc;
        d->aa_phase = 0;
    }
    double* dev = nullptr;
    HIPCHK(hipMalloc(&dev, 4 * n * sizeof(double)));
    hipError_t e = hipMemcpyAsync(dev, rho, n * sizeof(double), hipMemcpyHostToDevice, d->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(dev + n, u, 3 * n * sizeof(double), hipMemcpyHostToDevice, d->stream);
    if (e == hipSuccess) e = d->ops->init_eq(d->L, d->buf[d->cur], d->d_kind, dev, dev + n, d->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(d->stream);
    (void)hipFree(dev);
    if (e != hipSuccess) return fail(LBM_ERR_HIP, std::string("init_equilibrium: ") + hipGetErrorString(e));
    d->materialised = true;
    return LBM_OK;
}

int lbm_moments(lbm_domain* d, double* rho, double* u) {
    if (int rc = check_handle(d)) return rc;
    if (int rc = aa_settle(d)) return rc;
    if (int rc = materialise_boundaries(d)) return rc;
    const size_t n = static_cast<size_t>(d->L.xl) * d->L.yl * d->L.zl;
    double* dev = nullptr;
    HIPCHK(hipMalloc(&dev, 4 * n * sizeof(double)));
    hipError_t e = d->ops->moments(d->L, d->buf[d->cur], dev, dev + n, d->stream);
    if (e == hipSuccess && rho) e = hipMemcpyAsync(rho, dev, n * sizeof(double), hipMemcpyDeviceToHost, d->stream);
    if (e == hipSuccess && u) e = hipMemcpyAsync(u, dev + n, 3 * n * sizeof(double), hipMemcpyDeviceToHost, d->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(d->stream);
    (void)hipFree(dev);
    if (e != hipSuccess) return fail(LBM_ERR_HIP, std::string("moments: ") + hipGetErrorString(e));
    return LBM_OK;
}

int lbm_conservation(lbm_domain* d, double out[4]) {
    if (int rc = check_handle(d)) return rc;
    if (!d->finalized) return fail(LBM_ERR_STATE, "call lbm_finalize first");
    if (!out) return fail(LBM_ERR_ARG, "null output");
    if (int rc = aa_settle(d)) return rc;
    HIPCHK(d->ops->conservation(d->L, d->buf[d->cur], d->d_kind, d->d_partial, d->nred, d->d_out4,
                                d->stream));
    HIPCHK(hipMemcpyAsync(out, d->d_out4, 4 * sizeof(double), hipMemcpyDeviceToHost, d->stream));
    HIPCHK(hipStreamSynchronize(d->stream));
    return LBM_OK;
}

int lbm_copy_bandwidth(int device, uint64_t bytes, int reps, double out[2]) {
    if (!out || bytes < 16 || reps < 1 || reps > 1000) return fail(LBM_ERR_ARG, "bad copy arguments");
    HIPCHK(hipSetDevice(device));
    const int64_t n = static_cast<int64_t>(bytes / 16);
    void* a = nullptr;
    void* b = nullptr;
    if (hipMalloc(&a, n * 16) != hipSuccess) return fail(LBM_ERR_ALLOC, "hipMalloc failed");
    if (hipMalloc(&b, n * 16) != hipSuccess) {
        (void)hipFree(a);
        return fail(LBM_ERR_ALLOC, "hipMalloc failed");
    }
    hipEvent_t e0 = nullptr, e1 = nullptr;
    hipError_t e = hipMemset(a, 0, n * 16);
    if (e == hipSuccess) e = hipEventCreate(&e0);
    if (e == hipSuccess) e = hipEventCreate(&e1);
    const uint32_t grid = static_cast<uint32_t>((n + 255) / 256);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(k_copy16, dim3(grid), dim3(256), 0, nullptr, static_cast<const uint4*>(a),
                           static_cast<uint4*>(b), n);
        e = hipEventRecord(e0, nullptr);
    }
    for (int r = 0; e == hipSuccess && r < reps; ++r)
        hipLaunchKernelGGL(k_copy16, dim3(grid), dim3(256), 0, nullptr, static_cast<const uint4*>(a),
                           static_cast<uint4*>(b), n);
    if (e == hipSuccess) e = hipGetLastError();
    if (e == hipSuccess) e = hipEventRecord(e1, nullptr);
    if (e == hipSuccess) e = hipEventSynchronize(e1);
    float ms = 0;
    if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    (void)hipFree(a);
    (void)hipFree(b);
    if (e != hipSuccess) return fail(LBM_ERR_HIP, std::string("copy: ") + hipGetErrorString(e));
    out[0] = ms / reps;
    out[1] = 2.0 * static_cast<double>(n) * 16.0 / (out[0] * 1e-3) / 1e9;
    return LBM_OK;
}

int lbm_copy_planes_bandwidth(int device, int q, uint64_t plane_elems, int reps, double out[2]) {
    if (!out || (q != 15 && q != 19 && q != 27) || plane_elems < 2 || plane_elems % 2 || reps < 1 || reps > 1000)
        return fail(LBM_ERR_ARG, "bad plane copy arguments (q = 15, 19 or 27, even plane size)");
    auto launch = [&](const void* s, void* d, int64_t p2, uint32_t g) {
        if (q == 15)
            hipLaunchKernelGGL(k_copy_planes<15>, dim3(g), dim3(256), 0, nullptr, static_cast<const float2*>(s),
                               static_cast<float2*>(d), p2);
        else if (q == 19)
            hipLaunchKernelGGL(k_copy_planes<19>, dim3(g), dim3(256), 0, nullptr, static_cast<const float2*>(s),
                               static_cast<float2*>(d), p2);
        else
            hipLaunchKernelGGL(k_copy_planes<27>, dim3(g), dim3(256), 0, nullptr, static_cast<const float2*>(s),
                               static_cast<float2*>(d), p2);
    };
    HIPCHK(hipSetDevice(device));
    const int64_t plane2 = static_cast<int64_t>(plane_elems / 2);
    const size_t bytes = static_cast<size_t>(plane_elems) * 4 * q;
    void* a = nullptr;
    void* b = nullptr;
    if (hipMalloc(&a, bytes) != hipSuccess) return fail(LBM_ERR_ALLOC, "hipMalloc failed");
    if (hipMalloc(&b, bytes) != hipSuccess) {
        (void)hipFree(a);
        return fail(LBM_ERR_ALLOC, "hipMalloc failed");
    }
    hipEvent_t e0 = nullptr, e1 = nullptr;
    hipError_t e = hipMemset(a, 0, bytes);
    if (e == hipSuccess) e = hipEventCreate(&e0);
    if (e == hipSuccess) e = hipEventCreate(&e1);
    const uint32_t grid = static_cast<uint32_t>((plane2 + 255) / 256);
    if (e == hipSuccess) {
        launch(a, b, plane2, grid);
        e = hipEventRecord(e0, nullptr);
    }
    for (int r = 0; e == hipSuccess && r < reps; ++r) launch(a, b, plane2, grid);
    if (e == hipSuccess) e = hipGetLastError();
    if (e == hipSuccess) e = hipEventRecord(e1, nullptr);
    if (e == hipSuccess) e = hipEventSynchronize(e1);
    float ms = 0;
    if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    (void)hipFree(a);
    (void)hipFree(b);
    if (e != hipSuccess) return fail(LBM_ERR_HIP, std::string("plane copy: ") + hipGetErrorString(e));
    out[0] = ms / reps;
    out[1] = 2.0 * static_cast<double>(bytes) / (out[0] * 1e-3) / 1e9;
    return LBM_OK;
}

int lbm_get_info(lbm_domain* d, lbm_info* info) {
    if (!d || !info) return fail(LBM_ERR_ARG, "null argument");
    std::memset(info, 0, sizeof(*info));
    info->ncells_interior = static_cast<uint64_t>(d->L.xl) * d->L.yl * d->L.zl;
    info->ncells_fluid = d->n_fluid;
    info->ncells_near = static_cast<uint64_t>(d->nb);
    info->bytes_per_value = d->esize;
    info->field_bytes = d->field_bytes;
    info->device_bytes = d->device_bytes;
    info->vec = d->plan.V;
    info->block_x = d->plan.BX;
    info->block_y = d->plan.BY;
    info->nconds = d->nconds_dev;
    info->cell_tiles = (d->plan.cells || d->aa) ? 1 : 0;
    info->layout = d->aa ? LBM_LAYOUT_AA : LBM_LAYOUT_AB;
    info->steps_per_pass = (d->plan.pairs || d->plan.wpairs) ? 2 : 1;
    info->z_chunk = d->plan.tb_zc;
    return LBM_OK;
}

int lbm_set_index_width(lbm_domain* d, int bits) {
    if (int rc = check_handle(d)) return rc;
    if (bits != 32 && bits != 64) return fail(LBM_ERR_ARG, "index width must be 32 or 64");
    d->force64 = bits == 64;
    if (d->finalized) return make_plan(d);
    return LBM_OK;
}

int lbm_set_time_blocking(lbm_domain* d, int steps_per_pass, int z_chunk) {
    if (int rc = check_handle(d)) return rc;  // (planning queries the domain's device)
    if (steps_per_pass < 0 || steps_per_pass > 2)
        return fail(LBM_ERR_ARG, "steps_per_pass must be 0 (automatic), 1 or 2");
    if (z_chunk < 0) return fail(LBM_ERR_ARG, "z_chunk must be >= 0");
    const int op = d->tune_pass, oz = d->tune_zc;
    d->tune_pass = steps_per_pass;
    d->tune_zc = z_chunk;
    if (d->finalized) {
        if (int rc = make_plan(d)) {
            d->tune_pass = op;
            d->tune_zc = oz;
            make_plan(d);
            return rc;
        }
    }
    return LBM_OK;
}

}  // extern "C"
