// C ABI of the differentiation path (include/lbm_hip.h, lbm_adj_* / lbm_jvp).
//
// A tape holds the initial field g^0, checkpoints g^{k}, g^{2k}, ... of one
// forward run (k = ceil(sqrt(N)) by default) and k segment buffers. The
// backward sweep walks the segments from last to first: it recomputes the
// segment's states from its checkpoint into the segment buffers, then runs
// the adjoint step backwards through them (k_adj_bulk / k_adj_near), which
// reads g^{t-1}, reads mu^{t+1}, writes mu^t and accumulates fp64 gradients.
// Extra forward work: N steps; memory: ~2 sqrt(N) fields.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "domain_impl.hpp"
#include "dual.hpp"
#include "lbm_hip.h"
#include "ops_adjoint.hpp"

using namespace lbm_amd;

struct lbm_tape {
    lbm_domain* d = nullptr;
    const AdjOps* ops = nullptr;
    int objective = 0;
    int plane_x = 0;
    uint64_t T = 0, k = 0, nck = 0;
    std::vector<void*> ckpt;  // ckpt[s] = g^{s k}
    std::vector<void*> seg;   // seg[j] = g^{s k + j}, j = 1..k-1 during segment s
    void* mu[2] = {nullptr, nullptr};
    double* d_target = nullptr;
    double* d_grad_alpha = nullptr;
    double* d_grad_wall = nullptr;
    bool start_mat = true;
    bool forward_done = false;
    StepPlan p1{};  // one cell per thread
    double timing[4] = {0, 0, 0, 0};
};

namespace {

void tape_free(lbm_tape* t) {
    for (void* p : t->ckpt) (void)hipFree(p);
    for (void* p : t->seg) (void)hipFree(p);
    (void)hipFree(t->mu[0]);
    (void)hipFree(t->mu[1]);
    (void)hipFree(t->d_target);
    (void)hipFree(t->d_grad_alpha);
    (void)hipFree(t->d_grad_wall);
    t->ckpt.clear();
    t->seg.clear();
}

StepPlan plan_v1(const lbm_domain* d) {
    StepPlan p = d->plan;
    p.cells = false;  // the adjoint and dual steps run the row kernels
    p.V = 1;
    int bx = 1;
    while (bx < d->L.xl && bx < 256) bx *= 2;
    p.BX = bx;
    p.BY = 256 / bx;
    p.nbx = (d->L.xl + p.BX - 1) / p.BX;
    p.nby = (d->L.yl + p.BY - 1) / p.BY;
    p.nblocks_bulk = static_cast<int64_t>(p.nbx) * p.nby * d->L.zl;
    return p;
}

int forward_step(lbm_domain* d, const void* src, void* dst, bool mat) {
    const StepArgs a = d->args(src, dst, mat ? 1 : 0);
    const hipError_t e = d->ops->step(a, d->plan, d->stream, nullptr);
    if (e != hipSuccess) return fail(LBM_ERR_HIP, std::string("forward step: ") + hipGetErrorString(e));
    return LBM_OK;
}

// Device-resident porosity and gradients: interior arrays (x fastest) <->
// the padded element layout, in the mode's arithmetic type for alpha.
template <class C>
__global__ void __launch_bounds__(256) k_alpha_scatter(FieldLayout L, const double* __restrict__ in, C* __restrict__ out) {
    const int64_t i = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
    if (i >= static_cast<int64_t>(L.xl) * L.yl * L.zl) return;
    const int x = 1 + static_cast<int>(i % L.xl);
    const int64_t r = i / L.xl;
    const int y = 1 + static_cast<int>(r % L.yl), z = 1 + static_cast<int>(r / L.yl);
    out[cell_elem(L, x, y, z)] = static_cast<C>(in[i]);
}

__global__ void __launch_bounds__(256) k_grad_gather(FieldLayout L, const double* __restrict__ g, double dJ,
                                                     double* __restrict__ out) {
    const int64_t i = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
    if (i >= static_cast<int64_t>(L.xl) * L.yl * L.zl) return;
    const int x = 1 + static_cast<int>(i % L.xl);
    const int64_t r = i / L.xl;
    const int y = 1 + static_cast<int>(r % L.yl), z = 1 + static_cast<int>(r / L.yl);
    out[i] = dJ * g[cell_elem(L, x, y, z)];
}

uint32_t interior_blocks(const FieldLayout& L) {
    return static_cast<uint32_t>((static_cast<int64_t>(L.xl) * L.yl * L.zl + 255) / 256);
}

int upload_alpha(lbm_domain* d, const double* alpha) {
    if (!d->d_alpha) return fail(LBM_ERR_STATE, "the domain has no Brinkman field (lbm_set_brinkman before finalize)");
    const std::vector<char> a = alpha_in_layout(d, alpha);
    HIPCHK(hipMemcpyAsync(d->d_alpha, a.data(), a.size(), hipMemcpyHostToDevice, d->stream));
    HIPCHK(hipStreamSynchronize(d->stream));
    return LBM_OK;
}

}  // namespace

extern "C" {

int lbm_adj_create(lbm_domain* d, const lbm_adj_desc* desc, lbm_tape** out) {
    if (int rc = check_handle(d)) return rc;
    if (!desc || !out) return fail(LBM_ERR_ARG, "null argument");
    *out = nullptr;
    if (!d->finalized) return fail(LBM_ERR_STATE, "finalize the domain before creating a tape");
    if (d->aa) return fail(LBM_ERR_UNSUPPORTED, "the adjoint needs the two-buffer layout");
    const AdjOps* ops = adj_ops(d->Q, d->mode);
    if (!ops) return fail(LBM_ERR_UNSUPPORTED, "adjoint needs F32 or F64S storage");
    for (const HostCond& c : d->conds)
        if (c.kind == LBM_BC_FREESLIP)
            return fail(LBM_ERR_UNSUPPORTED,
                        "adjoint supports NoSlip, MovingWall, Inflow, Outflow and Pressure boundaries");
    if (desc->objective != LBM_OBJ_PLANE_VELOCITY_MISFIT && desc->objective != LBM_OBJ_KINETIC_ENERGY)
        return fail(LBM_ERR_ARG, "unknown objective");
    if (desc->objective == LBM_OBJ_PLANE_VELOCITY_MISFIT &&
        (desc->plane_x < 1 || desc->plane_x > d->L.xl || !desc->target_u))
        return fail(LBM_ERR_ARG, "plane objective needs plane_x in [1, xl] and target_u");
    if (desc->nsteps < 1 || desc->nsteps > 10000000) return fail(LBM_ERR_ARG, "nsteps in [1, 1e7]");

    lbm_tape* t = new lbm_tape();
    t->d = d;
    t->ops = ops;
    t->objective = desc->objective;
    t->plane_x = desc->plane_x;
    t->T = desc->nsteps;
    t->k = desc->checkpoint_every ? desc->checkpoint_every
                                  : static_cast<uint64_t>(std::ceil(std::sqrt(static_cast<double>(t->T))));
    if (t->k > t->T) t->k = t->T;
    t->nck = (t->T + t->k - 1) / t->k;
    t->p1 = plan_v1(d);
    auto bail = [&](int code, const std::string& m) {
        tape_free(t);
        delete t;
        return fail(code, m);
    };
    t->ckpt.assign(t->nck, nullptr);
    for (auto& p : t->ckpt)
        if (hipMalloc(&p, d->field_bytes) != hipSuccess) return bail(LBM_ERR_ALLOC, "checkpoint allocation failed");
    t->seg.assign(t->k, nullptr);
    for (uint64_t j = 1; j < t->k; ++j)
        if (hipMalloc(&t->seg[j], d->field_bytes) != hipSuccess) return bail(LBM_ERR_ALLOC, "segment allocation failed");
    const size_t mu_bytes = static_cast<size_t>(d->L.plane) * d->Q * ops->calc_bytes;
    for (auto& m : t->mu)
        if (hipMalloc(&m, mu_bytes) != hipSuccess) return bail(LBM_ERR_ALLOC, "adjoint buffer allocation failed");
    if (hipMalloc(&t->d_grad_alpha, sizeof(double) * d->L.plane) != hipSuccess ||
        hipMalloc(&t->d_grad_wall, sizeof(double) * 3 * (d->nb ? d->nb : 1)) != hipSuccess)
        return bail(LBM_ERR_ALLOC, "gradient allocation failed");
    const size_t ntarget = static_cast<size_t>(d->L.yl) * d->L.zl * 3;
    if (hipMalloc(&t->d_target, sizeof(double) * ntarget) != hipSuccess) return bail(LBM_ERR_ALLOC, "target allocation failed");
    hipError_t e = hipMemsetAsync(t->d_target, 0, sizeof(double) * ntarget, d->stream);
    if (e == hipSuccess && desc->target_u)
        e = hipMemcpyAsync(t->d_target, desc->target_u, sizeof(double) * ntarget, hipMemcpyHostToDevice, d->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(t->ckpt[0], d->buf[d->cur], d->field_bytes, hipMemcpyDeviceToDevice, d->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(d->stream);
    if (e != hipSuccess) return bail(LBM_ERR_HIP, std::string("tape setup: ") + hipGetErrorString(e));
    t->start_mat = d->materialised;
    *out = t;
    return LBM_OK;
}

int lbm_adj_destroy(lbm_tape* t) {
    if (!t) return LBM_OK;
    (void)hipSetDevice(t->d->desc.device);
    (void)hipStreamSynchronize(t->d->stream);
    tape_free(t);
    delete t;
    return LBM_OK;
}

int lbm_adj_forward(lbm_tape* t, const double* alpha, double* J) {
    if (!t) return fail(LBM_ERR_ARG, "null tape");
    lbm_domain* d = t->d;
    if (int rc = check_handle(d)) return rc;
    if (alpha)
        if (int rc = upload_alpha(d, alpha)) return rc;
    Events ev(2);
    HIPCHK(ev.create());
    HIPCHK(hipMemcpyAsync(d->buf[0], t->ckpt[0], d->field_bytes, hipMemcpyDeviceToDevice, d->stream));
    d->cur = 0;
    HIPCHK(hipEventRecord(ev[0], d->stream));
    for (uint64_t s = 1; s <= t->T; ++s) {
        if (int rc = forward_step(d, d->buf[d->cur], d->buf[d->cur ^ 1], s == 1 && t->start_mat)) return rc;
        d->cur ^= 1;
        if (s % t->k == 0 && s / t->k < t->nck)
            HIPCHK(hipMemcpyAsync(t->ckpt[s / t->k], d->buf[d->cur], d->field_bytes, hipMemcpyDeviceToDevice, d->stream));
    }
    HIPCHK(hipEventRecord(ev[1], d->stream));
    d->materialised = false;
    HIPCHK(t->ops->objective(d->L, d->buf[d->cur], d->d_kind, t->objective, t->plane_x, t->d_target,
                             d->d_partial, d->nred, d->d_out4, d->stream));
    double out[4];
    HIPCHK(hipMemcpyAsync(out, d->d_out4, sizeof(out), hipMemcpyDeviceToHost, d->stream));
    HIPCHK(hipStreamSynchronize(d->stream));
    float ms = 0;
    HIPCHK(hipEventElapsedTime(&ms, ev[0], ev[1]));
    t->timing[2] = ms;
    t->timing[3] = static_cast<double>(t->T);
    if (J) *J = out[0];
    t->forward_done = true;
    return LBM_OK;
}

int lbm_adj_forward_device(lbm_tape* t, const double* d_alpha, double* J) {
    if (!t) return fail(LBM_ERR_ARG, "null tape");
    lbm_domain* d = t->d;
    if (int rc = check_handle(d)) return rc;
    if (d_alpha) {
        if (!d->d_alpha)
            return fail(LBM_ERR_STATE, "the domain has no Brinkman field (lbm_set_brinkman before finalize)");
        if (d->mode == Mode::F64S)
            hipLaunchKernelGGL(k_alpha_scatter<double>, dim3(interior_blocks(d->L)), dim3(256), 0, d->stream, d->L,
                               d_alpha, static_cast<double*>(d->d_alpha));
        else
            hipLaunchKernelGGL(k_alpha_scatter<float>, dim3(interior_blocks(d->L)), dim3(256), 0, d->stream, d->L,
                               d_alpha, static_cast<float*>(d->d_alpha));
        HIPCHK(hipGetLastError());
    }
    return lbm_adj_forward(t, nullptr, J);
}

int lbm_adj_backward_device(lbm_tape* t, double dJ, double* d_grad_alpha, double grad_wall[3]) {
    if (!t) return fail(LBM_ERR_ARG, "null tape");
    lbm_domain* d = t->d;
    if (int rc = check_handle(d)) return rc;
    if (d_grad_alpha && !d->d_alpha) return fail(LBM_ERR_STATE, "no Brinkman field: no porosity gradient");
    if (int rc = lbm_adj_backward(t, dJ, nullptr, grad_wall)) return rc;
    if (d_grad_alpha) {
        hipLaunchKernelGGL(k_grad_gather, dim3(interior_blocks(d->L)), dim3(256), 0, d->stream, d->L,
                           t->d_grad_alpha, dJ, d_grad_alpha);
        HIPCHK(hipGetLastError());
        HIPCHK(hipStreamSynchronize(d->stream));
    }
    return LBM_OK;
}

int lbm_adj_backward(lbm_tape* t, double dJ, double* grad_alpha, double grad_wall[3]) {
    if (!t) return fail(LBM_ERR_ARG, "null tape");
    lbm_domain* d = t->d;
    if (int rc = check_handle(d)) return rc;
    if (!t->forward_done) return fail(LBM_ERR_STATE, "lbm_adj_backward needs a preceding lbm_adj_forward");
    const bool brink = d->d_alpha != nullptr;
    if (grad_alpha && !brink) return fail(LBM_ERR_STATE, "no Brinkman field: no porosity gradient");
    HIPCHK(hipMemsetAsync(t->d_grad_alpha, 0, sizeof(double) * d->L.plane, d->stream));
    HIPCHK(hipMemsetAsync(t->d_grad_wall, 0, sizeof(double) * 3 * (d->nb ? d->nb : 1), d->stream));
    const void* gT = d->buf[d->cur];

    AdjArgs a{};
    a.L = d->L;
    a.gcur = gT;
    a.kind = d->d_kind;
    a.blist = d->d_blist;
    a.nb = d->nb;
    a.conds = d->d_conds;
    a.tau = d->desc.tau;
    a.omega = 1.0 / d->desc.tau;
    a.alpha = d->d_alpha;
    for (int i = 0; i < 3; ++i) a.g[i] = d->g[i];
    a.grad_alpha = t->d_grad_alpha;
    a.grad_wall = t->d_grad_wall;
    a.objective = t->objective;
    a.plane_x = t->plane_x;
    a.target = t->d_target;

    Events ev(3 * t->nck);
    HIPCHK(ev.create());
    for (uint64_t si = t->nck; si-- > 0;) {
        const uint64_t start = si * t->k;
        const uint64_t end = std::min(start + t->k, t->T);
        auto state = [&](uint64_t s) -> void* { return s == start ? t->ckpt[si] : t->seg[s - start]; };
        HIPCHK(hipEventRecord(ev[3 * si + 0], d->stream));
        for (uint64_t s = start + 1; s < end; ++s)
            if (int rc = forward_step(d, state(s - 1), state(s), s == 1 && t->start_mat)) return rc;
        HIPCHK(hipEventRecord(ev[3 * si + 1], d->stream));
        for (uint64_t s = end; s > start; --s) {
            a.gprev = state(s - 1);
            a.terminal = (s == t->T) ? 1 : 0;
            a.mu_in = t->mu[(s + 1) & 1];
            a.mu_out = t->mu[s & 1];
            a.materialised = (s == 1 && t->start_mat) ? 1 : 0;
            HIPCHK(t->ops->backward_step(a, t->p1, brink, d->stream));
        }
        HIPCHK(hipEventRecord(ev[3 * si + 2], d->stream));
    }
    HIPCHK(hipStreamSynchronize(d->stream));
    double rec = 0, adj = 0;
    for (uint64_t si = 0; si < t->nck; ++si) {
        float ms;
        HIPCHK(hipEventElapsedTime(&ms, ev[3 * si], ev[3 * si + 1]));
        rec += ms;
        HIPCHK(hipEventElapsedTime(&ms, ev[3 * si + 1], ev[3 * si + 2]));
        adj += ms;
    }
    t->timing[0] = rec;
    t->timing[1] = adj;

    if (grad_alpha) {
        std::vector<double> g(static_cast<size_t>(d->L.plane));
        HIPCHK(hipMemcpy(g.data(), t->d_grad_alpha, sizeof(double) * g.size(), hipMemcpyDeviceToHost));
        int64_t i = 0;
        for (int z = 1; z <= d->L.zl; ++z)
            for (int y = 1; y <= d->L.yl; ++y)
                for (int x = 1; x <= d->L.xl; ++x, ++i) grad_alpha[i] = dJ * g[cell_elem(d->L, x, y, z)];
    }
    if (grad_wall) {
        std::vector<double> g(static_cast<size_t>(3 * (d->nb ? d->nb : 1)), 0.0);
        if (d->nb) HIPCHK(hipMemcpy(g.data(), t->d_grad_wall, sizeof(double) * 3 * d->nb, hipMemcpyDeviceToHost));
        double s[3] = {0, 0, 0};
        for (int64_t i = 0; i < d->nb; ++i)
            for (int k = 0; k < 3; ++k) s[k] += g[3 * i + k];
        for (int k = 0; k < 3; ++k) grad_wall[k] = dJ * s[k];
    }
    return LBM_OK;
}

int lbm_jvp(lbm_tape* t, const double* alpha, const double* dalpha, double* J, double* dJ) {
    if (!t) return fail(LBM_ERR_ARG, "null tape");
    lbm_domain* d = t->d;
    if (int rc = check_handle(d)) return rc;
    if (!d->d_alpha) return fail(LBM_ERR_STATE, "the domain has no Brinkman field");
    if (!dalpha) return fail(LBM_ERR_ARG, "null direction");
    const FieldLayout& L = d->L;
    // porosity as (value, tangent) pairs in the padded layout
    std::vector<Dual<double>> ad(static_cast<size_t>(L.plane), Dual<double>(0.0, 0.0));
    {
        std::vector<double> cur;
        const double* av = alpha;
        if (!av) {  // current device porosity
            const size_t cs = (d->mode == Mode::F64S) ? 8 : 4;
            std::vector<char> raw(static_cast<size_t>(L.plane) * cs);
            HIPCHK(hipMemcpy(raw.data(), d->d_alpha, raw.size(), hipMemcpyDeviceToHost));
            cur.resize(static_cast<size_t>(L.xl) * L.yl * L.zl);
            int64_t i = 0;
            for (int z = 1; z <= L.zl; ++z)
                for (int y = 1; y <= L.yl; ++y)
                    for (int x = 1; x <= L.xl; ++x, ++i) {
                        const int64_t e = cell_elem(L, x, y, z);
                        cur[i] = cs == 8 ? reinterpret_cast<double*>(raw.data())[e]
                                         : reinterpret_cast<float*>(raw.data())[e];
                    }
            av = cur.data();
        }
        int64_t i = 0;
        for (int z = 1; z <= L.zl; ++z)
            for (int y = 1; y <= L.yl; ++y)
                for (int x = 1; x <= L.xl; ++x, ++i) ad[cell_elem(L, x, y, z)] = Dual<double>(av[i], dalpha[i]);
    }
    const size_t fbytes = static_cast<size_t>(L.plane) * d->Q * sizeof(Dual<double>);
    void* buf[2] = {nullptr, nullptr};
    void* dal = nullptr;
    auto cleanup = [&] {
        (void)hipFree(buf[0]);
        (void)hipFree(buf[1]);
        (void)hipFree(dal);
    };
    if (hipMalloc(&buf[0], fbytes) != hipSuccess || hipMalloc(&buf[1], fbytes) != hipSuccess ||
        hipMalloc(&dal, ad.size() * sizeof(Dual<double>)) != hipSuccess) {
        cleanup();
        return fail(LBM_ERR_ALLOC, "dual buffers");
    }
    hipError_t e = hipMemcpyAsync(dal, ad.data(), ad.size() * sizeof(Dual<double>), hipMemcpyHostToDevice, d->stream);
    if (e == hipSuccess) e = t->ops->to_dual(L, t->ckpt[0], buf[0], d->stream);
    StepPlan p = t->p1;
    p.brink = true;
    int cur = 0;
    for (uint64_t s = 1; e == hipSuccess && s <= t->T; ++s) {
        StepArgs a = d->args(buf[cur], buf[cur ^ 1], (s == 1 && t->start_mat) ? 1 : 0);
        a.alpha = dal;
        e = t->ops->dual_step(a, p, d->stream, nullptr);
        cur ^= 1;
    }
    double out[4] = {0, 0, 0, 0};
    if (e == hipSuccess)
        e = t->ops->objective_dual(L, buf[cur], d->d_kind, t->objective, t->plane_x, t->d_target, d->d_partial,
                                   d->nred, d->d_out4, d->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(out, d->d_out4, sizeof(out), hipMemcpyDeviceToHost, d->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(d->stream);
    cleanup();
    if (e != hipSuccess) return fail(LBM_ERR_HIP, std::string("jvp: ") + hipGetErrorString(e));
    if (J) *J = out[0];
    if (dJ) *dJ = out[1];
    return LBM_OK;
}

int lbm_adj_timing(lbm_tape* t, double out[4]) {
    if (!t || !out) return fail(LBM_ERR_ARG, "null argument");
    for (int i = 0; i < 4; ++i) out[i] = t->timing[i];
    return LBM_OK;
}

}  // extern "C"
