// Launchers behind ModeOps, instantiated per (lattice, mode) by the mode
// translation units.
#pragma once

#include "kernels.hpp"
#include "kernels_aa.hpp"
#include "kernels_fields.hpp"
#include "kernels_tb.hpp"
#include "kernels_wall2.hpp"
#include "ops.hpp"

namespace lbm_amd {

template <class L, Mode M>
struct OpsImpl {
    template <int V, bool FLAGS, bool BRINK>
    static void launch_bulk(const StepArgs& a, const StepPlan& p, hipStream_t s) {
        dim3 block(p.BX, p.BY, 1);
        hipLaunchKernelGGL((k_bulk<L, M, V, FLAGS, BRINK>), dim3(static_cast<uint32_t>(p.nblocks_bulk)),
                           block, 0, s, a, p.nbx, p.nby);
    }

    template <int V>
    static bool try_bulk(const StepArgs& a, const StepPlan& p, hipStream_t s) {
        if (p.V != V) return false;
        if (p.brink) {
            if (p.flags)
                launch_bulk<V, true, true>(a, p, s);
            else
                launch_bulk<V, false, true>(a, p, s);
        } else {
            if (p.flags)
                launch_bulk<V, true, false>(a, p, s);
            else
                launch_bulk<V, false, false>(a, p, s);
        }
        return true;
    }

    template <bool FS, bool B32>
    static void launch_cells(const StepArgs& a, const StepPlan& p, hipStream_t s) {
        hipLaunchKernelGGL((k_cells<L, M, FS, B32>), dim3(static_cast<uint32_t>(p.nblocks_cells)), dim3(256), 0, s,
                           a, p.tpp, p.rxl);
    }

    template <int... Vs>
    static hipError_t step_impl(const StepArgs& a, const StepPlan& p, hipStream_t s, EventLog* log) {
        if (p.cells) {
            if (p.freeslip) {
                if (p.idx32)
                    launch_cells<true, true>(a, p, s);
                else
                    launch_cells<true, false>(a, p, s);
            } else {
                if (p.idx32)
                    launch_cells<false, true>(a, p, s);
                else
                    launch_cells<false, false>(a, p, s);
            }
            if (log) log->mark(s, 0);
            return hipGetLastError();
        }
        if (p.has_bulk) {
            bool done = (try_bulk<Vs>(a, p, s) || ...);
            if (!done) return hipErrorInvalidValue;
            if (log) log->mark(s, 0);
        }
        if (a.nb > 0) {
            const int bs = 128;
            const int64_t nblk = (a.nb + bs - 1) / bs;
            hipLaunchKernelGGL((k_near<L, M>), dim3(static_cast<uint32_t>(nblk)), dim3(bs), 0, s, a);
            if (log) log->mark(s, 1);
        }
        return hipGetLastError();
    }

    // the two-step pass: 4- and 2-byte storage whose LDS slots fit 64 KiB
    static constexpr bool tb_built() {
        return (M == Mode::F32 || M == Mode::F16 || M == Mode::BF16) && tb_lds_bytes<L, M>() <= 65536;
    }
    static hipError_t step2_impl(const StepArgs& a, const StepPlan& p, hipStream_t s, EventLog* log) {
        if constexpr (tb_built()) {
            const dim3 grid(static_cast<uint32_t>(p.tb_nblocks));
            if (p.brink)
                hipLaunchKernelGGL((k_bulk2<L, M, true>), grid, dim3(256), 0, s, a, p.tb_ntx, p.tb_nty, p.tb_zc,
                                   p.tb_nxcd);
            else
                hipLaunchKernelGGL((k_bulk2<L, M, false>), grid, dim3(256), 0, s, a, p.tb_ntx, p.tb_nty, p.tb_zc,
                                   p.tb_nxcd);
            if (log) log->mark(s, 0);
            return hipGetLastError();
        } else {
            return hipErrorInvalidValue;
        }
    }

    // the walled pass: fp32 storage (boundary values cross memory and LDS in
    // the arithmetic type) whose LDS slots fit 64 KiB
    static constexpr bool wall2_built() {
        return (M == Mode::F32 || M == Mode::F16 || M == Mode::BF16) &&
               static_cast<size_t>(TbSlots<L>::NSLOT) * TB_CELLS * sizeof(typename Io<M, L>::T) <= 65536;
    }
    template <int RULES>
    static void launch_wall2(const StepArgs& a, const StepPlan& p, hipStream_t s) {
        if (a.nb > 0) {
            // boundary links of the source buffer, from its fluid cells
            StepArgs e = a;
            e.dst = const_cast<void*>(a.src);
            hipLaunchKernelGGL((k_emit_rules<L, M, RULES>), dim3(static_cast<uint32_t>((a.nb + 255) / 256)), dim3(256),
                               0, s, e);
        }
        hipLaunchKernelGGL((k_wall2<L, M, RULES>), dim3(static_cast<uint32_t>(p.tb_nblocks)), dim3(256), 0, s, a,
                           p.tb_ntx, p.tb_nty, p.tb_zc, p.tb_nxcd);
    }
    static hipError_t step2w_impl(const StepArgs& a, const StepPlan& p, hipStream_t s, EventLog* log) {
        if constexpr (wall2_built()) {
            if (p.wrules == (RULE_MOVINGWALL))
                launch_wall2<RULE_MOVINGWALL>(a, p, s);
            else
                launch_wall2<RULES_ALL>(a, p, s);
            if (log) log->mark(s, 0);
            return hipGetLastError();
        } else {
            return hipErrorInvalidValue;
        }
    }
    static int step2w_occupancy(int rules) {
        int n = 0;
        if constexpr (wall2_built()) {
            const void* k = rules == RULE_MOVINGWALL ? reinterpret_cast<const void*>(&k_wall2<L, M, RULE_MOVINGWALL>)
                                                     : reinterpret_cast<const void*>(&k_wall2<L, M, RULES_ALL>);
            if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k, 256, 0) != hipSuccess) n = 0;
        }
        return n;
    }

    static int step2_occupancy(bool brink) {
        int n = 0;
        if constexpr (tb_built()) {
            const hipError_t e = brink ? hipOccupancyMaxActiveBlocksPerMultiprocessor(
                                             &n, reinterpret_cast<const void*>(&k_bulk2<L, M, true>), 256, 0)
                                       : hipOccupancyMaxActiveBlocksPerMultiprocessor(
                                             &n, reinterpret_cast<const void*>(&k_bulk2<L, M, false>), 256, 0);
            if (e != hipSuccess) n = 0;
        }
        return n;
    }

    static hipError_t get_pdf(const FieldLayout& Ly, const void* src, double* out, int nx, int ny,
                              int nz, hipStream_t s) {
        const int64_t n = static_cast<int64_t>(nx) * ny * nz;
        const int bs = 256;
        hipLaunchKernelGGL((k_get_pdf<L, M>), dim3(static_cast<uint32_t>((n + bs - 1) / bs)),
                           dim3(bs), 0, s, Ly, src, out, GhostDims{nx, ny, nz});
        return hipGetLastError();
    }
    static hipError_t set_pdf(const FieldLayout& Ly, void* dst, const double* in, int nx, int ny,
                              int nz, hipStream_t s) {
        const int64_t n = static_cast<int64_t>(nx) * ny * nz;
        const int bs = 256;
        hipLaunchKernelGGL((k_set_pdf<L, M>), dim3(static_cast<uint32_t>((n + bs - 1) / bs)),
                           dim3(bs), 0, s, Ly, dst, in, GhostDims{nx, ny, nz});
        return hipGetLastError();
    }
    static hipError_t moments(const FieldLayout& Ly, const void* src, double* rho, double* u,
                              hipStream_t s) {
        const int64_t n = static_cast<int64_t>(Ly.xl) * Ly.yl * Ly.zl;
        const int bs = 256;
        hipLaunchKernelGGL((k_moments<L, M>), dim3(static_cast<uint32_t>((n + bs - 1) / bs)),
                           dim3(bs), 0, s, Ly, src, rho, u);
        return hipGetLastError();
    }
    static hipError_t conservation(const FieldLayout& Ly, const void* src, const uint8_t* kind,
                                   double* partial, int nblocks, double* out, hipStream_t s) {
        hipLaunchKernelGGL((k_conservation<L, M>), dim3(nblocks), dim3(256), 0, s, Ly, src, kind,
                           partial);
        hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(256), 0, s, partial, nblocks, out);
        return hipGetLastError();
    }
    static hipError_t init_eq(const FieldLayout& Ly, void* dst, const uint8_t* kind,
                              const double* rho, const double* u, hipStream_t s) {
        const int64_t n = static_cast<int64_t>(Ly.xl) * Ly.yl * Ly.zl;
        const int bs = 256;
        hipLaunchKernelGGL((k_init_eq<L, M>), dim3(static_cast<uint32_t>((n + bs - 1) / bs)),
                           dim3(bs), 0, s, Ly, dst, kind, rho, u);
        return hipGetLastError();
    }
    static hipError_t materialise(const StepArgs& a, int nx, int ny, int nz, hipStream_t s) {
        const int64_t n = static_cast<int64_t>(nx) * ny * nz;
        const int bs = 128;
        hipLaunchKernelGGL((k_materialise<L, M>), dim3(static_cast<uint32_t>((n + bs - 1) / bs)),
                           dim3(bs), 0, s, a, GhostDims{nx, ny, nz});
        return hipGetLastError();
    }

    // phase 0: E -> O step, 1: O -> E step, 2: O -> E swap without a collision
    template <bool B32>
    static void launch_aa(const StepArgs& a, const StepPlan& p, int phase, hipStream_t s) {
        const dim3 grid(static_cast<uint32_t>(p.nblocks_cells));
        if (phase == 0)
            hipLaunchKernelGGL((k_aa_eo<L, M, B32>), grid, dim3(256), 0, s, a, p.tpp, p.rxl);
        else if (phase == 1)
            hipLaunchKernelGGL((k_aa_oe<L, M, B32>), grid, dim3(256), 0, s, a, p.tpp, p.rxl);
        else
            hipLaunchKernelGGL((k_aa_swap<L, M, B32>), grid, dim3(256), 0, s, a, p.tpp, p.rxl);
    }
    static hipError_t aa_step(const StepArgs& a, const StepPlan& p, int phase, hipStream_t s) {
        if (p.idx32)
            launch_aa<true>(a, p, phase, s);
        else
            launch_aa<false>(a, p, phase, s);
        return hipGetLastError();
    }
    static hipError_t aa_swap(const StepArgs& a, const StepPlan& p, hipStream_t s) {
        if (p.idx32)
            launch_aa<true>(a, p, 2, s);
        else
            launch_aa<false>(a, p, 2, s);
        return hipGetLastError();
    }

    template <int V0, int... Vs>
    static ModeOps make(int default_vec, int default_bx = 0) {
        ModeOps o{};
        o.mode = M;
        o.q = L::Q;
        o.default_vec = default_vec;
        o.default_bx = default_bx;
        const int v[] = {V0, Vs...};
        for (int i = 0; i < 4; ++i) o.vecs[i] = i < static_cast<int>(sizeof(v) / sizeof(int)) ? v[i] : 0;
        o.step = &step_impl<V0, Vs...>;
        o.step2 = tb_built() ? &step2_impl : nullptr;
        o.step2_blocks_per_cu = &step2_occupancy;
        o.step2w = wall2_built() ? &step2w_impl : nullptr;
        o.step2w_blocks_per_cu = &step2w_occupancy;
        o.get_pdf = &get_pdf;
        o.set_pdf = &set_pdf;
        o.moments = &moments;
        o.conservation = &conservation;
        o.init_eq = &init_eq;
        o.materialise = &materialise;
        o.aa_step = &aa_step;
        o.aa_swap = &aa_swap;
        return o;
    }
};

}  // namespace lbm_amd
