// Adjoint, objective and fp64 dual-number (forward-mode) kernels, D3Q15/19/27.
#include "kernels_adjoint.hpp"
#include "ops_adjoint.hpp"
#include "ops_impl.hpp"

namespace lbm_amd {

namespace {

template <class L, Mode M>
struct AdjImpl {
    using Calc = typename StorageOf<M>::Calc;

    static hipError_t backward_step(const AdjArgs& a, const StepPlan& p, bool brink, hipStream_t s) {
        if (p.has_bulk) {
            if (brink)
                hipLaunchKernelGGL((k_adj_bulk<L, M, true>), dim3(static_cast<uint32_t>(p.nblocks_bulk)),
                                   dim3(p.BX, p.BY), 0, s, a, p.nbx, p.nby);
            else
                hipLaunchKernelGGL((k_adj_bulk<L, M, false>), dim3(static_cast<uint32_t>(p.nblocks_bulk)),
                                   dim3(p.BX, p.BY), 0, s, a, p.nbx, p.nby);
        }
        if (a.nb > 0) {
            const uint32_t nblk = static_cast<uint32_t>((a.nb + 127) / 128);
            if (brink)
                hipLaunchKernelGGL((k_adj_near<L, M, true>), dim3(nblk), dim3(128), 0, s, a);
            else
                hipLaunchKernelGGL((k_adj_near<L, M, false>), dim3(nblk), dim3(128), 0, s, a);
        }
        return hipGetLastError();
    }

    template <Mode MM>
    static hipError_t objective_impl(const FieldLayout& Ly, const void* src, const uint8_t* kind, int obj,
                                     int plane, const double* target, double* partial, int nblocks,
                                     double* out, hipStream_t s) {
        hipLaunchKernelGGL((k_objective<L, MM>), dim3(nblocks), dim3(256), 0, s, Ly, src, kind, obj, plane,
                           target, partial);
        hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(256), 0, s, partial, nblocks, out);
        return hipGetLastError();
    }

    static hipError_t to_dual(const FieldLayout& Ly, const void* src, void* dst, hipStream_t s) {
        const uint32_t nblk = static_cast<uint32_t>((Ly.plane + 255) / 256);
        hipLaunchKernelGGL((k_to_dual<L, M>), dim3(nblk), dim3(256), 0, s, Ly, src,
                           static_cast<Dual<double>*>(dst), 0);
        return hipGetLastError();
    }

    static AdjOps make() {
        AdjOps o{};
        o.backward_step = &backward_step;
        o.objective = &objective_impl<M>;
        o.objective_dual = &objective_impl<Mode::DUAL64>;
        o.to_dual = &to_dual;
        o.dual_step = &OpsImpl<L, Mode::DUAL64>::template step_impl<1>;
        o.calc_bytes = sizeof(Calc);
        return o;
    }
};

const AdjOps kF32[3] = {AdjImpl<D3Q15, Mode::F32>::make(), AdjImpl<D3Q19, Mode::F32>::make(),
                        AdjImpl<D3Q27, Mode::F32>::make()};
const AdjOps kF64S[3] = {AdjImpl<D3Q15, Mode::F64S>::make(), AdjImpl<D3Q19, Mode::F64S>::make(),
                         AdjImpl<D3Q27, Mode::F64S>::make()};

}  // namespace

const AdjOps* adj_ops(int q, Mode m) {
    const int i = q == 15 ? 0 : (q == 19 ? 1 : (q == 27 ? 2 : -1));
    if (i < 0) return nullptr;
    if (m == Mode::F32) return &kF32[i];
    if (m == Mode::F64S) return &kF64S[i];
    return nullptr;
}

}  // namespace lbm_amd
