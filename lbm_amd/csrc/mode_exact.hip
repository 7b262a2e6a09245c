// EXACT64 (reference fp64 parity mode) kernels for D3Q15/19/27.
// Built with -ffp-contract=off: the reference's products and sums must not
// be fused into FMAs, or the bits differ from the reference run.
#include "ops_impl.hpp"

namespace lbm_amd {

namespace {
const ModeOps kQ15 = OpsImpl<D3Q15, Mode::EXACT64>::make<1, 2>(1);
const ModeOps kQ19 = OpsImpl<D3Q19, Mode::EXACT64>::make<1, 2>(1);
const ModeOps kQ27 = OpsImpl<D3Q27, Mode::EXACT64>::make<1, 2>(1);
}  // namespace

const ModeOps* ops_exact(int q) {
    switch (q) {
        case 15: return &kQ15;
        case 19: return &kQ19;
        case 27: return &kQ27;
        default: return nullptr;
    }
}

}  // namespace lbm_amd
