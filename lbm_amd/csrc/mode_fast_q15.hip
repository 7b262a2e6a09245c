// Shifted-storage kernels (F32, F16, BF16 with fp32 arithmetic; F64S with
// fp64 arithmetic) for D3Q15.
#include "ops_impl.hpp"

namespace lbm_amd {

namespace {
using Lat = D3Q15;
// fp32: 32 x 8 thread blocks (measured 1 % faster than full-row blocks)
const ModeOps kF32 = OpsImpl<Lat, Mode::F32>::make<1, 2, 4>(2, 32);
const ModeOps kF16 = OpsImpl<Lat, Mode::F16>::make<2, 4, 8>(4);
const ModeOps kBF16 = OpsImpl<Lat, Mode::BF16>::make<2, 4, 8>(4);
const ModeOps kF64S = OpsImpl<Lat, Mode::F64S>::make<1, 2>(1);
}  // namespace

const ModeOps* ops_fast_q15(Mode m) {
    switch (m) {
        case Mode::F32: return &kF32;
        case Mode::F16: return &kF16;
        case Mode::BF16: return &kBF16;
        case Mode::F64S: return &kF64S;
        default: return nullptr;
    }
}

}  // namespace lbm_amd
