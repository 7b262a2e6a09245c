// lbm_moments_noalloc (include/lbm_hip.h): lbm_moments with its device
// staging (rho, then u, of every interior cell: 4 doubles a cell) in a
// file-scope device array instead of a per-call allocation, for callers that
// cannot allocate device memory. Each GPU holds its own copy of the array;
// its address is looked up once per device with hipGetSymbolAddress and that
// address, never the array's name, is what kernels and copies are given.
// Every call on a device shares that device's array: calls that overlap on
// one GPU (two host threads at once, or work queued on other streams before
// a call's copies finish) race on it. A call writes every element it reads
// back, so nothing needs clearing between calls.
#include <hip/hip_runtime.h>

#include <cstddef>
#include <mutex>
#include <string>

#include "domain_impl.hpp"
#include "lbm_hip.h"

namespace {

constexpr size_t kNoAllocMaxCells = size_t{1} << 20;  // 32 MiB of staging per device
constexpr int kNoAllocMaxDevices = 64;

__device__ __align__(256) double scratch[4 * kNoAllocMaxCells];

std::mutex g_scratch_mutex;
double* g_scratch_ptr[kNoAllocMaxDevices] = {};

// This device's copy of scratch; the current device must be `device`.
int scratch_for(int device, double** out) {
    if (device < 0 || device >= kNoAllocMaxDevices) return fail(LBM_ERR_ARG, "device ordinal out of range");
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (!g_scratch_ptr[device]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return fail(LBM_ERR_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        g_scratch_ptr[device] = static_cast<double*>(p);
    }
    *out = g_scratch_ptr[device];
    return LBM_OK;
}

}  // namespace

extern "C" {

int lbm_moments_noalloc(lbm_domain* d, double* rho, double* u) {
    if (int rc = check_handle(d)) return rc;  // makes the domain's device current
    const size_t n = static_cast<size_t>(d->L.xl) * d->L.yl * d->L.zl;
    if (n > kNoAllocMaxCells)
        return fail(LBM_ERR_UNSUPPORTED, "lbm_moments_noalloc: " + std::to_string(n) +
                                             " interior cells exceed its staging of " +
                                             std::to_string(kNoAllocMaxCells) + " cells; use lbm_moments");
    double* dev = nullptr;
    if (int rc = scratch_for(d->desc.device, &dev)) return rc;
    if (int rc = aa_settle(d)) return rc;
    if (int rc = materialise_boundaries(d)) return rc;
    hipError_t e = d->ops->moments(d->L, d->buf[d->cur], dev, dev + n, d->stream);
    if (e == hipSuccess && rho) e = hipMemcpyAsync(rho, dev, n * sizeof(double), hipMemcpyDeviceToHost, d->stream);
    if (e == hipSuccess && u) e = hipMemcpyAsync(u, dev + n, 3 * n * sizeof(double), hipMemcpyDeviceToHost, d->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(d->stream);
    if (e != hipSuccess) return fail(LBM_ERR_HIP, std::string("moments_noalloc: ") + hipGetErrorString(e));
    return LBM_OK;
}

}  // extern "C"
