// Internal host-side state of a domain, shared by the C ABI translation
// units (capi.hip: domain lifecycle and stepping; capi_adjoint.hip:
// differentiation). Not part of the public interface.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "domain_view.hpp"
#include "lbm_hip.h"
#include "lbm_hip/lattice.hpp"
#include "ops.hpp"
#include "storage.hpp"

namespace lbm_amd {

// Message of the last failure on this thread (lbm_last_error).
std::string& last_error();

inline int fail(int code, const std::string& msg) {
    last_error() = msg;
    return code;
}

#define HIPCHK(expr)                                                                  \
    do {                                                                              \
        hipError_t e_ = (expr);                                                       \
        if (e_ != hipSuccess)                                                         \
            return fail(LBM_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

// HIP events owned by a scope: destroyed when it ends, error returns included.
class Events {
public:
    explicit Events(size_t n) : ev_(n, nullptr) {}
    ~Events() {
        for (hipEvent_t e : ev_)
            if (e) (void)hipEventDestroy(e);
    }
    Events(const Events&) = delete;
    Events& operator=(const Events&) = delete;
    hipError_t create() {
        for (hipEvent_t& e : ev_) {
            const hipError_t r = hipEventCreate(&e);
            if (r != hipSuccess) {
                e = nullptr;
                return r;
            }
        }
        return hipSuccess;
    }
    hipEvent_t operator[](size_t i) const { return ev_[i]; }
    hipEvent_t* data() { return ev_.data(); }

private:
    std::vector<hipEvent_t> ev_;
};

// Device memory owned by a scope: freed when it ends, error returns included.
struct DeviceTemp {
    void* p = nullptr;
    DeviceTemp() = default;
    ~DeviceTemp() {
        if (p) (void)hipFree(p);
    }
    DeviceTemp(const DeviceTemp&) = delete;
    DeviceTemp& operator=(const DeviceTemp&) = delete;
};

struct HostCond {
    int kind;
    double vel[3];
    double rho;
};

// Lattice tables as plain arrays for host code.
struct HostLattice {
    int Q = 0;
    int c[27][3];
    double w[27];
};

template <class L>
inline HostLattice make_lattice() {
    HostLattice h;
    h.Q = L::Q;
    for (int q = 0; q < L::Q; ++q) {
        h.c[q][0] = L::cx(q);
        h.c[q][1] = L::cy(q);
        h.c[q][2] = L::cz(q);
        h.w[q] = weight<L>(q);
    }
    return h;
}

inline HostLattice lattice_for(int q) {
    if (q == 15) return make_lattice<D3Q15>();
    if (q == 19) return make_lattice<D3Q19>();
    return make_lattice<D3Q27>();
}

inline size_t elem_size(Mode m) {
    switch (m) {
        case Mode::EXACT64:
        case Mode::F64S: return 8;
        case Mode::F32: return 4;
        default: return 2;
    }
}

inline int64_t roundup(int64_t v, int64_t a) { return (v + a - 1) / a * a; }

// L2 domains (XCDs) of a gfx950 device with ncu compute units: 32 CUs and
// one L2 each; the hardware deals a launch's blocks to them in turn.
inline int xcd_count(int ncu) { return (ncu >= 32 && ncu % 32 == 0) ? ncu / 32 : 1; }

}  // namespace lbm_amd

// internal header: the domain struct below is spelled with these names
using namespace lbm_amd;

struct lbm_domain {
    lbm_desc desc{};
    Mode mode = Mode::EXACT64;
    const ModeOps* ops = nullptr;
    HostLattice lat;
    int Q = 0;
    int nx = 0, ny = 0, nz = 0;
    int64_t ncell = 0;
    FieldLayout L{};
    bool compact = false;  // rows packed at pitch xl + 2 (no line alignment)
    bool aa = false;       // one buffer updated in place (AA pattern), buf[0] only
    int aa_phase = 0;      // AA: 0 natural representation, 1 swapped
    size_t esize = 0;
    size_t field_bytes = 0;
    void* buf[2] = {nullptr, nullptr};
    int cur = 0;
    uint8_t* d_kind = nullptr;
    uint32_t* d_lmask = nullptr;  // per element: non-fluid sources of a NEAR cell (bit q)
    uint8_t* d_lcond = nullptr;   // per element: their shared condition (same allocation, after d_lmask)
    CellCoord* d_blist = nullptr;
    int64_t nb = 0;
    Cond* d_conds = nullptr;
    int nconds_dev = 0;
    void* d_alpha = nullptr;
    double g[3] = {0, 0, 0};
    std::vector<uint16_t> handler;  // per ghosted cell: 0 fluid, else 1 + index in conds
    std::vector<HostCond> conds;
    int mask_cond = -1;
    std::vector<double> h_alpha;  // interior, x fastest (empty: none)
    bool finalized = false;
    bool materialised = true;
    hipStream_t stream = nullptr;
    StepPlan plan{};
    int tune_vec = 0, tune_bx = 0;
    int tune_pass = 0, tune_zc = 0;  // steps per pass (0: automatic), planes per block of the pass
    bool force64 = false;            // 64-bit element indices even where a plane is under 4 GiB
    double* d_partial = nullptr;
    double* d_out4 = nullptr;
    int nred = 0;
    uint64_t n_fluid = 0;
    size_t device_bytes = 0;

    int64_t hidx(int64_t x, int64_t y, int64_t z) const { return x + nx * (y + ny * z); }
    bool interior(int64_t x, int64_t y, int64_t z) const {
        return x >= 1 && x <= L.xl && y >= 1 && y <= L.yl && z >= 1 && z <= L.zl;
    }
    int64_t wrapc(int64_t v, int axis) const {
        const int64_t l = axis == 0 ? L.xl : (axis == 1 ? L.yl : L.zl);
        if (!L.per[axis]) return v;
        if (v < 1) return v + l;
        if (v > l) return v - l;
        return v;
    }
    StepArgs args(const void* src, void* dst, int mat) const {
        StepArgs a{};
        a.L = L;
        a.src = src;
        a.dst = dst;
        a.kind = d_kind;
        a.lmask = d_lmask;
        a.lcond = d_lcond;
        a.blist = d_blist;
        a.nb = nb;
        a.conds = d_conds;
        a.materialised = mat;
        a.tau = desc.tau;
        a.omega = 1.0 / desc.tau;
        a.alpha = d_alpha;
        a.g[0] = g[0];
        a.g[1] = g[1];
        a.g[2] = g[2];
        return a;
    }
};

namespace lbm_amd {

// Interior porosity (x fastest) -> the mode's arithmetic type in the padded
// element layout (fp64 for F64S, fp32 otherwise).
inline std::vector<char> alpha_in_layout(const lbm_domain* d, const double* alpha) {
    const FieldLayout& L = d->L;
    const size_t cs = (d->mode == Mode::F64S) ? 8 : 4;
    std::vector<char> a(static_cast<size_t>(L.plane) * cs, 0);
    int64_t i = 0;
    for (int z = 1; z <= L.zl; ++z)
        for (int y = 1; y <= L.yl; ++y)
            for (int x = 1; x <= L.xl; ++x, ++i) {
                const int64_t e = cell_elem(L, x, y, z);
                if (cs == 8)
                    reinterpret_cast<double*>(a.data())[e] = alpha[i];
                else
                    reinterpret_cast<float*>(a.data())[e] = static_cast<float>(alpha[i]);
            }
    return a;
}

// In-place layout: brings the field back to the natural representation if
// the last step left it swapped (defined in capi.hip).
int aa_settle(lbm_domain* d);
// Writes the boundary links' rule values into the current field, as the
// reference's boundary pass leaves them (defined in capi.hip).
int materialise_boundaries(lbm_domain* d);

inline int check_handle(lbm_domain* d) {
    if (!d) return fail(LBM_ERR_ARG, "null domain handle");
    hipError_t e = hipSetDevice(d->desc.device);
    if (e != hipSuccess) return fail(LBM_ERR_HIP, std::string("hipSetDevice: ") + hipGetErrorString(e));
    return LBM_OK;
}

}  // namespace lbm_amd
