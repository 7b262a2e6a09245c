// lbm_hip/lbm.hpp -- header-only C++ face of the C ABI (lbm_hip.h) with the
// reference's class names and call shapes, so that the reference's driver
// loop (src/main.cpp:30-73) compiles against it unchanged:
//
//   auto collision = lbm::BGKCollision<model>(tau);              // collision.hpp:55-58
//   auto domain = lbm::io::parse_scenario_file<model>(xml, cfg, collision);
//   domain->set_nonfluid_cells_nullcollide();                    // domain.hpp:101-113
//   for (t = 1..T) { domain->stream(); domain->swap(); domain->collide(); }
//
// On the GPU one time step is one fused kernel pass, so stream() enqueues
// the step, swap() flips nothing the caller can see and collide() waits for
// the step to finish, so that a clock read after collide() (src/main.cpp:53)
// times the step. step(n) enqueues n steps without waiting (sync() waits).
// Errors are thrown as std::runtime_error, which the reference driver's catch
// block reports. The headers under include/lbm_hip/compat carry the
// reference's header names, so src/main.cpp compiles against this file
// unchanged.
#pragma once

#include <array>
#include <cstddef>
#include <cstdint>
#include <list>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "lbm_hip.h"
#include "lbm_hip/lattice.hpp"

namespace lbm {

// include/lbmdefinitions.h:11-47
template <size_t SIZE>
using double_array = std::array<double, SIZE>;
template <size_t SIZE>
using int_array = std::array<int, SIZE>;
template <size_t SIZE>
using uint_array = std::array<std::uint64_t, SIZE>;
static constexpr double C_S = lbm_amd::kCS;

inline void check(int rc) {
    if (rc != LBM_OK) throw std::runtime_error(lbm_last_error());
}

namespace model {
// include/model.h:13-139: same names, sizes, tables and index helpers.
template <class L>
struct lattice_model {
    static constexpr size_t D = 3;
    static constexpr size_t Q = L::Q;
    static constexpr const char* name = L::name;
    static double velocity(size_t q, size_t d) {
        return d == 0 ? L::cx(static_cast<int>(q)) : (d == 1 ? L::cy(static_cast<int>(q)) : L::cz(static_cast<int>(q)));
    }
    static double weight(size_t q) { return lbm_amd::weight<L>(static_cast<int>(q)); }
    static int inv(int q) { return static_cast<int>(Q) - 1 - q; }
    static size_t velocity_index(int u, int v, int w) {
        for (int q = 0; q < static_cast<int>(Q); ++q)
            if (L::cx(q) == u && L::cy(q) == v && L::cz(q) == w) return static_cast<size_t>(q);
        throw std::out_of_range("velocity not in the lattice");
    }
};
struct d3q15 : lattice_model<lbm_amd::D3Q15> {};
struct d3q19 : lattice_model<lbm_amd::D3Q19> {};
struct d3q27 : lattice_model<lbm_amd::D3Q27> {};
}  // namespace model

// Storage / execution options of a GPU domain (no reference counterpart).
struct DeviceOptions {
    int storage = LBM_STORAGE_F64;  // the reference's fp64 arithmetic by default
    int device = 0;
    std::array<int, 3> periodic{{0, 0, 0}};
    int layout = LBM_LAYOUT_AB;  // LBM_LAYOUT_AA: one buffer updated in place
    // lbm_set_time_blocking's steps_per_pass: 0 (automatic) runs pairs of
    // steps as one pass over the field where that measured faster, 2 wherever
    // the domain allows it, 1 never; bit-identical to single steps
    int steps_per_pass = 0;
};

template <class M>
class Domain;

// include/collision.h:33-41,63-71
template <class M>
class FluidCollision {
public:
    virtual ~FluidCollision() = default;
    virtual double tau() const = 0;
    bool is_fluid() const { return true; }
};

template <class M>
class BGKCollision : public FluidCollision<M> {
    double tau_;

public:
    explicit BGKCollision(double tau) : tau_(tau) {}
    double tau() const override { return tau_; }
};

// include/collision.h:47-59 and the six boundary classes of
// include/boundary.h:9-70. On the device a boundary is a kind id plus
// parameters; these objects only carry them to setBoundaryCondition.
template <class M>
class NonFluidCollision {
protected:
    Domain<M>& domain;

public:
    const int kind;
    const std::array<double, 4> params;  // vx, vy, vz, rho
    NonFluidCollision(Domain<M>& d, int k, std::array<double, 4> p) : domain(d), kind(k), params(p) {}
    virtual ~NonFluidCollision() = default;
    bool is_fluid() const { return false; }
};

template <class M>
struct NoSlipBoundary : NonFluidCollision<M> {
    explicit NoSlipBoundary(Domain<M>& d) : NonFluidCollision<M>(d, LBM_BC_NOSLIP, {0, 0, 0, 1}) {}
};
template <class M>
struct MovingWallBoundary : NonFluidCollision<M> {
    MovingWallBoundary(Domain<M>& d, const double_array<3>& u)
        : NonFluidCollision<M>(d, LBM_BC_MOVINGWALL, {u[0], u[1], u[2], 1}) {}
};
template <class M>
struct FreeSlipBoundary : NonFluidCollision<M> {
    explicit FreeSlipBoundary(Domain<M>& d) : NonFluidCollision<M>(d, LBM_BC_FREESLIP, {0, 0, 0, 1}) {}
};
template <class M>
struct OutflowBoundary : NonFluidCollision<M> {
    explicit OutflowBoundary(Domain<M>& d, double rho_ref = 1.0)
        : NonFluidCollision<M>(d, LBM_BC_OUTFLOW, {0, 0, 0, rho_ref}) {}
};
template <class M>
struct InflowBoundary : NonFluidCollision<M> {
    InflowBoundary(Domain<M>& d, const double_array<3>& u, double rho_ref = 1.0)
        : NonFluidCollision<M>(d, LBM_BC_INFLOW, {u[0], u[1], u[2], rho_ref}) {}
};
template <class M>
struct PressureBoundary : NonFluidCollision<M> {
    PressureBoundary(Domain<M>& d, double rho_in)
        : NonFluidCollision<M>(d, LBM_BC_PRESSURE, {0, 0, 0, rho_in}) {}
};

// include/boundary.h:72-91: owner of condition objects.
template <class M>
class BoundaryKeeper {
    static std::list<std::unique_ptr<NonFluidCollision<M>>>& store() {
        static std::list<std::unique_ptr<NonFluidCollision<M>>> s;
        return s;
    }

public:
    template <class C, class... Args>
    static NonFluidCollision<M>& get_collision(Args&&... args) {
        store().push_back(std::make_unique<C>(std::forward<Args>(args)...));
        return *store().back();
    }
};

// include/cell.h:11-69 as a value: the Q pdfs of one cell with the
// reference's moment and equilibrium evaluation order.
template <class M>
class Cell {
    double_array<M::Q> pdf_{};
    bool fluid_ = true;

public:
    Cell() = default;
    Cell(const double* p, bool fluid) : fluid_(fluid) {
        for (size_t q = 0; q < M::Q; ++q) pdf_[q] = p[q];
    }
    bool is_fluid() const { return fluid_; }
    double& operator[](size_t i) { return pdf_[i]; }
    const double& operator[](size_t i) const { return pdf_[i]; }
    double density() const {  // collision.hpp:7-14
        double density = 0;
        for (size_t q = 0; q < M::Q; ++q) density += pdf_[q];
        return density;
    }
    double_array<3> velocity(double density) const {  // collision.hpp:17-31
        double_array<3> v = {0.0, 0.0, 0.0};
        for (size_t d = 0; d < 3; ++d)
            for (size_t q = 0; q < M::Q; ++q) v[d] += pdf_[q] * M::velocity(q, d);
        v[0] /= density;
        v[1] /= density;
        v[2] /= density;
        return v;
    }
    double_array<M::Q> equilibrium(double density, const double_array<3>& u) const {  // :34-51
        double_array<M::Q> feq;
        for (size_t q = 0; q < M::Q; ++q) {
            double cu = 0.0, uu = 0.0;
            for (size_t d = 0; d < 3; ++d) {
                cu += M::velocity(q, d) * u[d];
                uu += u[d] * u[d];
            }
            feq[q] = M::weight(q) * density *
                     (1 + cu / (C_S * C_S) + (cu) * (cu) / (2 * C_S * C_S * C_S * C_S) -
                      uu / (2 * C_S * C_S));
        }
        return feq;
    }
};

// include/domain.h:10-56 over one device domain.
template <class M>
class Domain {
    lbm_domain* h_ = nullptr;
    size_t xl_, yl_, zl_;
    double xo_, yo_, zo_, xs_, ys_, zs_;
    bool finalized_ = false;
    // host snapshot of the collide field for cell() reads
    mutable std::vector<double> snap_;
    mutable std::vector<uint8_t> fluid_;
    mutable bool snap_valid_ = false;
    uint64_t steps_ = 0;

    void ensure_final() {
        if (!finalized_) set_nonfluid_cells_nullcollide();
    }
    void load_snapshot() const {
        if (snap_valid_) return;
        snap_.resize((xl_ + 2) * (yl_ + 2) * (zl_ + 2) * M::Q);
        check(lbm_get_pdf(h_, snap_.data()));
        snap_valid_ = true;
    }

public:
    Domain(size_t xl, size_t yl, size_t zl, FluidCollision<M>& collision, double xorigin = 0,
           double yorigin = 0, double zorigin = 0, double xspacing = 1, double yspacing = 1,
           double zspacing = 1, const DeviceOptions& opt = DeviceOptions())
        : xl_(xl), yl_(yl), zl_(zl), xo_(xorigin), yo_(yorigin), zo_(zorigin), xs_(xspacing),
          ys_(yspacing), zs_(zspacing) {
        lbm_desc d;
        check(lbm_desc_init(&d));
        d.q = static_cast<int>(M::Q);
        d.storage = opt.storage;
        d.device = opt.device;
        for (int k = 0; k < 3; ++k) d.periodic[k] = opt.periodic[k];
        d.layout = opt.layout;
        d.xl = xl;
        d.yl = yl;
        d.zl = zl;
        d.tau = collision.tau();
        d.origin[0] = xorigin;
        d.origin[1] = yorigin;
        d.origin[2] = zorigin;
        d.spacing[0] = xspacing;
        d.spacing[1] = yspacing;
        d.spacing[2] = zspacing;
        check(lbm_domain_create(&d, &h_));
        check(lbm_set_time_blocking(h_, opt.steps_per_pass, 0));
        fluid_.assign((xl + 2) * (yl + 2) * (zl + 2), 1);
    }
    ~Domain() { lbm_domain_destroy(h_); }
    Domain(const Domain&) = delete;
    Domain& operator=(const Domain&) = delete;

    size_t xlength() const { return xl_; }
    size_t ylength() const { return yl_; }
    size_t zlength() const { return zl_; }
    double xorigin() const { return xo_; }
    double yorigin() const { return yo_; }
    double zorigin() const { return zo_; }
    double xspacing() const { return xs_; }
    double yspacing() const { return ys_; }
    double zspacing() const { return zs_; }
    lbm_domain* handle() const { return h_; }
    uint64_t steps() const { return steps_; }

    // include/domain.hpp:61-70
    int idx(int x, int y, int z) const {
        return x + static_cast<int>(xl_ + 2) * y + static_cast<int>((xl_ + 2) * (yl_ + 2)) * z;
    }
    bool in_bounds(int x, int y, int z) const {
        return x > 0 && x < static_cast<int>(xl_) + 1 && y > 0 && y < static_cast<int>(yl_) + 1 &&
               z > 0 && z < static_cast<int>(zl_) + 1;
    }

    // include/domain.hpp:175-194
    void setBoundaryCondition(NonFluidCollision<M>& c, size_t x0, size_t xE, size_t y0, size_t yE,
                              size_t z0, size_t zE) {
        check(lbm_set_boundary_box(h_, c.kind, c.params.data(), x0, xE, y0, yE, z0, zE));
        for (size_t z = z0; z <= zE; ++z)
            for (size_t y = y0; y <= yE; ++y)
                for (size_t x = x0; x <= xE; ++x) fluid_[idx(x, y, z)] = 0;
        snap_valid_ = false;
    }
    // the mask effect of include/io/vtk.hpp:137-149
    void set_solid_mask(const std::vector<uint8_t>& fluid_mask) {
        check(lbm_set_solid_mask(h_, fluid_mask.data()));
        size_t i = 0;
        for (size_t z = 1; z <= zl_; ++z)
            for (size_t y = 1; y <= yl_; ++y)
                for (size_t x = 1; x <= xl_; ++x, ++i)
                    if (!fluid_mask[i]) fluid_[idx(x, y, z)] = 0;
    }
    // include/domain.hpp:101-113
    void set_nonfluid_cells_nullcollide() {
        if (finalized_) return;
        check(lbm_finalize(h_));
        finalized_ = true;
    }

    // src/main.cpp:50-52
    void stream() {
        ensure_final();
        check(lbm_step(h_, 1));
        ++steps_;
        snap_valid_ = false;
    }
    void swap() {}
    void collide() { check(lbm_sync(h_)); }
    void step(uint64_t n) {
        ensure_final();
        check(lbm_step(h_, n));
        steps_ += n;
        snap_valid_ = false;
    }
    void sync() { check(lbm_sync(h_)); }

    // include/domain.hpp:74-83 (read access to the collide field)
    Cell<M> cell(int x, int y, int z) const {
        load_snapshot();
        return Cell<M>(snap_.data() + static_cast<size_t>(idx(x, y, z)) * M::Q, fluid_[idx(x, y, z)] != 0);
    }
    // write access: whole-field upload (cell(x,y,z)[q] = v in the reference)
    void set_cell(int x, int y, int z, const double_array<M::Q>& pdf) {
        load_snapshot();
        for (size_t q = 0; q < M::Q; ++q) snap_[static_cast<size_t>(idx(x, y, z)) * M::Q + q] = pdf[q];
        check(lbm_set_pdf(h_, snap_.data()));
    }
    // densities and velocities of interior cells, x fastest (include/io/vtk.hpp:62-73)
    void moments(std::vector<double>& rho, std::vector<double>& u) const {
        rho.resize(xl_ * yl_ * zl_);
        u.resize(3 * xl_ * yl_ * zl_);
        check(lbm_moments(h_, rho.data(), u.data()));
    }
    double_array<4> conservation() const {
        double_array<4> out{};
        check(lbm_conservation(h_, out.data()));
        return out;
    }
};

template <class M>
using Domain_ptr = std::unique_ptr<Domain<M>>;

}  // namespace lbm
