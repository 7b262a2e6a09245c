/*
 * CPU oracle for the collide-stream time step -- TEST INFRASTRUCTURE ONLY.
 *
 * This is a plain-C restatement of the reference solver's time step
 * (src/main.cpp:48-54 -> include/domain.hpp:116-172), kept in the reference's
 * own structure: two ghosted array-of-structures pdf fields, a per-cell handler
 * id, pull streaming into the second field, a field swap, then a two-pass
 * collide (BGK on fluid cells, boundary rules on the others), all in fp64 and
 * in the reference's operation order. It exists to check the HIP path and to
 * time the reference algorithm on host cores; only the test suite, the smoke
 * check and the CPU-baseline leg of the benchmark may load it. The shipped
 * library never calls it.
 *
 * Extensions that the reference does not have (periodic axes, a Brinkman
 * porosity field with a body force) are marked as such where they appear.
 */
#ifndef LBM_ORACLE_H
#define LBM_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Handler kinds; numbering matches the shipped C-ABI (include/lbm_hip.h). */
enum {
    ORA_NOSLIP = 1,     /* include/boundary.hpp:14-31 */
    ORA_MOVINGWALL = 2, /* include/boundary.hpp:35-68 */
    ORA_FREESLIP = 3,   /* include/boundary.hpp:72-115 */
    ORA_OUTFLOW = 4,    /* include/boundary.hpp:119-150 */
    ORA_INFLOW = 5,     /* include/boundary.hpp:154-181 */
    ORA_PRESSURE = 6    /* include/boundary.hpp:185-214 */
};

typedef struct ora_domain ora_domain;

/* Domain<M>(xl,yl,zl,BGKCollision(tau)) -- include/domain.hpp:87-98 and
 * include/collision.hpp:55-58. periodic[3] (may be NULL) is an extension. */
ora_domain* ora_create(int q, int64_t xl, int64_t yl, int64_t zl, double tau,
                       const int* periodic);
void ora_destroy(ora_domain* d);

/* Tables as the oracle holds them (include/model.h). */
void ora_lattice(int q, int* c /* q*3 */, double* w /* q */);
double ora_cs(void);

/* setBoundaryCondition -- include/domain.hpp:175-194; params = vx,vy,vz,rho. */
int ora_set_boundary_box(ora_domain* d, int kind, const double* params, int64_t x0,
                         int64_t xe, int64_t y0, int64_t ye, int64_t z0, int64_t ze);
/* VTK mask reader's effect -- include/io/vtk.hpp:137-149: interior cells with
 * mask == 0 get one shared NoSlip handler. mask is xl*yl*zl, x fastest. */
int ora_set_solid_mask(ora_domain* d, const uint8_t* mask);
/* set_nonfluid_cells_nullcollide -- include/domain.hpp:101-113, plus the
 * check that no unassigned ghost cell is a pull source. */
int ora_finalize(ora_domain* d);

/* Extension: Brinkman porosity alpha per interior cell (xl*yl*zl, x fastest)
 * and a body acceleration g. alpha==NULL removes the field. */
int ora_set_brinkman(ora_domain* d, const double* alpha, const double* g);

/* nsteps x { stream(); swap(); collide(); } -- src/main.cpp:50-52. */
void ora_step(ora_domain* d, int64_t nsteps);
/* The three phases separately, for phase-level tests. */
void ora_stream(ora_domain* d);
void ora_swap(ora_domain* d);
void ora_collide(ora_domain* d);

/* collide_field as doubles, ghosted AoS order idx(x,y,z)*Q + q. */
void ora_get_pdf(const ora_domain* d, double* out);
void ora_set_pdf(ora_domain* d, const double* in);
/* density()/velocity() of interior cells, x fastest -- include/io/vtk.hpp:62-73. */
void ora_moments(const ora_domain* d, double* rho, double* u3);
/* Sums over interior fluid cells: mass, momentum x, y, z. */
void ora_conservation(const ora_domain* d, double* out4);
/* Handler kind per ghosted cell: 0 fluid, 255 null, else condition kind. */
void ora_get_kinds(const ora_domain* d, uint8_t* out);

int ora_num_threads(void);

#ifdef __cplusplus
}
#endif
#endif
