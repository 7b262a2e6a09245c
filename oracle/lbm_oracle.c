/*
 * CPU oracle for the collide-stream time step -- TEST INFRASTRUCTURE ONLY.
 * See lbm_oracle.h for what it is, who may use it and what it is not.
 *
 * Every function follows the reference file:line named in its comment, in
 * the reference's data layout (ghosted AoS, two fields), loop order and fp64
 * operation order, so that results are bit-identical to a run of the
 * reference built without FMA contraction (-ffp-contract=off is set in the
 * Makefile next to this file).
 */
#include "lbm_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define MAXQ 27

/* ---------------------------------------------------------------------------
 * Lattice tables -- include/model.h:13-139. Velocities are small integers
 * stored as doubles in the reference; weights are k.0/n literals.
 */
static const int C15[15][3] = {
    {-1, -1, -1}, {1, -1, -1}, {0, 0, -1}, {-1, 1, -1}, {1, 1, -1},
    {0, -1, 0},   {-1, 0, 0},  {0, 0, 0},  {1, 0, 0},   {0, 1, 0},
    {-1, -1, 1},  {1, -1, 1},  {0, 0, 1},  {-1, 1, 1},  {1, 1, 1}};
static const int W15[15] = {1, 1, 8, 1, 1, 8, 8, 16, 8, 8, 1, 1, 8, 1, 1};

static const int C19[19][3] = {
    {0, -1, -1}, {-1, 0, -1}, {0, 0, -1}, {1, 0, -1}, {0, 1, -1},
    {-1, -1, 0}, {0, -1, 0},  {1, -1, 0}, {-1, 0, 0}, {0, 0, 0},
    {1, 0, 0},   {-1, 1, 0},  {0, 1, 0},  {1, 1, 0},  {0, -1, 1},
    {-1, 0, 1},  {0, 0, 1},   {1, 0, 1},  {0, 1, 1}};
static const int W19[19] = {1, 1, 2, 1, 1, 1, 2, 1, 2, 12, 2, 1, 2, 1, 1, 1, 2, 1, 1};

static const int W27[27] = {1,  4, 1,  4, 16, 4,  1, 4,  1, 4,  16, 4, 16, 64,
                            16, 4, 16, 4, 1,  4,  1, 4, 16, 4,  1,  4, 1};

/* include/lbmdefinitions.h:47 -- a long double literal narrowed to double. */
static const double CS = 0.57735026919L;

typedef struct {
    int kind;
    double vel[3];
    double rho;
} ora_cond;

struct ora_domain {
    int Q;
    int64_t xl, yl, zl, nx, ny, nz, ncell;
    double tau;
    int periodic[3];
    double c[MAXQ][3]; /* velocities as doubles, like the reference */
    int ci[MAXQ][3];
    double w[MAXQ];
    double* field[2]; /* AoS pdfs: field[k][cell*Q + q] */
    int cur;          /* index of collide_field */
    /* Handler per ghosted cell: 0 = BGK fluid, 0xFFFF = NullCollision,
     * otherwise 1 + index into conds. The reference keeps one handler pointer
     * per cell per field; both are set together (domain.hpp:189-190) except
     * for NullCollision, which only matters for cells no fluid cell reads. */
    uint16_t* handler;
    ora_cond* conds;
    int nconds, capconds;
    int mask_cond; /* condition index of the shared mask NoSlip, or -1 */
    double* alpha; /* extension: Brinkman porosity per ghosted cell, or NULL */
    double g[3];
};

#define H_FLUID 0
#define H_NULL 0xFFFF

double ora_cs(void) { return CS; }

static int fill_tables(int q, int (*ci)[3], double* w) {
    if (q == 15) {
        for (int i = 0; i < 15; ++i) {
            for (int d = 0; d < 3; ++d) ci[i][d] = C15[i][d];
            w[i] = (double)W15[i] / 72.0;
        }
    } else if (q == 19) {
        for (int i = 0; i < 19; ++i) {
            for (int d = 0; d < 3; ++d) ci[i][d] = C19[i][d];
            w[i] = (double)W19[i] / 36.0;
        }
    } else if (q == 27) {
        int i = 0;
        for (int z = -1; z <= 1; ++z)
            for (int y = -1; y <= 1; ++y)
                for (int x = -1; x <= 1; ++x, ++i) {
                    ci[i][0] = x;
                    ci[i][1] = y;
                    ci[i][2] = z;
                    w[i] = (double)W27[i] / 216.0;
                }
    } else {
        return -1;
    }
    return 0;
}

void ora_lattice(int q, int* c, double* w) {
    int ci[MAXQ][3];
    double ww[MAXQ];
    if (fill_tables(q, ci, ww)) return;
    for (int i = 0; i < q; ++i) {
        for (int d = 0; d < 3; ++d) c[3 * i + d] = ci[i][d];
        w[i] = ww[i];
    }
}

int ora_num_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

static inline int64_t IDX(const ora_domain* d, int64_t x, int64_t y, int64_t z) {
    return x + d->nx * (y + d->ny * z); /* include/domain.hpp:61-64 */
}

/* include/domain.hpp:67-70: strictly interior. */
static inline int in_bounds(const ora_domain* d, int64_t x, int64_t y, int64_t z) {
    return x > 0 && x < d->xl + 1 && y > 0 && y < d->yl + 1 && z > 0 && z < d->zl + 1;
}

/* Extension: wrap a coordinate into [1, l] along a periodic axis. */
static inline int64_t wrap(int64_t v, int64_t l, int per) {
    if (!per) return v;
    if (v < 1) return v + l;
    if (v > l) return v - l;
    return v;
}

static inline int is_fluid_h(uint16_t h) { return h == H_FLUID; }

ora_domain* ora_create(int q, int64_t xl, int64_t yl, int64_t zl, double tau,
                       const int* periodic) {
    if (xl < 1 || yl < 1 || zl < 1) return NULL;
    ora_domain* d = (ora_domain*)calloc(1, sizeof(ora_domain));
    if (!d) return NULL;
    if (fill_tables(q, d->ci, d->w)) {
        free(d);
        return NULL;
    }
    d->Q = q;
    for (int i = 0; i < q; ++i)
        for (int k = 0; k < 3; ++k) d->c[i][k] = (double)d->ci[i][k];
    d->xl = xl;
    d->yl = yl;
    d->zl = zl;
    d->nx = xl + 2;
    d->ny = yl + 2;
    d->nz = zl + 2;
    d->ncell = d->nx * d->ny * d->nz;
    d->tau = tau;
    for (int k = 0; k < 3; ++k) d->periodic[k] = periodic ? (periodic[k] != 0) : 0;
    d->mask_cond = -1;
    size_t n = (size_t)d->ncell * (size_t)q;
    d->field[0] = (double*)malloc(n * sizeof(double));
    d->field[1] = (double*)malloc(n * sizeof(double));
    d->handler = (uint16_t*)calloc((size_t)d->ncell, sizeof(uint16_t));
    if (!d->field[0] || !d->field[1] || !d->handler) {
        ora_destroy(d);
        return NULL;
    }
    /* Cell ctor copies the weights into every cell of both fields
     * (include/cell.hpp:9-15, include/domain.hpp:87-98). */
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < d->ncell; ++i)
        for (int k = 0; k < q; ++k) {
            d->field[0][i * q + k] = d->w[k];
            d->field[1][i * q + k] = d->w[k];
        }
    d->cur = 0;
    return d;
}

void ora_destroy(ora_domain* d) {
    if (!d) return;
    free(d->field[0]);
    free(d->field[1]);
    free(d->handler);
    free(d->conds);
    free(d->alpha);
    free(d);
}

static int add_cond(ora_domain* d, int kind, const double* p) {
    if (d->nconds == d->capconds) {
        int nc = d->capconds ? 2 * d->capconds : 8;
        ora_cond* n = (ora_cond*)realloc(d->conds, (size_t)nc * sizeof(ora_cond));
        if (!n) return -1;
        d->conds = n;
        d->capconds = nc;
    }
    ora_cond* c = &d->conds[d->nconds];
    c->kind = kind;
    c->vel[0] = p ? p[0] : 0.0;
    c->vel[1] = p ? p[1] : 0.0;
    c->vel[2] = p ? p[2] : 0.0;
    c->rho = p ? p[3] : 1.0;
    return d->nconds++;
}

/* include/domain.hpp:175-194 (inclusive box, ghosted coordinates, later calls
 * overwrite earlier ones). Each call creates a new condition object, like
 * BoundaryKeeper::get_collision (include/boundary.h:85-91). */
int ora_set_boundary_box(ora_domain* d, int kind, const double* params, int64_t x0,
                         int64_t xe, int64_t y0, int64_t ye, int64_t z0, int64_t ze) {
    if (kind < ORA_NOSLIP || kind > ORA_PRESSURE) return -1;
    if (xe < x0 || ye < y0 || ze < z0) return -2;
    if (x0 < 0 || y0 < 0 || z0 < 0 || xe > d->xl + 1 || ye > d->yl + 1 || ze > d->zl + 1)
        return -2;
    if (d->nconds >= 0xFFFE) return -3;
    int ci = add_cond(d, kind, params);
    if (ci < 0) return -3;
    uint16_t h = (uint16_t)(ci + 1);
    for (int64_t z = z0; z <= ze; ++z)
        for (int64_t y = y0; y <= ye; ++y)
            for (int64_t x = x0; x <= xe; ++x) d->handler[IDX(d, x, y, z)] = h;
    return 0;
}

/* include/io/vtk.hpp:137-149: one shared NoSlip object for all mask==0 cells,
 * interior cells only, mask in x-fastest interior order. */
int ora_set_solid_mask(ora_domain* d, const uint8_t* mask) {
    if (d->mask_cond < 0) {
        d->mask_cond = add_cond(d, ORA_NOSLIP, NULL);
        if (d->mask_cond < 0) return -3;
    }
    uint16_t h = (uint16_t)(d->mask_cond + 1);
    int64_t i = 0;
    for (int64_t z = 1; z <= d->zl; ++z)
        for (int64_t y = 1; y <= d->yl; ++y)
            for (int64_t x = 1; x <= d->xl; ++x, ++i)
                if (!mask[i]) d->handler[IDX(d, x, y, z)] = h;
    return 0;
}

int ora_set_brinkman(ora_domain* d, const double* alpha, const double* g) {
    free(d->alpha);
    d->alpha = NULL;
    d->g[0] = g ? g[0] : 0.0;
    d->g[1] = g ? g[1] : 0.0;
    d->g[2] = g ? g[2] : 0.0;
    if (!alpha) return 0;
    d->alpha = (double*)calloc((size_t)d->ncell, sizeof(double));
    if (!d->alpha) return -3;
    int64_t i = 0;
    for (int64_t z = 1; z <= d->zl; ++z)
        for (int64_t y = 1; y <= d->yl; ++y)
            for (int64_t x = 1; x <= d->xl; ++x, ++i) d->alpha[IDX(d, x, y, z)] = alpha[i];
    return 0;
}

/* Source cell of the pull for interior cell (x,y,z) and direction q. */
static inline int64_t src_index(const ora_domain* d, int64_t x, int64_t y, int64_t z, int q) {
    int64_t sx = wrap(x - d->ci[q][0], d->xl, d->periodic[0]);
    int64_t sy = wrap(y - d->ci[q][1], d->yl, d->periodic[1]);
    int64_t sz = wrap(z - d->ci[q][2], d->zl, d->periodic[2]);
    return IDX(d, sx, sy, sz);
}

/* include/domain.hpp:101-113 and include/cell.hpp:75-92. The vicinity test
 * uses the same neighbour map as streaming (wrapped along periodic axes).
 * Also rejects a domain where an interior fluid cell would pull from a ghost
 * cell that no boundary covers: the reference would read a BGK-collided ghost
 * that is never streamed into, which no shipped scenario relies on. */
int ora_finalize(ora_domain* d) {
    int bad = 0;
    for (int64_t z = 1; z <= d->zl; ++z)
        for (int64_t y = 1; y <= d->yl; ++y)
            for (int64_t x = 1; x <= d->xl; ++x) {
                int64_t i = IDX(d, x, y, z);
                if (!is_fluid_h(d->handler[i])) continue;
                for (int q = 0; q < d->Q; ++q) {
                    int64_t sx = wrap(x - d->ci[q][0], d->xl, d->periodic[0]);
                    int64_t sy = wrap(y - d->ci[q][1], d->yl, d->periodic[1]);
                    int64_t sz = wrap(z - d->ci[q][2], d->zl, d->periodic[2]);
                    if (!in_bounds(d, sx, sy, sz) && is_fluid_h(d->handler[IDX(d, sx, sy, sz)]))
                        bad = 1;
                }
            }
    if (bad) return -4;
    for (int64_t z = 1; z <= d->zl; ++z)
        for (int64_t y = 1; y <= d->yl; ++y)
            for (int64_t x = 1; x <= d->xl; ++x) {
                int64_t i = IDX(d, x, y, z);
                if (is_fluid_h(d->handler[i])) continue;
                int vic = 0;
                for (int q = 0; q < d->Q && !vic; ++q) {
                    int64_t nx = wrap(x + d->ci[q][0], d->xl, d->periodic[0]);
                    int64_t ny = wrap(y + d->ci[q][1], d->yl, d->periodic[1]);
                    int64_t nz = wrap(z + d->ci[q][2], d->zl, d->periodic[2]);
                    if (in_bounds(d, nx, ny, nz) && is_fluid_h(d->handler[IDX(d, nx, ny, nz)]))
                        vic = 1;
                }
                if (!vic) d->handler[i] = H_NULL;
            }
    return 0;
}

/* ---------------------------------------------------------------------------
 * Collision::compute_density / compute_velocity / compute_feq --
 * include/collision.hpp:7-51. Sums start at 0 and run over ascending q (and
 * ascending d); velocities are doubles; divisions are true divisions.
 */
static inline double compute_density(const ora_domain* d, const double* f) {
    double density = 0;
    for (int q = 0; q < d->Q; ++q) density += f[q];
    return density;
}

static inline void compute_velocity(const ora_domain* d, const double* f, double density,
                                    double* v) {
    v[0] = 0.0;
    v[1] = 0.0;
    v[2] = 0.0;
    for (int k = 0; k < 3; ++k)
        for (int q = 0; q < d->Q; ++q) v[k] += f[q] * d->c[q][k];
    v[0] /= density;
    v[1] /= density;
    v[2] /= density;
}

static inline void compute_feq(const ora_domain* d, double density, const double* v,
                               double* feq) {
    for (int q = 0; q < d->Q; ++q) {
        double c_dot_u = 0.0;
        double u_dot_u = 0.0;
        for (int k = 0; k < 3; ++k) {
            c_dot_u += d->c[q][k] * v[k];
            u_dot_u += v[k] * v[k];
        }
        feq[q] = d->w[q] * density *
                 (1 + c_dot_u / (CS * CS) + (c_dot_u) * (c_dot_u) / (2 * CS * CS * CS * CS) -
                  u_dot_u / (2 * CS * CS));
    }
}

/* BGKCollision::collide -- include/collision.hpp:61-70. With a Brinkman field
 * (extension) the equilibrium velocity becomes (1 - alpha) * (u + tau * g). */
static inline void bgk_collide(const ora_domain* d, double* f, int64_t i) {
    double feq[MAXQ], v[3];
    const double density = compute_density(d, f);
    compute_velocity(d, f, density, v);
    if (d->alpha) {
        const double s = 1.0 - d->alpha[i];
        for (int k = 0; k < 3; ++k) v[k] = s * (v[k] + d->tau * d->g[k]);
    }
    compute_feq(d, density, v, feq);
    for (int q = 0; q < d->Q; ++q) f[q] -= (f[q] - feq[q]) / d->tau;
}

static int velocity_index(const ora_domain* d, int u, int v, int w) {
    for (int q = 0; q < d->Q; ++q)
        if (d->ci[q][0] == u && d->ci[q][1] == v && d->ci[q][2] == w) return q;
    return -1;
}

/* Boundary rules -- include/boundary.hpp. Neighbour coordinates are wrapped
 * along periodic axes (extension; identity otherwise). */
static void boundary_collide(const ora_domain* d, double* field, int64_t x, int64_t y,
                             int64_t z, const ora_cond* cond) {
    const int Q = d->Q;
    double* cell = field + IDX(d, x, y, z) * Q;
    for (int q = 0; q < Q; ++q) {
        const int dx = d->ci[q][0], dy = d->ci[q][1], dz = d->ci[q][2];
        const int64_t nx = wrap(x + dx, d->xl, d->periodic[0]);
        const int64_t ny = wrap(y + dy, d->yl, d->periodic[1]);
        const int64_t nz = wrap(z + dz, d->zl, d->periodic[2]);
        if (!(in_bounds(d, nx, ny, nz) && is_fluid_h(d->handler[IDX(d, nx, ny, nz)]))) continue;
        const double* nb = field + IDX(d, nx, ny, nz) * Q;
        const int qi = Q - 1 - q; /* inv(q), include/model.h */
        switch (cond->kind) {
            case ORA_NOSLIP: /* boundary.hpp:26-28 */
                cell[q] = nb[qi];
                break;
            case ORA_MOVINGWALL: { /* boundary.hpp:55-65 */
                const double density = compute_density(d, nb);
                double c_dot_u = 0;
                for (int k = 0; k < 3; ++k) c_dot_u += d->c[q][k] * cond->vel[k];
                const double finv = nb[qi];
                cell[q] = finv + 2.0 * d->w[q] * density * c_dot_u / (CS * CS);
                break;
            }
            case ORA_FREESLIP: { /* boundary.hpp:91-110 */
                const int64_t xs = wrap(x + dx, d->xl, d->periodic[0]);
                const int64_t ys = wrap(y + dy, d->yl, d->periodic[1]);
                const int64_t zs = wrap(z + dz, d->zl, d->periodic[2]);
                const int64_t cand[6][3] = {{xs, y, z},  {x, ys, z},  {x, y, zs},
                                            {x, ys, zs}, {xs, y, zs}, {xs, ys, z}};
                const int refl[6][3] = {{-dx, dy, dz},  {dx, -dy, dz},  {dx, dy, -dz},
                                        {dx, -dy, -dz}, {-dx, dy, -dz}, {-dx, -dy, dz}};
                for (int b = 0; b < 6; ++b) {
                    const int64_t j = IDX(d, cand[b][0], cand[b][1], cand[b][2]);
                    if (is_fluid_h(d->handler[j])) {
                        cell[q] = field[j * Q +
                                        velocity_index(d, refl[b][0], refl[b][1], refl[b][2])];
                        break;
                    }
                }
                break;
            }
            case ORA_OUTFLOW:    /* boundary.hpp:140-147 */
            case ORA_PRESSURE: { /* boundary.hpp:206-211 */
                double v[3], feq[MAXQ];
                compute_velocity(d, nb, cond->rho, v);
                compute_feq(d, cond->rho, v, feq);
                cell[q] = feq[q] + feq[qi] - nb[qi];
                break;
            }
            case ORA_INFLOW: { /* boundary.hpp:176-178 */
                double feq[MAXQ];
                compute_feq(d, cond->rho, cond->vel, feq);
                cell[q] = feq[q];
                break;
            }
            default:
                break;
        }
    }
}

/* Domain::stream -- include/domain.hpp:116-141 (pull into stream_field). */
void ora_stream(ora_domain* d) {
    const int Q = d->Q;
    const double* src = d->field[d->cur];
    double* dst = d->field[d->cur ^ 1];
#pragma omp parallel for collapse(2) schedule(static)
    for (int64_t z = 1; z <= d->zl; ++z)
        for (int64_t y = 1; y <= d->yl; ++y)
            for (int64_t x = 1; x <= d->xl; ++x) {
                const int64_t i = IDX(d, x, y, z);
                if (!is_fluid_h(d->handler[i])) continue;
                for (int q = 0; q < Q; ++q) dst[i * Q + q] = src[src_index(d, x, y, z, q) * Q + q];
            }
}

/* Domain::swap -- include/domain.hpp:169-172. */
void ora_swap(ora_domain* d) { d->cur ^= 1; }

/* Domain::collide -- include/domain.hpp:144-166: BGK over every fluid cell of
 * the ghosted box, then the boundary rules over every other cell. Ghost cells
 * that sit at a ghost coordinate of a periodic axis are skipped in the second
 * pass (extension: nothing reads them). */
void ora_collide(ora_domain* d) {
    const int Q = d->Q;
    double* field = d->field[d->cur];
#pragma omp parallel for collapse(2) schedule(static)
    for (int64_t z = 0; z < d->nz; ++z)
        for (int64_t y = 0; y < d->ny; ++y)
            for (int64_t x = 0; x < d->nx; ++x) {
                const int64_t i = IDX(d, x, y, z);
                if (is_fluid_h(d->handler[i])) bgk_collide(d, field + i * Q, i);
            }
#pragma omp parallel for collapse(2) schedule(static)
    for (int64_t z = 0; z < d->nz; ++z)
        for (int64_t y = 0; y < d->ny; ++y)
            for (int64_t x = 0; x < d->nx; ++x) {
                const uint16_t h = d->handler[IDX(d, x, y, z)];
                if (is_fluid_h(h) || h == H_NULL) continue;
                if ((d->periodic[0] && (x == 0 || x == d->nx - 1)) ||
                    (d->periodic[1] && (y == 0 || y == d->ny - 1)) ||
                    (d->periodic[2] && (z == 0 || z == d->nz - 1)))
                    continue;
                boundary_collide(d, field, x, y, z, &d->conds[h - 1]);
            }
}

void ora_step(ora_domain* d, int64_t nsteps) {
    for (int64_t t = 0; t < nsteps; ++t) {
        ora_stream(d);
        ora_swap(d);
        ora_collide(d);
    }
}

void ora_get_pdf(const ora_domain* d, double* out) {
    memcpy(out, d->field[d->cur], (size_t)d->ncell * (size_t)d->Q * sizeof(double));
}

/* Cell::operator[] writes go to collide_field (include/domain.hpp:74-83). */
void ora_set_pdf(ora_domain* d, const double* in) {
    memcpy(d->field[d->cur], in, (size_t)d->ncell * (size_t)d->Q * sizeof(double));
}

/* include/io/vtk.hpp:62-73: density() then velocity(density) per interior
 * cell of collide_field, x fastest. */
void ora_moments(const ora_domain* d, double* rho, double* u3) {
    const double* field = d->field[d->cur];
#pragma omp parallel for collapse(2) schedule(static)
    for (int64_t z = 1; z <= d->zl; ++z)
        for (int64_t y = 1; y <= d->yl; ++y)
            for (int64_t x = 1; x <= d->xl; ++x) {
                const int64_t o = (x - 1) + d->xl * ((y - 1) + d->yl * (z - 1));
                const double* f = field + IDX(d, x, y, z) * d->Q;
                double v[3];
                const double density = compute_density(d, f);
                compute_velocity(d, f, density, v);
                if (rho) rho[o] = density;
                if (u3) {
                    u3[3 * o + 0] = v[0];
                    u3[3 * o + 1] = v[1];
                    u3[3 * o + 2] = v[2];
                }
            }
}

void ora_init_equilibrium(ora_domain* d, const double* rho, const double* u3) {
    double* field = d->field[d->cur];
#pragma omp parallel for collapse(2) schedule(static)
    for (int64_t z = 1; z <= d->zl; ++z)
        for (int64_t y = 1; y <= d->yl; ++y)
            for (int64_t x = 1; x <= d->xl; ++x) {
                const int64_t i = IDX(d, x, y, z);
                if (!is_fluid_h(d->handler[i])) continue;
                const int64_t o = (x - 1) + d->xl * ((y - 1) + d->yl * (z - 1));
                compute_feq(d, rho[o], u3 + 3 * o, field + i * d->Q);
            }
}

/* Mass and momentum summed over interior fluid cells in x-fastest order. */
void ora_conservation(const ora_domain* d, double* out4) {
    const double* field = d->field[d->cur];
    double s[4] = {0, 0, 0, 0};
    for (int64_t z = 1; z <= d->zl; ++z)
        for (int64_t y = 1; y <= d->yl; ++y)
            for (int64_t x = 1; x <= d->xl; ++x) {
                const int64_t i = IDX(d, x, y, z);
                if (!is_fluid_h(d->handler[i])) continue;
                const double* f = field + i * d->Q;
                for (int q = 0; q < d->Q; ++q) {
                    s[0] += f[q];
                    s[1] += f[q] * d->c[q][0];
                    s[2] += f[q] * d->c[q][1];
                    s[3] += f[q] * d->c[q][2];
                }
            }
    for (int k = 0; k < 4; ++k) out4[k] = s[k];
}

void ora_get_kinds(const ora_domain* d, uint8_t* out) {
    for (int64_t i = 0; i < d->ncell; ++i) {
        const uint16_t h = d->handler[i];
        out[i] = h == H_FLUID ? 0 : (h == H_NULL ? 255 : (uint8_t)d->conds[h - 1].kind);
    }
}
