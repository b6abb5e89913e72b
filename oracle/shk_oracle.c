/*
 * shk_oracle.c -- CPU FP64 restatement of the Shiokaze MAC liquid pressure
 * projection.  TEST INFRASTRUCTURE ONLY: the parity checker for the HIP product
 * path and the CPU baseline timed by bench.py.  The product never links it.
 *
 * Built with -ffp-contract=off so that every expression is rounded exactly as
 * written (the reference is compiled for x86-64 without FMA contraction).
 *
 * Parity pins (each function is checked by tests/ against something other than
 * a retyping of its own formula):
 *   fraction        closed forms, symmetry, special cases          (test_oracle_fractions)
 *   get_area        exact polygon areas, half-plane cuts, brute force (test_oracle_fractions)
 *   area/fluid fraction  wall rule, eps clamp, clamped indexing      (test_oracle_fractions)
 *   label           brute-force relation on tiny grids, symmetry     (test_oracle_system)
 *   assemble        symmetry, M-matrix, row sums, constant rows,
 *                   rhs = -div for linear fields, accuracytest3 closed
 *                   form with second-order convergence              (test_oracle_system,
 *                                                                    test_oracle_accuracy)
 *   pcg             scipy direct-solve fixtures, CG identities,
 *                   all modes agree on the iterates where they must  (test_oracle_solver)
 *   mic0            L L^T == A on the pattern up to the modification (test_oracle_solver)
 *   update/divergence  post-divergence == -(b - A p) identity        (test_oracle_system)
 *   project() branches (test_oracle_branches):
 *     curvature        exact on quadratics; 2/r on a sphere's distance field
 *     surface tension  Laplace's law: a static drop projects to p = 2*sigma/R, u ~ 0
 *     volume correction  projected divergence equals the rhs constant on every
 *                      unknown; sign, zero and integral behaviour of the controller
 *                      (its gains 2.3/(25*0.01) and kp^2/16 are the reference's
 *                      constants, pinned only through those properties)
 *     solid density    density 1 leaves areas unchanged and scales the velocity by
 *                      the area; density 0 opens every face; affine in between
 *     first order      nonzero fractions -> 1, zeros kept
 *     warm start       zero previous pressure == cold start bit for bit; any previous
 *                      pressure gives the cold solution to the solver tolerance
 *   2D (test_oracle_2d; the 2D projection runs the functions above on the grid held
 *   as nx x 1 x ny, whose y faces are walls):
 *     area_fraction2d  walls, open faces, exact cuts of a straight solid boundary
 *     curvature2d      exact on quadratics; 1/r on a circle's distance field
 *     the embedding    its system equals a direct 4-neighbour assembly of the 2D
 *                      rows; accuracytest2's closed form converges at second order;
 *                      Laplace's law sigma/R for a static disc
 */
#include "shk_oracle.h"

#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#ifdef _OPENMP
#include <omp.h>
#endif

#define CELL(i, j, k) ((int64_t)(i) + (int64_t)nx * ((int64_t)(j) + (int64_t)ny * (int64_t)(k)))
#define NODE(i, j, k) ((int64_t)(i) + (int64_t)(nx + 1) * ((int64_t)(j) + (int64_t)(ny + 1) * (int64_t)(k)))

static inline int64_t face_index(int nx, int ny, int dim, int i, int j, int k) {
    const int64_t fx = nx + (dim == 0), fy = ny + (dim == 1);
    return (int64_t)i + fx * ((int64_t)j + fy * (int64_t)k);
}

static inline double store_real(double v, int real_is_double) {
    return real_is_double ? v : (double)(float)v;
}

static double now_ms(void) {
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return ts.tv_sec * 1e3 + ts.tv_nsec * 1e-6;
}

/* ---------------------------------------------------------------- a1, a2 */

double shk_or_fraction(double phi0, double phi1) {
    /* utility.h:174-182 */
    if (phi0 * phi1 >= 0.0) {
        if (phi0 < 0.0 || phi1 < 0.0) return 1.0;
        return 0.0;
    }
    const double lo = (phi1 < phi0) ? phi1 : phi0;
    const double gap = fabs(phi1 - phi0);
    return -lo / (gap < DBL_MIN ? DBL_MIN : gap);
}

double shk_or_get_area(double q00, double q10, double q11, double q01) {
    /* utility.h:191-226: walk the unit square's corners (0,0),(1,0),(1,1),(0,1);
     * emit each negative corner and each linear sign crossing on the edge to the
     * next corner, then take the shoelace area of the emitted polygon. */
    static const int corner[4][2] = {{0, 0}, {1, 0}, {1, 1}, {0, 1}};
    const double v[4] = {q00, q10, q11, q01};
    double px[8], py[8];
    int np = 0;
    for (int n = 0; n < 4; ++n) {
        const int m = (n + 1) & 3;
        if (v[n] < 0.0) {
            px[np] = corner[n][0];
            py[np] = corner[n][1];
            ++np;
        }
        if (v[n] * v[m] < 0) {
            const double y0 = v[n], y1 = v[m];
            if (y0 - y1) {
                const double a = y0 / (y0 - y1);
                px[np] = (1.0 - a) * corner[n][0] + a * corner[m][0];
                py[np] = (1.0 - a) * corner[n][1] + a * corner[m][1];
                ++np;
            }
        }
    }
    double sum = 0.0;
    for (int m = 0; m < np; ++m) {
        const int m1 = (m + 1) % np;
        sum += px[m] * py[m1] - py[m] * px[m1];
    }
    return 0.5 * sum;
}

/* ---------------------------------------------------------------- a3, a4 */

void shk_or_area_fraction(int nx, int ny, int nz, const double *solid, double eps_solid,
                          int real_is_double, double *ax, double *ay, double *az) {
    /* macutility3.cpp:98-171.  Dense evaluation equals the reference's band +
     * flood fill: away from the band all four corners share a sign, so the face
     * is fully open (1) or fully closed (0) there, which is what the fill gives. */
    double *out[3] = {ax, ay, az};
    const int n[3] = {nx, ny, nz};
    for (int dim = 0; dim < 3; ++dim) {
        const int fx = nx + (dim == 0), fy = ny + (dim == 1), fz = nz + (dim == 2);
#pragma omp parallel for schedule(static)
        for (int k = 0; k < fz; ++k) {
            for (int j = 0; j < fy; ++j) {
                for (int i = 0; i < fx; ++i) {
                    const int pi[3] = {i, j, k};
                    double area;
                    if (pi[dim] == 0 || pi[dim] == n[dim]) {
                        area = 0.0;
                    } else {
                        double q00, q10, q11, q01;
                        if (dim == 0) {
                            q00 = solid[NODE(i, j, k)];
                            q10 = solid[NODE(i, j + 1, k)];
                            q11 = solid[NODE(i, j + 1, k + 1)];
                            q01 = solid[NODE(i, j, k + 1)];
                        } else if (dim == 1) {
                            q00 = solid[NODE(i, j, k)];
                            q10 = solid[NODE(i + 1, j, k)];
                            q11 = solid[NODE(i + 1, j, k + 1)];
                            q01 = solid[NODE(i, j, k + 1)];
                        } else {
                            q00 = solid[NODE(i, j, k)];
                            q10 = solid[NODE(i + 1, j, k)];
                            q11 = solid[NODE(i + 1, j + 1, k)];
                            q01 = solid[NODE(i, j + 1, k)];
                        }
                        area = 1.0 - shk_or_get_area(q00, q10, q11, q01);
                    }
                    if (area != 0.0 && area < eps_solid) area = eps_solid;
                    out[dim][face_index(nx, ny, dim, i, j, k)] = store_real(area, real_is_double);
                }
            }
        }
    }
}

static inline int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

void shk_or_fluid_fraction(int nx, int ny, int nz, const double *fluid, double eps_fluid,
                           int real_is_double, double *rx, double *ry, double *rz) {
    /* macutility3.cpp:172-200 with shape3::clamp (shape.h:808-816). */
    double *out[3] = {rx, ry, rz};
    for (int dim = 0; dim < 3; ++dim) {
        const int fx = nx + (dim == 0), fy = ny + (dim == 1), fz = nz + (dim == 2);
#pragma omp parallel for schedule(static)
        for (int k = 0; k < fz; ++k) {
            for (int j = 0; j < fy; ++j) {
                for (int i = 0; i < fx; ++i) {
                    const int i0 = clampi(i, 0, nx - 1), j0 = clampi(j, 0, ny - 1), k0 = clampi(k, 0, nz - 1);
                    const int i1 = clampi(i - (dim == 0), 0, nx - 1), j1 = clampi(j - (dim == 1), 0, ny - 1),
                              k1 = clampi(k - (dim == 2), 0, nz - 1);
                    double rho = shk_or_fraction(fluid[CELL(i0, j0, k0)], fluid[CELL(i1, j1, k1)]);
                    if (rho != 0.0 && rho < eps_fluid) rho = eps_fluid;
                    out[dim][face_index(nx, ny, dim, i, j, k)] = store_real(rho, real_is_double);
                }
            }
        }
    }
}

/* ---------------------------------------------------------------- project() branches */

void shk_or_solid_density_blend(int nx, int ny, int nz, const double *du, const double *dv, const double *dw,
                                int real_is_double, double *ax, double *ay, double *az) {
    /* macpressuresolver3.cpp:75-80: area' = density*area + (1 - density).  With
     * Real = float the product is a float product; the sum is double, stored as float. */
    const double *dens[3] = {du, dv, dw};
    double *areas[3] = {ax, ay, az};
    for (int dim = 0; dim < 3; ++dim) {
        const int64_t nf = (int64_t)(nx + (dim == 0)) * (ny + (dim == 1)) * (nz + (dim == 2));
        for (int64_t f = 0; f < nf; ++f) {
            const double d = dens[dim][f], a = areas[dim][f];
            const double prod = real_is_double ? d * a : (double)((float)d * (float)a);
            areas[dim][f] = store_real(prod + (1.0 - d), real_is_double);
        }
    }
}

void shk_or_first_order(int nx, int ny, int nz, int second_order_solid, int second_order_fluid,
                        double *ax, double *ay, double *az, double *rx, double *ry, double *rz) {
    /* macpressuresolver3.cpp:83-93: without second-order accuracy every nonzero
     * fraction becomes 1. */
    double *areas[3] = {ax, ay, az}, *rhos[3] = {rx, ry, rz};
    for (int dim = 0; dim < 3; ++dim) {
        const int64_t nf = (int64_t)(nx + (dim == 0)) * (ny + (dim == 1)) * (nz + (dim == 2));
        for (int64_t f = 0; f < nf; ++f) {
            if (!second_order_solid && areas[dim][f] != 0.0) areas[dim][f] = 1.0;
            if (!second_order_fluid && rhos[dim][f] != 0.0) rhos[dim][f] = 1.0;
        }
    }
}

void shk_or_curvature(int nx, int ny, int nz, double dx, const double *fluid, int real_is_double, double *curv) {
    /* macpressuresolver3.cpp:100-108: 7-point Laplacian of the liquid level set with
     * clamped neighbours.  The six neighbours are summed left to right in Real
     * (float), the centre term and the division are double, stored as Real. */
    for (int k = 0; k < nz; ++k)
        for (int j = 0; j < ny; ++j)
            for (int i = 0; i < nx; ++i) {
                const int64_t nb[6] = {CELL(clampi(i - 1, 0, nx - 1), j, k), CELL(clampi(i + 1, 0, nx - 1), j, k),
                                       CELL(i, clampi(j - 1, 0, ny - 1), k), CELL(i, clampi(j + 1, 0, ny - 1), k),
                                       CELL(i, j, clampi(k - 1, 0, nz - 1)), CELL(i, j, clampi(k + 1, 0, nz - 1))};
                double sum;
                if (real_is_double) {
                    sum = fluid[nb[0]];
                    for (int q = 1; q < 6; ++q) sum = sum + fluid[nb[q]];
                } else {
                    float fs = (float)fluid[nb[0]];
                    for (int q = 1; q < 6; ++q) fs = fs + (float)fluid[nb[q]];
                    sum = fs;
                }
                const double value = (sum - 6.0 * fluid[CELL(i, j, k)]) / (dx * dx);
                curv[CELL(i, j, k)] = store_real(value, real_is_double);
            }
}

void shk_or_surface_tension(int nx, int ny, int nz, double dt, double dx, double kappa, const double *fluid,
                            const double *curv, const double *rx, const double *ry, const double *rz,
                            int real_is_double, double *u, double *v, double *w) {
    /* macpressuresolver3.cpp:110-122: on faces with 0 < rho < 1, the curvature
     * interpolated with theta (the liquid side's fraction) gives the jump
     * -sgn*dt/(dx*rho)*kappa*face_c, added to the face velocity (rounded to Real
     * first by array3::increment's `const T&` parameter). */
    const double *rhos[3] = {rx, ry, rz};
    double *vel[3] = {u, v, w};
    for (int dim = 0; dim < 3; ++dim) {
        const int fx = nx + (dim == 0), fy = ny + (dim == 1), fz = nz + (dim == 2);
        for (int k = 0; k < fz; ++k)
            for (int j = 0; j < fy; ++j)
                for (int i = 0; i < fx; ++i) {
                    const int64_t f = face_index(nx, ny, dim, i, j, k);
                    const double rho = rhos[dim][f];
                    if (!(rho != 0.0 && rho < 1.0)) continue;
                    const int64_t c1 = CELL(clampi(i, 0, nx - 1), clampi(j, 0, ny - 1), clampi(k, 0, nz - 1));
                    const int64_t c0 = CELL(clampi(i - (dim == 0), 0, nx - 1), clampi(j - (dim == 1), 0, ny - 1),
                                            clampi(k - (dim == 2), 0, nz - 1));
                    const double sgn = fluid[c1] < 0.0 ? -1.0 : 1.0;
                    const double theta = sgn < 0 ? 1.0 - rho : rho;
                    const double face_c = theta * curv[c1] + (1.0 - theta) * curv[c0];
                    const double inc = -sgn * dt / (dx * rho) * kappa * face_c;
                    if (real_is_double)
                        vel[dim][f] += inc;
                    else
                        vel[dim][f] = (double)((float)vel[dim][f] + (float)inc);
                }
    }
}

double shk_or_volume_correction(double gain, double current_volume, double target_volume, double dt,
                                double *y_prev) {
    /* macpressuresolver3.cpp:222-230: PI control of the relative volume error x with
     * integral y (state kept between projections); returns the constant added to
     * every rhs row, 0 when off. */
    if (!(gain != 0.0 && target_volume != 0.0)) return 0.0;
    const double x = (current_volume - target_volume) / target_volume;
    const double y = *y_prev + x * dt;
    *y_prev = y;
    const double kp = gain * 2.3 / (25.0 * 0.01);
    const double ki = kp * kp / 16.0;
    return -(kp * x + ki * y) / (x + 1.0);
}

void shk_or_density_scale(int nx, int ny, int nz, const double *du, const double *dv, const double *dw,
                          const double *ax, const double *ay, const double *az, int real_is_double,
                          double *u, double *v, double *w) {
    /* macpressuresolver3.cpp:287-293: velocity *= 1 - (1 - area)*density on the faces
     * the update left active (area != 0 and rho != 0; the others are 0 already).
     * The scale is computed in double and stored in a Real. */
    const double *dens[3] = {du, dv, dw}, *areas[3] = {ax, ay, az};
    double *vel[3] = {u, v, w};
    for (int dim = 0; dim < 3; ++dim) {
        const int64_t nf = (int64_t)(nx + (dim == 0)) * (ny + (dim == 1)) * (nz + (dim == 2));
        for (int64_t f = 0; f < nf; ++f) {
            const double scale = store_real(1.0 - (1.0 - areas[dim][f]) * dens[dim][f], real_is_double);
            vel[dim][f] = real_is_double ? scale * vel[dim][f] : (double)((float)scale * (float)vel[dim][f]);
        }
    }
}

/* ---------------------------------------------------------------- 2D (macpressuresolver2) */

void shk_or_area_fraction2d(int nx, int ny, const double *solid, double eps_solid, int real_is_double,
                            double *ax, double *ay) {
    /* macutility2.cpp:104-111: area = 1 - fraction of the face's two end nodes
     * (x faces: (i,j)-(i,j+1); y faces: (i,j)-(i+1,j)), 0 on the domain walls, eps
     * clamp.  Nodal index i + (nx+1) j. */
    for (int j = 0; j < ny; ++j)
        for (int i = 0; i <= nx; ++i) {
            double area;
            if (i == 0 || i == nx)
                area = 0.0;
            else
                area = 1.0 - shk_or_fraction(solid[i + (int64_t)(nx + 1) * j], solid[i + (int64_t)(nx + 1) * (j + 1)]);
            if (area != 0.0 && area < eps_solid) area = eps_solid;
            ax[i + (int64_t)(nx + 1) * j] = store_real(area, real_is_double);
        }
    for (int j = 0; j <= ny; ++j)
        for (int i = 0; i < nx; ++i) {
            double area;
            if (j == 0 || j == ny)
                area = 0.0;
            else
                area = 1.0 - shk_or_fraction(solid[i + (int64_t)(nx + 1) * j], solid[i + 1 + (int64_t)(nx + 1) * j]);
            if (area != 0.0 && area < eps_solid) area = eps_solid;
            ay[i + (int64_t)nx * j] = store_real(area, real_is_double);
        }
}

void shk_or_curvature2d(int nx, int ny, double dx, const double *fluid, int real_is_double, double *curv) {
    /* macpressuresolver2.cpp:93-102: clamped 5-point Laplacian, the four neighbours
     * summed left to right in Real, centre term and division in double. */
    for (int j = 0; j < ny; ++j)
        for (int i = 0; i < nx; ++i) {
            const int64_t nb[4] = {clampi(i - 1, 0, nx - 1) + (int64_t)nx * j, clampi(i + 1, 0, nx - 1) + (int64_t)nx * j,
                                   i + (int64_t)nx * clampi(j - 1, 0, ny - 1), i + (int64_t)nx * clampi(j + 1, 0, ny - 1)};
            double sum;
            if (real_is_double) {
                sum = fluid[nb[0]];
                for (int q = 1; q < 4; ++q) sum = sum + fluid[nb[q]];
            } else {
                float fs = (float)fluid[nb[0]];
                for (int q = 1; q < 4; ++q) fs = fs + (float)fluid[nb[q]];
                sum = fs;
            }
            const double value = (sum - 4.0 * fluid[i + (int64_t)nx * j]) / (dx * dx);
            curv[i + (int64_t)nx * j] = store_real(value, real_is_double);
        }
}

/* ---------------------------------------------------------------- a6 */

/* Neighbour table of macpressuresolver3.cpp:139-141 / 184-187:
 * query offset, face offset, direction, sign. */
static const int NQ_Q[6][3] = {{1, 0, 0}, {-1, 0, 0}, {0, 1, 0}, {0, -1, 0}, {0, 0, 1}, {0, 0, -1}};
static const int NQ_F[6][3] = {{1, 0, 0}, {0, 0, 0}, {0, 1, 0}, {0, 0, 0}, {0, 0, 1}, {0, 0, 0}};
static const int NQ_DIR[6] = {0, 0, 1, 1, 2, 2};
static const int NQ_SGN[6] = {1, -1, 1, -1, 1, -1};

static inline int out_of_bounds(int nx, int ny, int nz, int i, int j, int k) {
    return i < 0 || i >= nx || j < 0 || j >= ny || k < 0 || k >= nz;
}

int64_t shk_or_label(int nx, int ny, int nz, const double *fluid,
                     const double *ax, const double *ay, const double *az,
                     const double *rx, const double *ry, const double *rz,
                     int64_t *index_map) {
    const double *areas[3] = {ax, ay, az};
    const double *rhos[3] = {rx, ry, rz};
    int64_t index = 0;
    for (int k = 0; k < nz; ++k)
        for (int j = 0; j < ny; ++j)
            for (int i = 0; i < nx; ++i) {
                int inside = 0;
                if (fluid[CELL(i, j, k)] < 0.0) {
                    for (int nq = 0; nq < 6; ++nq) {
                        const int qi = i + NQ_Q[nq][0], qj = j + NQ_Q[nq][1], qk = k + NQ_Q[nq][2];
                        if (out_of_bounds(nx, ny, nz, qi, qj, qk)) continue;
                        if (!(fluid[CELL(qi, qj, qk)] < 0.0)) continue;
                        const int dim = NQ_DIR[nq];
                        const int64_t f = face_index(nx, ny, dim, i + NQ_F[nq][0], j + NQ_F[nq][1], k + NQ_F[nq][2]);
                        if (areas[dim][f] != 0.0 && rhos[dim][f] != 0.0) {
                            inside = 1;
                            break;
                        }
                    }
                }
                index_map[CELL(i, j, k)] = inside ? index++ : -1;
            }
    return index;
}

/* ---------------------------------------------------------------- a7 */

typedef struct {
    int64_t col;
    double val;
} entry_t;

static int assemble_row(int nx, int ny, int nz, double dt, double dx, const double *fluid,
                        const double *const areas[3], const double *const rhos[3],
                        const double *const vel[3], const double *const svel[3],
                        const int64_t *index_map, int i, int j, int k,
                        entry_t *ent, double *rhs_out) {
    const int64_t n_index = index_map[CELL(i, j, k)];
    double rhs = 0.0, diagonal = 0.0;
    int ne = 0;
    for (int nq = 0; nq < 6; ++nq) {
        const int dim = NQ_DIR[nq];
        const int qi = i + NQ_Q[nq][0], qj = j + NQ_Q[nq][1], qk = k + NQ_Q[nq][2];
        const int64_t f = face_index(nx, ny, dim, i + NQ_F[nq][0], j + NQ_F[nq][1], k + NQ_F[nq][2]);
        const double area = areas[dim][f];
        if (!out_of_bounds(nx, ny, nz, qi, qj, qk)) {
            if (area != 0.0) {
                const double rho = rhos[dim][f];
                if (rho != 0.0) {
                    const double value = dt * area / (dx * dx * rho);
                    if (fluid[CELL(qi, qj, qk)] < 0.0) {
                        const double off = -value;
                        if (off != 0.0) {
                            ent[ne].col = index_map[CELL(qi, qj, qk)];
                            ent[ne].val = off;
                            ++ne;
                        }
                    }
                    diagonal += value;
                }
            }
        }
        rhs += -NQ_SGN[nq] * area * vel[dim][f] / dx;
        if (area < 1.0) {
            const double us = svel[dim][f];
            rhs += -NQ_SGN[nq] * (1.0 - area) * us / dx;
        }
    }
    if (diagonal != 0.0) {
        ent[ne].col = n_index;
        ent[ne].val = diagonal;
        ++ne;
    }
    /* RCMatrix rows are kept sorted by column (RCMatrix.cpp:256-279). */
    for (int a = 1; a < ne; ++a) {
        entry_t e = ent[a];
        int b = a - 1;
        while (b >= 0 && ent[b].col > e.col) {
            ent[b + 1] = ent[b];
            --b;
        }
        ent[b + 1] = e;
    }
    *rhs_out = rhs;
    return ne;
}

int64_t shk_or_assemble(int nx, int ny, int nz, double dt, double dx, const double *fluid,
                        const double *ax, const double *ay, const double *az,
                        const double *rx, const double *ry, const double *rz,
                        const double *u, const double *v, const double *w,
                        const double *us, const double *vs, const double *ws,
                        const int64_t *index_map, int64_t N,
                        int64_t *rowptr, int64_t *cols, double *vals, double *rhs) {
    const double *const areas[3] = {ax, ay, az};
    const double *const rhos[3] = {rx, ry, rz};
    const double *const vel[3] = {u, v, w};
    const double *const svel[3] = {us, vs, ws};
    /* cells of each row, in row order */
    int64_t *cell_of_row = (int64_t *)malloc(sizeof(int64_t) * (N > 0 ? N : 1));
    const int64_t G = (int64_t)nx * ny * nz;
    for (int64_t c = 0; c < G; ++c)
        if (index_map[c] >= 0) cell_of_row[index_map[c]] = c;
    int *count = (int *)malloc(sizeof(int) * (N > 0 ? N : 1));
#pragma omp parallel for schedule(static)
    for (int64_t r = 0; r < N; ++r) {
        const int64_t c = cell_of_row[r];
        const int i = (int)(c % nx), j = (int)((c / nx) % ny), k = (int)(c / ((int64_t)nx * ny));
        entry_t ent[7];
        double b;
        count[r] = assemble_row(nx, ny, nz, dt, dx, fluid, areas, rhos, vel, svel, index_map, i, j, k, ent, &b);
    }
    rowptr[0] = 0;
    for (int64_t r = 0; r < N; ++r) rowptr[r + 1] = rowptr[r] + count[r];
#pragma omp parallel for schedule(static)
    for (int64_t r = 0; r < N; ++r) {
        const int64_t c = cell_of_row[r];
        const int i = (int)(c % nx), j = (int)((c / nx) % ny), k = (int)(c / ((int64_t)nx * ny));
        entry_t ent[7];
        double b;
        const int ne = assemble_row(nx, ny, nz, dt, dx, fluid, areas, rhos, vel, svel, index_map, i, j, k, ent, &b);
        for (int e = 0; e < ne; ++e) {
            cols[rowptr[r] + e] = ent[e].col;
            vals[rowptr[r] + e] = ent[e].val;
        }
        rhs[r] = b;
    }
    free(count);
    free(cell_of_row);
    return N > 0 ? rowptr[N] : 0;
}

/* ---------------------------------------------------------------- a10-a13 */

void shk_or_spmv(int64_t N, const int64_t *rowptr, const int64_t *cols, const double *vals,
                 const double *x, double *y) {
    /* sparse_matrix.h:264-275 */
    for (int64_t i = 0; i < N; ++i) {
        double acc = 0;
        for (int64_t p = rowptr[i]; p < rowptr[i + 1]; ++p) acc += vals[p] * x[cols[p]];
        y[i] = acc;
    }
}

static double ddot(int64_t n, const double *x, const double *y) {
    /* textbook sequential ddot (reference CBLAS loop order) */
    double r = 0.0;
    for (int64_t i = 0; i < n; ++i) r += x[i] * y[i];
    return r;
}

static void daxpy(int64_t n, double alpha, const double *x, double *y) {
    for (int64_t i = 0; i < n; ++i) y[i] += alpha * x[i];
}

static double abs_max(int64_t n, const double *x) {
    /* fabs(x[idamax]): first index of the largest |x|; NaNs never compare larger */
    double best = 0.0;
    int64_t idx = 0;
    for (int64_t i = 0; i < n; ++i) {
        const double a = fabs(x[i]);
        if (a > best) {
            best = a;
            idx = i;
        }
    }
    return n > 0 ? fabs(x[idx]) : 0.0;
}

static int row_has(const int64_t *rowptr, const int64_t *cols, int64_t row, int64_t col) {
    for (int64_t p = rowptr[row]; p < rowptr[row + 1]; ++p)
        if (cols[p] == col) return 1;
    return 0;
}

int64_t shk_or_mic0(int64_t N, const int64_t *rowptr, const int64_t *cols, const double *vals,
                    double omega, double min_diag_ratio,
                    int64_t *colstart, int64_t *rowidx, double *value, double *invdiag) {
    /* Modified incomplete Cholesky, level zero (pcg_solver.h:88-189 restated; the
     * algorithm is Bridson, "Fluid Simulation for Computer Graphics", sec. 4.3.2).
     * Column k of L holds the strictly-lower entries A[j,k], j > k, which for a
     * symmetric A are row k's entries right of the diagonal. Fill-in that the
     * zero-level pattern cannot hold is moved, scaled by omega, onto the diagonal. */
    double *adiag = (double *)calloc((size_t)(N > 0 ? N : 1), sizeof(double));
    int64_t nnz = 0;
    for (int64_t k = 0; k < N; ++k) {
        colstart[k] = nnz;
        invdiag[k] = 0.0;
        for (int64_t p = rowptr[k]; p < rowptr[k + 1]; ++p) {
            if (cols[p] > k) {
                rowidx[nnz] = cols[p];
                value[nnz] = vals[p];
                ++nnz;
            } else if (cols[p] == k) {
                adiag[k] = invdiag[k] = vals[p];
            }
        }
    }
    colstart[N] = nnz;
    for (int64_t k = 0; k < N; ++k) {
        if (adiag[k] == 0) continue;
        const double piv = (invdiag[k] < min_diag_ratio * adiag[k]) ? adiag[k] : invdiag[k];
        invdiag[k] = 1 / sqrt(piv);
        for (int64_t p = colstart[k]; p < colstart[k + 1]; ++p) value[p] *= invdiag[k];
        for (int64_t p = colstart[k]; p < colstart[k + 1]; ++p) {
            const int64_t j = rowidx[p];
            const double m = value[p];
            double missing = 0;
            for (int64_t q = colstart[k]; q < colstart[k + 1]; ++q) {
                const int64_t a = rowidx[q];
                if (a < j) {
                    if (!row_has(rowptr, cols, j, a)) missing += value[q];
                } else if (a > j) {
                    int64_t b = colstart[j];
                    while (b < colstart[j + 1] && rowidx[b] < a) ++b;
                    if (b < colstart[j + 1] && rowidx[b] == a)
                        value[b] -= m * value[q];
                    else
                        missing += value[q];
                }
            }
            invdiag[j] -= m * m;
            invdiag[j] -= omega * m * missing;
        }
    }
    free(adiag);
    return nnz;
}

static void solve_lower(int64_t N, const int64_t *colstart, const int64_t *rowidx, const double *value,
                        const double *invdiag, const double *rhs, double *out) {
    /* L y = rhs (pcg_solver.h:195-207) */
    memcpy(out, rhs, sizeof(double) * (size_t)N);
    for (int64_t i = 0; i < N; ++i) {
        out[i] *= invdiag[i];
        for (int64_t p = colstart[i]; p < colstart[i + 1]; ++p) out[rowidx[p]] -= value[p] * out[i];
    }
}

static void solve_lower_transpose(int64_t N, const int64_t *colstart, const int64_t *rowidx,
                                  const double *value, const double *invdiag, double *x) {
    /* L^T x = x (pcg_solver.h:210-223) */
    for (int64_t i = N - 1; i >= 0; --i) {
        for (int64_t p = colstart[i]; p < colstart[i + 1]; ++p) x[i] -= value[p] * x[rowidx[p]];
        x[i] *= invdiag[i];
    }
}

int shk_or_pcg(int64_t N, const int64_t *rowptr, const int64_t *cols, const double *vals_in,
               const double *b, double *x, double tol, int64_t max_iterations, int mode,
               double mic_omega, double mic_min_diag_ratio,
               int64_t *iters, double *reresid) {
    /* PCGSolver::solve (pcg_solver.h:246-295) driven as pcg.cpp:45-73 drives it. */
    if (tol < 1e-30) tol = 1e-30; /* set_solver_parameters, pcg_solver.h:235 */
    const size_t nb = sizeof(double) * (size_t)(N > 0 ? N : 1);
    const int64_t nnz = rowptr[N];
    /* pcg.cpp:47-52 copies the RC matrix into the solver's own storage and
     * pcg_solver.h:273 builds a second, compressed copy; keep one such copy so the
     * timed work includes it. */
    double *vals = (double *)malloc(sizeof(double) * (size_t)(nnz > 0 ? nnz : 1));
    memcpy(vals, vals_in, sizeof(double) * (size_t)nnz);
    double *r = (double *)malloc(nb), *z = (double *)malloc(nb), *s = (double *)malloc(nb);
    double *diag = NULL;
    int64_t *colstart = NULL, *rowidx = NULL;
    double *lval = NULL, *linv = NULL;
    int status = 1;
    memset(x, 0, nb);
    memcpy(r, b, sizeof(double) * (size_t)N);
    double res = abs_max(N, r);
    const double res0 = res;
    if (res == 0) {
        *iters = 0;
        *reresid = res;
        status = 0;
        goto done;
    }
    {
        const double tol_abs = tol * res;
        if (mode == SHK_OR_REFERENCE || mode == SHK_OR_MIC0) {
            colstart = (int64_t *)malloc(sizeof(int64_t) * (size_t)(N + 1));
            rowidx = (int64_t *)malloc(sizeof(int64_t) * (size_t)(nnz > 0 ? nnz : 1));
            lval = (double *)malloc(sizeof(double) * (size_t)(nnz > 0 ? nnz : 1));
            linv = (double *)malloc(nb);
            shk_or_mic0(N, rowptr, cols, vals, mic_omega, mic_min_diag_ratio, colstart, rowidx, lval, linv);
        }
        if (mode == SHK_OR_JACOBI) {
            diag = (double *)malloc(nb);
            for (int64_t i = 0; i < N; ++i) {
                diag[i] = 0.0;
                for (int64_t p = rowptr[i]; p < rowptr[i + 1]; ++p)
                    if (cols[p] == i) diag[i] = vals[p];
            }
        }
#define APPLY_PRECONDITIONER(rv, zv)                                                   \
    do {                                                                               \
        if (mode == SHK_OR_REFERENCE) {                                                \
            solve_lower(N, colstart, rowidx, lval, linv, rv, zv);                      \
            solve_lower_transpose(N, colstart, rowidx, lval, linv, zv);                \
            memcpy(zv, rv, sizeof(double) * (size_t)N);                                \
        } else if (mode == SHK_OR_MIC0) {                                              \
            solve_lower(N, colstart, rowidx, lval, linv, rv, zv);                      \
            solve_lower_transpose(N, colstart, rowidx, lval, linv, zv);                \
        } else if (mode == SHK_OR_JACOBI) {                                            \
            for (int64_t ii = 0; ii < N; ++ii) zv[ii] = rv[ii] / diag[ii];             \
        } else {                                                                       \
            memcpy(zv, rv, sizeof(double) * (size_t)N);                                \
        }                                                                              \
    } while (0)
        APPLY_PRECONDITIONER(r, z);
        double rho = ddot(N, z, r);
        if (rho == 0) {
            *iters = 0;
            *reresid = res;
            status = 3;
            goto done;
        }
        if (rho != rho) {
            *iters = 0;
            *reresid = res;
            status = 2;
            goto done;
        }
        memcpy(s, z, sizeof(double) * (size_t)N);
        int64_t it;
        for (it = 0; it < max_iterations; ++it) {
            shk_or_spmv(N, rowptr, cols, vals, s, z);
            const double alpha = rho / ddot(N, s, z);
            daxpy(N, alpha, s, x);
            daxpy(N, -alpha, z, r);
            res = abs_max(N, r);
            if (res <= tol_abs) {
                *iters = it + 1;
                *reresid = res / res0;
                status = 0;
                goto done;
            }
            APPLY_PRECONDITIONER(r, z);
            const double rho_new = ddot(N, z, r);
            const double beta = rho_new / rho;
            daxpy(N, beta, s, z);
            double *t = s;
            s = z;
            z = t;
            rho = rho_new;
        }
#undef APPLY_PRECONDITIONER
        *iters = it;
        *reresid = res / res0;
        status = 1;
    }
done:
    free(vals);
    free(r);
    free(z);
    free(s);
    free(diag);
    free(colstart);
    free(rowidx);
    free(lval);
    free(linv);
    return status;
}

/* ---------------------------------------------------------------- threaded CG */

/* The same Krylov recurrences as shk_or_pcg in modes SHK_OR_CG and SHK_OR_JACOBI
 * (pcg_solver.h:246-295 with z = r, or z = r / diag(A)), spread over the OpenMP
 * threads: the CG-equivalent CPU baseline on T cores, and the FP64 truth solves of
 * large grids.  Deterministic for a fixed thread count: every thread reduces its own
 * static row range and the partials are summed in thread order.  Each iteration is
 * three passes over the rows: q = A s with s.q; x += a s, r -= a q, z, z.r and max|r|;
 * s = z + b s.  Column indices are narrowed to 32 bits once at the start (N < 2^31). */
int shk_or_pcg_threaded(int64_t N, const int64_t *rowptr, const int64_t *cols, const double *vals,
                        const double *b, double *x, double tol, int64_t max_iterations, int mode,
                        int64_t *iters, double *reresid) {
    if (mode != SHK_OR_CG && mode != SHK_OR_JACOBI) return -1;
    if (N >= ((int64_t)1 << 31)) return -1;
    if (tol < 1e-30) tol = 1e-30;
    int nt = 1;
#ifdef _OPENMP
    nt = omp_get_max_threads();
#endif
    const int64_t nnz = N > 0 ? rowptr[N] : 0;
    int32_t *c32 = (int32_t *)malloc(sizeof(int32_t) * (size_t)(nnz > 0 ? nnz : 1));
    const size_t nb = sizeof(double) * (size_t)(N > 0 ? N : 1);
    double *r = (double *)malloc(nb), *z = (double *)malloc(nb), *s = (double *)malloc(nb),
           *q = (double *)malloc(nb), *inv = (double *)malloc(nb); /* diag(A) */
    double *part = (double *)calloc((size_t)(3 * nt), sizeof(double));
    int status = 1;
    double res0 = 0.0, res = 0.0, rho = 0.0;
#pragma omp parallel
    {
        int t = 0;
#ifdef _OPENMP
        t = omp_get_thread_num();
#endif
        const int64_t lo = N * t / nt, hi = N * (t + 1) / nt;
        double m = 0.0, zr = 0.0;
        for (int64_t i = lo; i < hi; ++i) {
            double d = 0.0;
            for (int64_t p = rowptr[i]; p < rowptr[i + 1]; ++p) {
                c32[p] = (int32_t)cols[p];
                if (cols[p] == i) d = vals[p];
            }
            inv[i] = d;
            x[i] = 0.0;
            r[i] = b[i];
            z[i] = mode == SHK_OR_JACOBI ? b[i] / d : b[i];
            s[i] = z[i];
            zr += z[i] * r[i];
            if (fabs(r[i]) > m) m = fabs(r[i]);
        }
        part[3 * t] = m;
        part[3 * t + 1] = zr;
    }
    for (int t = 0; t < nt; ++t) {
        if (part[3 * t] > res0) res0 = part[3 * t];
        rho += part[3 * t + 1];
    }
    res = res0;
    int64_t it = 0;
    if (res0 == 0.0) {
        status = 0;
        goto done;
    }
    if (rho == 0.0 || rho != rho) {
        status = rho != rho ? 2 : 3;
        goto done;
    }
    {
        const double tol_abs = tol * res0;
        for (it = 0; it < max_iterations; ++it) {
            double sq = 0.0;
#pragma omp parallel
            {
                int t = 0;
#ifdef _OPENMP
                t = omp_get_thread_num();
#endif
                const int64_t lo = N * t / nt, hi = N * (t + 1) / nt;
                double acc_sq = 0.0;
                for (int64_t i = lo; i < hi; ++i) {
                    double acc = 0.0;
                    for (int64_t p = rowptr[i]; p < rowptr[i + 1]; ++p) acc += vals[p] * s[c32[p]];
                    q[i] = acc;
                    acc_sq += s[i] * acc;
                }
                part[3 * t] = acc_sq;
            }
            for (int t = 0; t < nt; ++t) sq += part[3 * t];
            const double alpha = rho / sq;
            double rho_new = 0.0;
            res = 0.0;
#pragma omp parallel
            {
                int t = 0;
#ifdef _OPENMP
                t = omp_get_thread_num();
#endif
                const int64_t lo = N * t / nt, hi = N * (t + 1) / nt;
                double m = 0.0, zr = 0.0;
                for (int64_t i = lo; i < hi; ++i) {
                    x[i] += alpha * s[i];
                    const double ri = r[i] - alpha * q[i];
                    r[i] = ri;
                    const double zi = mode == SHK_OR_JACOBI ? ri / inv[i] : ri;
                    z[i] = zi;
                    zr += zi * ri;
                    if (fabs(ri) > m) m = fabs(ri);
                }
                part[3 * t] = m;
                part[3 * t + 1] = zr;
            }
            for (int t = 0; t < nt; ++t) {
                if (part[3 * t] > res) res = part[3 * t];
                rho_new += part[3 * t + 1];
            }
            if (res != res) {
                status = 2;
                goto done;
            }
            if (res <= tol_abs) {
                ++it;
                status = 0;
                goto done;
            }
            const double beta = rho_new / rho;
            rho = rho_new;
#pragma omp parallel for schedule(static)
            for (int64_t i = 0; i < N; ++i) s[i] = z[i] + beta * s[i];
        }
        status = 1;
    }
done:
    *iters = it;
    *reresid = res0 > 0 ? res / res0 : res;
    free(c32);
    free(r);
    free(z);
    free(s);
    free(q);
    free(inv);
    free(part);
    return status;
}

/* ---------------------------------------------------------------- a14, a15 */

void shk_or_scatter(int nx, int ny, int nz, const int64_t *index_map, const double *x,
                    int real_is_double, double *pressure) {
    const int64_t G = (int64_t)nx * ny * nz;
    for (int64_t c = 0; c < G; ++c)
        pressure[c] = index_map[c] >= 0 ? store_real(x[index_map[c]], real_is_double) : 0.0;
}

void shk_or_update_velocity(int nx, int ny, int nz, double dt, double dx,
                            const double *pressure,
                            const double *ax, const double *ay, const double *az,
                            const double *rx, const double *ry, const double *rz,
                            const double *us, const double *vs, const double *ws,
                            int real_is_double, double *u, double *v, double *w) {
    /* macpressuresolver3.cpp:269-284.  With Real = float the pressure difference
     * is a float subtraction, the increment is rounded to float by
     * array3::subtract's `const T&` parameter, and the update is a float
     * subtraction (array3.h:679-687). */
    const double *areas[3] = {ax, ay, az};
    const double *rhos[3] = {rx, ry, rz};
    const double *svel[3] = {us, vs, ws};
    double *vel[3] = {u, v, w};
    const int n[3] = {nx, ny, nz};
    for (int dim = 0; dim < 3; ++dim) {
        const int fx = nx + (dim == 0), fy = ny + (dim == 1), fz = nz + (dim == 2);
#pragma omp parallel for schedule(static)
        for (int k = 0; k < fz; ++k)
            for (int j = 0; j < fy; ++j)
                for (int i = 0; i < fx; ++i) {
                    const int64_t f = face_index(nx, ny, dim, i, j, k);
                    const int pi[3] = {i, j, k};
                    const double rho = rhos[dim][f];
                    const double solid_u = svel[dim][f];
                    if (areas[dim][f] != 0.0 && rho != 0.0) {
                        if (pi[dim] == 0 || pi[dim] == n[dim]) {
                            vel[dim][f] = store_real(solid_u, real_is_double);
                        } else {
                            const double p1 = pressure[CELL(i, j, k)];
                            const double p0 = pressure[CELL(i - (dim == 0), j - (dim == 1), k - (dim == 2))];
                            if (real_is_double) {
                                vel[dim][f] -= dt * (p1 - p0) / (rho * dx);
                            } else {
                                const float dp = (float)p1 - (float)p0;
                                const float inc = (float)(dt * (double)dp / (rho * dx));
                                vel[dim][f] = (double)((float)vel[dim][f] - inc);
                            }
                        }
                    } else {
                        vel[dim][f] = 0.0;
                    }
                }
    }
}

double shk_or_divergence(int nx, int ny, int nz, double dx,
                         const double *ax, const double *ay, const double *az,
                         const double *u, const double *v, const double *w,
                         const double *us, const double *vs, const double *ws,
                         const int64_t *index_map, double *div) {
    const double *areas[3] = {ax, ay, az};
    const double *vel[3] = {u, v, w};
    const double *svel[3] = {us, vs, ws};
    double best = 0.0;
    for (int k = 0; k < nz; ++k)
        for (int j = 0; j < ny; ++j)
            for (int i = 0; i < nx; ++i) {
                const int64_t row = index_map[CELL(i, j, k)];
                if (row < 0) continue;
                double acc = 0.0;
                for (int nq = 0; nq < 6; ++nq) {
                    const int dim = NQ_DIR[nq];
                    const int64_t f = face_index(nx, ny, dim, i + NQ_F[nq][0], j + NQ_F[nq][1], k + NQ_F[nq][2]);
                    const double area = areas[dim][f];
                    acc += -NQ_SGN[nq] * area * vel[dim][f] / dx;
                    if (area < 1.0) acc += -NQ_SGN[nq] * (1.0 - area) * svel[dim][f] / dx;
                }
                if (div) div[row] = -acc;
                if (fabs(acc) > best) best = fabs(acc);
            }
    return best;
}

/* ---------------------------------------------------------------- whole project() */

int shk_or_project(int nx, int ny, int nz, double dt, double dx,
                   const double *solid, const double *fluid,
                   double *u, double *v, double *w,
                   const double *us, const double *vs, const double *ws,
                   double tol, int64_t max_iterations, int mode, int real_is_double,
                   double eps_fluid, double eps_solid,
                   double *pressure, int64_t *n_unknowns, int64_t *iters, double *reresid,
                   double *t_ms) {
    const int64_t G = (int64_t)nx * ny * nz;
    const int64_t F0 = (int64_t)(nx + 1) * ny * nz, F1 = (int64_t)nx * (ny + 1) * nz, F2 = (int64_t)nx * ny * (nz + 1);
    const double t0 = now_ms();
    double *ax = (double *)malloc(sizeof(double) * F0), *ay = (double *)malloc(sizeof(double) * F1),
           *az = (double *)malloc(sizeof(double) * F2);
    double *rx = (double *)malloc(sizeof(double) * F0), *ry = (double *)malloc(sizeof(double) * F1),
           *rz = (double *)malloc(sizeof(double) * F2);
    int64_t *index_map = (int64_t *)malloc(sizeof(int64_t) * G);
    shk_or_area_fraction(nx, ny, nz, solid, eps_solid, real_is_double, ax, ay, az);
    shk_or_fluid_fraction(nx, ny, nz, fluid, eps_fluid, real_is_double, rx, ry, rz);
    const int64_t N = shk_or_label(nx, ny, nz, fluid, ax, ay, az, rx, ry, rz, index_map);
    *n_unknowns = N;
    int status = 0;
    *iters = 0;
    *reresid = 0.0;
    memset(pressure, 0, sizeof(double) * (size_t)G);
    double t1 = now_ms(), t2 = t1, t3 = t1;
    if (N) {
        int64_t *rowptr = (int64_t *)malloc(sizeof(int64_t) * (size_t)(N + 1));
        int64_t *cols = (int64_t *)malloc(sizeof(int64_t) * (size_t)(7 * N));
        double *vals = (double *)malloc(sizeof(double) * (size_t)(7 * N));
        double *rhs = (double *)malloc(sizeof(double) * (size_t)N);
        double *x = (double *)malloc(sizeof(double) * (size_t)N);
        shk_or_assemble(nx, ny, nz, dt, dx, fluid, ax, ay, az, rx, ry, rz, u, v, w, us, vs, ws, index_map, N,
                        rowptr, cols, vals, rhs);
        t1 = now_ms();
        status = shk_or_pcg(N, rowptr, cols, vals, rhs, x, tol, max_iterations, mode, 0.97, 0.25, iters, reresid);
        t2 = now_ms();
        shk_or_scatter(nx, ny, nz, index_map, x, real_is_double, pressure);
        shk_or_update_velocity(nx, ny, nz, dt, dx, pressure, ax, ay, az, rx, ry, rz, us, vs, ws, real_is_double, u,
                               v, w);
        t3 = now_ms();
        free(rowptr);
        free(cols);
        free(vals);
        free(rhs);
        free(x);
    }
    if (t_ms) {
        t_ms[0] = t1 - t0;
        t_ms[1] = t2 - t1;
        t_ms[2] = t3 - t2;
        t_ms[3] = t3 - t0;
    }
    free(ax);
    free(ay);
    free(az);
    free(rx);
    free(ry);
    free(rz);
    free(index_map);
    return status;
}

int shk_or_set_threads(int n) {
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
    return omp_get_max_threads();
#else
    (void)n;
    return 1;
#endif
}
