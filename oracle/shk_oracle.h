/*
 * shk_oracle.h -- CPU FP64 restatement of the Shiokaze MAC liquid pressure
 * projection (macpressuresolver3 + pcg linsolver).
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the parity checker for the HIP
 * product path and the CPU baseline that bench.py times beside it.  Nothing in
 * the product (shiokaze_amd/) links, loads or calls it.
 *
 * Every function cites the reference file:line it restates.  Fields are dense,
 * x-fastest:
 *   cell   (i,j,k), i<nx, j<ny, k<nz        -> i + nx*(j + ny*k)
 *   face d (i,j,k) on a grid one larger along d -> i + fx*(j + fy*k),
 *          fx = nx+(d==0), fy = ny+(d==1)
 *   node   (i,j,k), i<=nx, j<=ny, k<=nz     -> i + (nx+1)*(j + (ny+1)*k)
 * Arrays are double on this interface.  `real_is_double == 0` reproduces the
 * reference's `using Real = float` storage (include/shiokaze/core/config.h:34):
 * face fractions, the scattered pressure and the updated velocity are rounded
 * through float exactly where the reference stores them into Real arrays.
 */
#ifndef SHK_ORACLE_H
#define SHK_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* utility::fraction  (include/shiokaze/utility/utility.h:174-182) */
double shk_or_fraction(double phi0, double phi1);
/* utility::get_area  (include/shiokaze/utility/utility.h:191-226); arguments are
 * isosurf[0][0], isosurf[1][0], isosurf[1][1], isosurf[0][1]. */
double shk_or_get_area(double q00, double q10, double q11, double q01);

/* macutility3::compute_area_fraction (src/utility/macutility3.cpp:98-171) with
 * enclose_domain_boundary = true.  solid: nodal (nx+1)(ny+1)(nz+1). */
void shk_or_area_fraction(int nx, int ny, int nz, const double *solid, double eps_solid,
                          int real_is_double, double *ax, double *ay, double *az);
/* macutility3::compute_fluid_fraction (src/utility/macutility3.cpp:172-200).
 * fluid: cell-centred level set. */
void shk_or_fluid_fraction(int nx, int ny, int nz, const double *fluid, double eps_fluid,
                           int real_is_double, double *rx, double *ry, double *rz);

/* Unknown labelling, macpressuresolver3.cpp:134-167.  index_map[cell] = -1 or
 * the row index (numbered in x-fastest linear order).  Returns N. */
int64_t shk_or_label(int nx, int ny, int nz, const double *fluid,
                     const double *ax, const double *ay, const double *az,
                     const double *rx, const double *ry, const double *rz,
                     int64_t *index_map);

/* Assembly, macpressuresolver3.cpp:176-216 with RCMatrix::add_to_element
 * (src/math/RCMatrix.cpp:256-279: exact zeros are not stored, rows sorted).
 * rowptr[N+1], cols/vals sized >= 7N.  Returns nnz. */
int64_t shk_or_assemble(int nx, int ny, int nz, double dt, double dx, const double *fluid,
                        const double *ax, const double *ay, const double *az,
                        const double *rx, const double *ry, const double *rz,
                        const double *u, const double *v, const double *w,
                        const double *us, const double *vs, const double *ws,
                        const int64_t *index_map, int64_t N,
                        int64_t *rowptr, int64_t *cols, double *vals, double *rhs);

/* Preconditioner / solver variants of shk_or_pcg. */
enum {
    SHK_OR_REFERENCE = 0, /* pcg_solver.h:246-295 verbatim semantics: MIC(0) factor and
                             two triangular solves per iteration whose result is then
                             overwritten (z = r, pcg_solver.h:379-384) */
    SHK_OR_CG = 1,        /* the same iterates without the discarded work */
    SHK_OR_JACOBI = 2,    /* z = r / diag(A) */
    SHK_OR_MIC0 = 3       /* z = (L L^T)^-1 r with the MIC(0) factor actually applied */
};
/* PCGSolver::solve (local/include/pcgsolver/pcg_solver.h:246-295), x0 = 0, stop when
 * max|r| <= tol * max|b|.  Returns 0 converged, 1 iteration cap, 2 NaN, 3 rho == 0.
 * iters / reresid follow the reference's outputs (reresid = max|r| / max|b|). */
int shk_or_pcg(int64_t N, const int64_t *rowptr, const int64_t *cols, const double *vals,
               const double *b, double *x, double tol, int64_t max_iterations, int mode,
               double mic_omega, double mic_min_diag_ratio,
               int64_t *iters, double *reresid);

/* y = A x (sparse_matrix.h:264-275) */
void shk_or_spmv(int64_t N, const int64_t *rowptr, const int64_t *cols, const double *vals,
                 const double *x, double *y);

/* The CG and Jacobi-PCG modes of shk_or_pcg on all OpenMP threads (three fused
 * passes per iteration, deterministic per thread count): the CG-equivalent CPU
 * baseline on T cores and the FP64 truth solves of large grids.  Returns -1 for
 * other modes or N >= 2^31. */
int shk_or_pcg_threaded(int64_t N, const int64_t *rowptr, const int64_t *cols, const double *vals,
                        const double *b, double *x, double tol, int64_t max_iterations, int mode,
                        int64_t *iters, double *reresid);

/* MIC(0) factor in CSC lower form (pcg_solver.h:88-189 restated): colstart[N+1],
 * rowidx/value sized nnz; invdiag[N].  Exposed for property tests. */
int64_t shk_or_mic0(int64_t N, const int64_t *rowptr, const int64_t *cols, const double *vals,
                    double omega, double min_diag_ratio,
                    int64_t *colstart, int64_t *rowidx, double *value, double *invdiag);

/* Scatter x to a dense cell array (macpressuresolver3.cpp:262-265); non-unknowns
 * are 0.  With real_is_double == 0 the values are rounded through float. */
void shk_or_scatter(int nx, int ny, int nz, const int64_t *index_map, const double *x,
                    int real_is_double, double *pressure);

/* Velocity update (macpressuresolver3.cpp:269-284) in place on u, v, w. */
void shk_or_update_velocity(int nx, int ny, int nz, double dt, double dx,
                            const double *pressure,
                            const double *ax, const double *ay, const double *az,
                            const double *rx, const double *ry, const double *rz,
                            const double *us, const double *vs, const double *ws,
                            int real_is_double, double *u, double *v, double *w);

/* Weighted divergence of the face field on the unknowns: the negative of the
 * rhs that macpressuresolver3.cpp:209-213 would assemble from it.  div[row]
 * sized N.  Returns max |div|. */
double shk_or_divergence(int nx, int ny, int nz, double dx,
                         const double *ax, const double *ay, const double *az,
                         const double *u, const double *v, const double *w,
                         const double *us, const double *vs, const double *ws,
                         const int64_t *index_map, double *div);

/* Whole project() (macpressuresolver3.cpp:54-299) for the default parameters, used
 * as the CPU baseline: fractions, labels, assembly, solve, scatter, update.
 * Work arrays are allocated inside.  Returns the solver status; timings in ms
 * are written to t_ms[4] = {fractions+labels+assembly, solve, scatter+update, total}. */
int shk_or_project(int nx, int ny, int nz, double dt, double dx,
                   const double *solid, const double *fluid,
                   double *u, double *v, double *w,
                   const double *us, const double *vs, const double *ws,
                   double tol, int64_t max_iterations, int mode, int real_is_double,
                   double eps_fluid, double eps_solid,
                   double *pressure, int64_t *n_unknowns, int64_t *iters, double *reresid,
                   double *t_ms);

/* OpenMP threads for the fraction / assembly / update loops (the reference
 * threads those loops with NUM_THREAD = cpu count, wscript:75; the solve is
 * serial there and here).  n <= 0 only queries.  Returns the thread count. */
/* project() branches (macpressuresolver3.cpp:75-124, 220-236, 287-293) */
void shk_or_solid_density_blend(int nx, int ny, int nz, const double *du, const double *dv, const double *dw,
                                int real_is_double, double *ax, double *ay, double *az);
void shk_or_first_order(int nx, int ny, int nz, int second_order_solid, int second_order_fluid,
                        double *ax, double *ay, double *az, double *rx, double *ry, double *rz);
void shk_or_curvature(int nx, int ny, int nz, double dx, const double *fluid, int real_is_double, double *curv);
void shk_or_surface_tension(int nx, int ny, int nz, double dt, double dx, double kappa, const double *fluid,
                            const double *curv, const double *rx, const double *ry, const double *rz,
                            int real_is_double, double *u, double *v, double *w);
double shk_or_volume_correction(double gain, double current_volume, double target_volume, double dt,
                                double *y_prev);
void shk_or_density_scale(int nx, int ny, int nz, const double *du, const double *dv, const double *dw,
                          const double *ax, const double *ay, const double *az, int real_is_double,
                          double *u, double *v, double *w);

/* 2D (macutility2.cpp:104-111, macpressuresolver2.cpp:93-102) */
void shk_or_area_fraction2d(int nx, int ny, const double *solid, double eps_solid, int real_is_double,
                            double *ax, double *ay);
void shk_or_curvature2d(int nx, int ny, double dx, const double *fluid, int real_is_double, double *curv);

int shk_or_set_threads(int n);

/* the velocity half of the liquid step around the projection (shk_liquid_oracle.c;
 * macliquid3.cpp:342-391): dense float face arrays with one uint8 activity mask each */
int64_t shk_or_extrapolate(int nx, int ny, int nz, int dim, float *a, uint8_t *m, int count);
void shk_or_set_weno(int on);
int64_t shk_or_redistance(int nx, int ny, int nz, double dx, float *phi, uint8_t *band, int width, double rate,
                          int rk2);
int64_t shk_or_extrapolate_levelset(int nx, int ny, int nz, double dx, const float *solid, const uint8_t *solid_band,
                                    float *phi, uint8_t *band, float background, int wall, int horizontal);
void shk_or_constrain_velocity(int nx, int ny, int nz, double dx, const float *solid, float *u, float *v,
                               float *w, const uint8_t *mu, const uint8_t *mv, const uint8_t *mw);
double shk_or_compute_max_u(int nx, int ny, int nz, const float *u, const float *v, const float *w,
                            const uint8_t *mu, const uint8_t *mv, const uint8_t *mw);
void shk_or_advect_velocity(int nx, int ny, int nz, double dx, double dt, float *u, float *v, float *w,
                            const uint8_t *mu, const uint8_t *mv, const uint8_t *mw, const float *fluid,
                            int maccormack, int trim);
void shk_or_advect_levelset(int nx, int ny, int nz, double dx, double dt, float *phi, const uint8_t *mphi,
                            const float *u, const float *v, const float *w, const uint8_t *mu, const uint8_t *mv,
                            const uint8_t *mw, float background, int maccormack, int trim);
void shk_or_add_gravity(int nx, int ny, int nz, double dt, double gx, double gy, double gz, float *u, float *v,
                        float *w, uint8_t *mu, uint8_t *mv, uint8_t *mw);
void shk_or_project_activity(int nx, int ny, int nz, const double *au, const double *av, const double *aw,
                             const double *ru, const double *rv, const double *rw, uint8_t *mu, uint8_t *mv,
                             uint8_t *mw);

#ifdef __cplusplus
}
#endif
#endif
