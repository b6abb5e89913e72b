/*
 * shk_projection.h -- C-ABI of libshk_projection.so, the MI355X (gfx950)
 * MAC liquid pressure projection.
 *
 * The library replaces two Shiokaze plugins behind their own contracts:
 *   - macproject3_interface (include/shiokaze/projection/macproject3_interface.h:39-116),
 *     implemented by macpressuresolver3 (src/projection/macpressuresolver3.cpp:40-372);
 *   - RCMatrix_solver_interface<size_t,double>::solve
 *     (include/shiokaze/linsolver/RCMatrix_solver.h:77), implemented by the "pcg"
 *     plugin (src/linsolver/pcg.cpp:32-82).
 *
 * Conventions
 *   - Plain pointers and sizes only.  Field pointers passed to shk_project /
 *     shk_solve_csr / shk_divergence are DEVICE pointers (hipMalloc'd or any
 *     allocation visible to the context's device) unless the context was given
 *     "HostPointers=1", in which case they are host pointers and the library
 *     stages them (the PCIe copies are then part of the call).
 *   - Dense x-fastest layouts, the same as the reference's array3 / macarray3
 *     shapes (include/shiokaze/math/shape.h:700-716):
 *       cell  nx*ny*nz            index i + nx*(j + ny*k)
 *       u     (nx+1)*ny*nz        index i + (nx+1)*(j + ny*k)
 *       v     nx*(ny+1)*nz        index i + nx*(j + (ny+1)*k)
 *       w     nx*ny*(nz+1)        index i + nx*(j + ny*k)
 *       solid (nx+1)*(ny+1)*(nz+1) nodal level set
 *     Velocity, level sets and pressure are float (the reference's Real,
 *     include/shiokaze/core/config.h:34); the linear system is double.
 *   - Every entry point returns an shk_status; nothing calls exit() or throws.
 *     shk_last_error() returns the message of the most recent failure.
 *   - One context owns one HIP stream and its device workspace.  A context is
 *     not thread-safe; separate contexts may be used from separate threads.
 *     Calls are synchronous on return.
 */
#ifndef SHK_PROJECTION_H
#define SHK_PROJECTION_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SHK_ABI_VERSION 3

typedef struct shk_ctx shk_ctx;

typedef enum shk_status {
    SHK_OK = 0,
    /* iteration cap reached; the reference returns silently with the residual
       (local/include/pcgsolver/pcg_solver.h:292-294) */
    SHK_NOT_CONVERGED = 1,
    /* NaN in the Krylov scalars (the reference prints "NaN Error !!" and
       returns 0 iterations, pcg_solver.h:267-271) */
    SHK_NAN = 2,
    SHK_BAD_ARG = 3,
    SHK_HIP_ERROR = 4,
    SHK_NO_DEVICE = 5,
    /* rho == 0 or p.Ap <= 0: the system is singular or not SPD on the unknowns
       (pcg_solver.h:263-266 returns false with 0 iterations for rho == 0) */
    SHK_BREAKDOWN = 6,
    SHK_OUT_OF_MEMORY = 7
} shk_status;

/* Result of a solve.  count / reresid are RCMatrix_solver_interface::Result's
 * fields (include/shiokaze/linsolver/RCMatrix_solver.h:44-57); the rest are the
 * figures this library adds. */
typedef struct shk_result {
    int64_t count;          /* Krylov iterations (all outer steps summed) */
    double reresid;         /* max|r|/max|b| of the residual the stop test used */
    double true_relres_inf; /* max|b - A x| / max|b|, FP64 */
    double true_relres_l2;  /* ||b - A x||_2 / ||b||_2, FP64 */
    double b_inf;           /* max|b| */
    int64_t unknowns;       /* N, the number of rows */
    int64_t active_bricks;  /* 8x8x8 bricks that hold at least one unknown */
    int32_t outer;          /* refinement steps (1 for the FP64 path) */
    int32_t status;         /* shk_status of the solve */
    int32_t fallback;       /* 1 if the learned flexible CG handed over to plain PCG */
    int32_t precision;      /* 0 = fp64, 1 = mixed */
    double ms_setup;        /* fractions, labels, bricks, coefficients, rhs */
    double ms_solve;        /* Krylov loop incl. refinement */
    double ms_update;       /* pressure scatter + velocity update */
    double volume_correct_rhs; /* the constant added to every rhs row by the volume
                                  correction (macpressuresolver3.cpp:231-235), 0 if off */
    /* RCMatrix_utility::report of the assembled matrix (RCMatrix_utility.h:98-146,
     * called at macpressuresolver3.cpp:236), filled by shk_project / shk_project2d
     * when Report=1 (report_valid = 1), else zero. */
    int64_t report_nnz;              /* nonzero entries */
    int32_t report_min_row;          /* fewest nonzeros in a row */
    int32_t report_max_row;          /* most nonzeros in a row */
    double report_min_diag;          /* smallest diagonal entry */
    double report_max_diag;          /* largest diagonal entry */
    double report_max_offdiag_ratio; /* worst max|off-diagonal| / diagonal of a row */
    double report_symmetry_error;    /* max |A_ij - A_ji| (0: one coupling per face) */
    int32_t report_has_nan;          /* 1 if a matrix entry is NaN */
    int32_t report_valid;
} shk_result;

/* ------------------------------------------------------------------ context */

/* macproject3_interface::initialize(shape3, dx) (macproject3_interface.h:104-112).
 * device < 0 selects the current HIP device. */
int shk_create(int nx, int ny, int nz, double dx, int device, shk_ctx **out);
void shk_destroy(shk_ctx *ctx);

/* configure() keys (same names and defaults as the reference):
 *   Residual (1e-4)          pcg.cpp:40, 76       relative max-norm stop
 *   MaxIterations (30000)    pcg.cpp:43, 77
 *   ModifiedIC, MinDiagRatio pcg.cpp:41-42        accepted; the GPU path has no MIC(0)
 *   SecondOrderAccurateFluid (1), SecondOrderAccurateSolid (1)
 *                            macpressuresolver3.cpp:302-303
 *   EpsFluid (1e-2), EpsSolid (1e-2)   macutility3.cpp:502-503
 *   Gain (1)                 macpressuresolver3.cpp:304, 350  volume-correction rate;
 *                            active once shk_set_target_volume gives a nonzero target
 *   WarmStart (0)            macpressuresolver3.cpp:305, 238-260  solve for the change
 *                            from the previous projection's pressure
 * Keys of this library:
 *   Precision      fp64 | mixed            (fp64)
 *   Preconditioner none | jacobi | unet | mg   (jacobi)
 *                  none: the reference's effective iteration (plain CG, z = r);
 *                  jacobi: diag(A) as symmetric scaling; unet: flexible CG with the
 *                  learned network below; mg: CG with an aggregation-multigrid V-cycle
 *                  (mixed precision, the counterpart of the reference's "amg" LinSolver,
 *                  src/linsolver/amg.cpp:43-166).  shk_solve_csr takes none, jacobi
 *                  and mg (pairwise algebraic aggregation of the CSR matrix); unet
 *                  returns SHK_BAD_ARG there.
 *   MgOmega, MgAlpha  multigrid damped-Jacobi weight (0.9) and coarse-correction
 *                  factor (1.5), each in (0, 2)
 *   OuterMax       refinement steps for mixed (3)
 *   InnerResidual  relative stop of each mixed inner solve (1e-6, floor)
 *   CheckInterval  iterations launched between host convergence polls (32)
 *   HostPointers   0 | 1                   (0)
 *   UnetWeights    path of a trained preconditioner blob (for Preconditioner=unet)
 *   UnetSmoothing  smoothing steps after each network application inside the
 *                  solve, 1..64 (10); the training loss uses 2
 *   UnetChebyshev  ratio > 1: the steps use Chebyshev weights for the eigenvalue
 *                  interval [2 / ratio, 2] of the Jacobi-scaled operator; 0: every
 *                  step uses the weights file's omega (16)
 *   Report         0 | 1: RCMatrix_utility::report statistics of every assembled
 *                  system into shk_result.report_* (RCMatrix_utility.h:98-146) (0)
 *   ProfileKernels 0 | 1: time every Krylov kernel launch with HIP events on the
 *                  context's stream (read with shk_kernel_stats; resets the counters)
 * Unknown keys and unparsable values return SHK_BAD_ARG. */
int shk_configure(shk_ctx *ctx, const char *key, const char *value);

/* Most recent error message of this context ("" if none).  ctx may be NULL for
 * errors of shk_create. */
const char *shk_last_error(const shk_ctx *ctx);

/* ------------------------------------------------------------------ projection */

/* macproject3_interface::project (macproject3_interface.h:79-87, implemented at
 * macpressuresolver3.cpp:54-299).  u, v, w are updated in place.  solid is the
 * nodal solid level set, fluid the cell-centred liquid level set, us/vs/ws the
 * solid face velocities.  surface_tension != 0 adds the ghost-fluid surface-tension
 * jump to the interface faces first (macpressuresolver3.cpp:95-124), also when the
 * scene has no unknowns; otherwise a scene without unknowns leaves the velocity
 * untouched (macpressuresolver3.cpp:170).  The reference's rigidbodies argument is
 * not used by its pressure solver and has no counterpart here. */
int shk_project(shk_ctx *ctx, double dt, float *u, float *v, float *w,
                const float *solid, const float *us, const float *vs, const float *ws,
                const float *fluid, double surface_tension, shk_result *result);

/* ------------------------------------------------------------------ 2D projection */

/* macproject2_interface::initialize(shape2, dx) for the 2D solver macpressuresolver2
 * (src/projection/macpressuresolver2.cpp:40-300).  The context holds the nx x ny grid
 * as nx x 1 x ny (the 2D y axis on the library's z axis) in 64x1x8 bricks.  Every
 * context call applies, with these 2D conventions:
 *   cell  nx*ny        index i + nx*j           (pressure, fluid level set)
 *   u     (nx+1)*ny    index i + (nx+1)*j
 *   v     nx*(ny+1)    index i + nx*j
 *   solid (nx+1)*(ny+1) nodal, index i + (nx+1)*j
 * shk_set_solid_density, shk_export_fractions and shk_divergence take the x and y
 * face arrays in their first two slots and NULL in the third.  Preconditioner none,
 * jacobi (either precision) and mg (2x1x2 aggregation) are available in 2D; unet is
 * 3D only. */
int shk_create2d(int nx, int ny, double dx, int device, shk_ctx **out);

/* macproject2_interface::project (macpressuresolver2.cpp:52-281): the same steps as
 * shk_project with 4-neighbour cells, the 2D area fraction 1 - fraction of the face's
 * two solid nodes (macutility2.cpp:104-111) and the 5-point curvature.  u, v are
 * updated in place. */
int shk_project2d(shk_ctx *ctx, double dt, float *u, float *v, const float *solid, const float *us,
                  const float *vs, const float *fluid, double surface_tension, shk_result *result);

/* macproject3_interface::set_target_volume (macproject3_interface.h:51,
 * macpressuresolver3.cpp:46-49): the liquid volumes the next projections' PI
 * volume correction works from (with Gain != 0 and target != 0). */
int shk_set_target_volume(shk_ctx *ctx, double current_volume, double target_volume);

/* macpressuresolver3::serialize / initialize(const filestream &)
 * (macpressuresolver3.cpp:319-344): the projection's state that a resumed run needs,
 * in the reference's field order: shape, dx, target and current volume, the PI
 * controller's integral and, after a warm-started projection, the previous pressure
 * (here per cell, FP64).  shk_state_bytes gives the size, shk_save_state writes it to a
 * host buffer of at least that size, shk_load_state restores it into a context of the
 * same shape and dx (SHK_BAD_ARG otherwise).  The bytes are what a plugin writes through
 * its filestream. */
int shk_state_bytes(const shk_ctx *ctx, size_t *bytes);
int shk_save_state(shk_ctx *ctx, void *buf, size_t capacity);
int shk_load_state(shk_ctx *ctx, const void *buf, size_t bytes);

/* macproject3_interface::set_solid_density (macproject3_interface.h:58,
 * macpressuresolver3.cpp:51-53): per-face solid density (three face arrays of the
 * u, v, w shapes; device pointers, or host pointers with HostPointers=1) that lets
 * liquid pass solid faces: area' = density*area + (1 - density), and the updated
 * velocity is scaled by 1 - (1 - area')*density (macpressuresolver3.cpp:75-80,
 * 287-293).  The context keeps the pointers, as the reference keeps its pointer;
 * all three NULL switches it off. */
int shk_set_solid_density(shk_ctx *ctx, const float *du, const float *dv, const float *dw);

/* get_pressure() (macproject3_interface.h:93): nx*ny*nz float device array of the
 * last projection (0 on non-unknowns).  Owned by the context. */
const float *shk_pressure(const shk_ctx *ctx);

/* Copies get_pressure() into dst (a device buffer of nx*ny*nz floats, or a host
 * buffer when the context has HostPointers=1). */
int shk_copy_pressure(shk_ctx *ctx, float *dst);

/* Copies the last projection's face fractions ("areas" from
 * macutility3::compute_area_fraction, "rhos" from compute_fluid_fraction) into
 * caller buffers of the face sizes above.  Any pointer may be NULL. */
int shk_export_fractions(shk_ctx *ctx, float *area_u, float *area_v, float *area_w,
                         float *rho_u, float *rho_v, float *rho_w);

/* Copies the last projection's linear system in dense cell layout: label[c] = 1
 * for unknowns else 0; diag, the three low-face couplings (A[c, c-e_d] = -c_d[c]
 * when both cells are unknowns), rhs and the FP64 solution x.  Any pointer may be NULL. */
int shk_export_system(shk_ctx *ctx, uint8_t *label, double *diag, double *cx, double *cy,
                      double *cz, double *rhs, double *x);

/* FP64 weighted divergence max over the last projection's unknowns of a face
 * field: the negative of the rhs macpressuresolver3.cpp:209-213 would assemble
 * from (u, v, w).  With use_fp64_update = 1 the faces are first updated in FP64
 * with the FP64 solution (u - dt*(p_i - p_{i-1})/(rho*dx)), which is the
 * post-projection divergence without float storage rounding. */
int shk_divergence(shk_ctx *ctx, double dt, const float *u, const float *v, const float *w,
                   const float *us, const float *vs, const float *ws, int use_fp64_update,
                   double *max_abs_div);

/* ------------------------------------------------------------------ linear solver */

/* RCMatrix_solver_interface<size_t,double>::solve (RCMatrix_solver.h:77) as the
 * pcg plugin does it (pcg.cpp:45-73): x0 = 0, stop when max|r| <= Residual*max|b|.
 * CSR with sorted int64 columns, rows = number of rows; device pointers (host pointers
 * with HostPointers=1).  Preconditioner=mg is the amg plugin's solve (amg.cpp:43-166):
 * CG preconditioned by a V-cycle over a pairwise-aggregation hierarchy built on the
 * host from the matrix and kept for the next call with the same matrix (a content
 * hash of rowptr, cols and vals decides); a matrix whose coarsest level (where the
 * aggregation stops shrinking) keeps more than 1,024 rows returns SHK_BAD_ARG, one whose
 * coarsest matrix is not positive definite SHK_BREAKDOWN. */
int shk_solve_csr(shk_ctx *ctx, int64_t rows, const int64_t *rowptr, const int64_t *cols,
                  const double *vals, const double *b, double *x, shk_result *result);

/* No counterpart in the reference.  shk_solve_csr for callers that cannot allocate
 * device memory during a solve (code under HIP graph capture, serving loops,
 * harnesses that forbid hipMalloc): the same kernels and result, with every piece of
 * per-call device workspace (partials, the five CG vectors, the diagonal and the
 * Jacobi-scaled values) carved from one static per-device array of
 * shk_noalloc_scratch_bytes() bytes (128 MiB) instead of context buffers.  It calls
 * no hipMalloc or hipFree; the context itself is allocated by shk_create as usual.
 * Preconditioner none or jacobi; a system whose workspace
 * (6 * rows + nnz [jacobi] + 7,360 doubles) exceeds the array returns SHK_BAD_ARG, as
 * do ProfileKernels=1 and HostPointers=1.  Every call on one device shares that
 * device's array: calls that overlap on one GPU (threads, contexts or streams) race. */
int shk_solve_csr_noalloc(shk_ctx *ctx, int64_t rows, const int64_t *rowptr, const int64_t *cols,
                          const double *vals, const double *b, double *x, shk_result *result);
size_t shk_noalloc_scratch_bytes(void);

/* Accumulated launch count and total milliseconds of one Krylov kernel
 * ("spmv_dot": direction update fused with the stencil and s.As, or "axpy_norm")
 * since ProfileKernels was last set.
 * Only launches that did work are counted. */
int shk_kernel_stats(shk_ctx *ctx, const char *kernel, int64_t *launches, double *total_ms);

/* ------------------------------------------------------------------ multigrid preconditioner */

/* z = M r with the multigrid V-cycle of the last projection (Preconditioner=mg): dense
 * FP32 device grids (nx*ny*nz) in the Jacobi-scaled unknowns (r^ = D^-1/2 r,
 * z^ = D^1/2 z).  The hierarchy: 2x2x2 aggregates with Galerkin coarse operators,
 * one damped-Jacobi pre- and post-smoothing step (w = 0.9), coarse-correction factor
 * 1.5, dense solve on the coarsest level (<= 4 cells per axis).  A test and
 * inspection entry point; shk_project applies the same V-cycle inside CG. */
int shk_mg_apply(shk_ctx *ctx, const float *r, float *z);

/* ------------------------------------------------------------------ learned preconditioner */

/* A 4-level 3D U-Net (16/32/64/128 channels, BF16 MFMA convolutions, FP32
 * accumulation) used as the preconditioner of flexible CG when a context has
 * Preconditioner=unet and UnetWeights=<file written by shk_unet_save>.  It has no
 * counterpart in the reference, whose pcg plugin applies no preconditioner
 * (pcg_solver.h:379-384); its right-hand sides and labels are the reference
 * system of the projection above.
 *
 * All grids below are dense, x-fastest, FP32 device arrays of the Jacobi-scaled
 * system A^ = D^-1/2 A D^-1/2 (unit diagonal) with fine dims that are multiples
 * of 16; the network works on the half-resolution grid. */
typedef struct shk_unet shk_unet;

int shk_unet_create(uint64_t seed, int device, shk_unet **out);
void shk_unet_destroy(shk_unet *net);
const char *shk_unet_last_error(const shk_unet *net);
int shk_unet_num_params(const shk_unet *net, int64_t *count);
/* Copies the FP32 parameters (grads = 0) or their accumulated gradients (grads = 1)
 * into a host array of shk_unet_num_params floats. */
int shk_unet_get_params(shk_unet *net, float *host, int grads);
/* Sets the parameters from a host array and clears the gradients. */
int shk_unet_set_params(shk_unet *net, const float *host);
int shk_unet_save(shk_unet *net, const char *path);
int shk_unet_load(shk_unet *net, const char *path);

/* The smoothing of shk_unet_train_step's M and of shk_unet_apply: steps steps (1..64;
 * 2 on creation), each with omega, or, chebyshev > 1, with the Chebyshev weights of the
 * solver's UnetChebyshev for the eigenvalue interval [2 / chebyshev, 2]. */
int shk_unet_set_smoothing(shk_unet *net, int steps, double chebyshev);
/* One training sample: the A^-norm error loss of M(r) against xstar = A^-1 r,
 * L = (M(r) - xstar)' A^ (M(r) - xstar) / (xstar' r), where
 *   z0 = mask * s * P U(R r / s, R mask, R cx, R cy, R cz)   (R: 2x2x2 average,
 *        P: cell-centred trilinear prolongation, s: RMS of r over the mask)
 *   z_{i+1} = z_i + w_i * mask * (r - A^ z_i), the smoothing steps (w_i = omega:
 *        two steps unless shk_unet_set_smoothing says otherwise);  M(r) = z_last.
 * Gradients accumulate; with lr > 0 an Adam step is taken and the gradients are
 * cleared.  omega is kept with the network and saved with its weights. */
int shk_unet_train_step(shk_unet *net, int nx, int ny, int nz, const float *r, const float *mask, const float *cx,
                        const float *cy, const float *cz, const float *xstar, double omega, float lr,
                        double *loss);
/* z = M(r) (the preconditioner application). */
int shk_unet_apply(shk_unet *net, int nx, int ny, int nz, const float *r, const float *mask, const float *cx,
                   const float *cy, const float *cz, double omega, float *z);

/* r = b - A^ x on a dense grid with the couplings above (device pointers; r aliases no
 * input).  Builds training residuals. */
int shk_unet_residual(int nx, int ny, int nz, const float *cx, const float *cy, const float *cz, const float *x,
                      const float *b, float *r);

/* Dense FP32 copies (nx*ny*nz, device) of the last projection's scaled system:
 * b^ = D^-1/2 b, mask (1 on unknowns), the low-face couplings of A^ and
 * x^ = D^1/2 x (the scaled solution).  Any pointer may be NULL. */
int shk_export_scaled(shk_ctx *ctx, float *rhat, float *mask, float *hx, float *hy, float *hz, float *xhat);

/* One application of the learned preconditioner of the last projection with
 * Preconditioner=unet (features, network, prolongation and smoothing on the bricks,
 * as inside its flexible CG): z = M r on dense nx*ny*nz float device arrays of the
 * Jacobi-scaled unknowns.  No counterpart in the reference; for tests and tools. */
int shk_unet_precondition(shk_ctx *ctx, const float *r, float *z);

/* Layer-level access to the network's convolution kernels (device pointers):
 * x and y are channels-last BF16 grids, w is W[tap][ci][co] FP32 with
 * tap = (dz+1)*9 + (dy+1)*3 + (dx+1); cin a multiple of 8, cout of 16.
 * y = act(conv3x3x3(x, w) + bias) with zero padding (act = ReLU when relu). */
int shk_nn_conv3d(int nx, int ny, int nz, const void *x, int cin, const float *w, const float *bias, int cout,
                  int relu, void *y);
/* The same convolution on the input [upsample(coarse), x] (coarse: the half-size grid
 * with c1 channels, nearest upsampling; x: cin - c1 channels; coarse = NULL and c1 = 0:
 * x alone with cin channels), as the decoder layers read it, by kernel 0 (the choice the
 * network's inference makes for this grid), 1 (the tiled kernel the training uses),
 * 2 (the row-reuse kernel; cout 16 or 32) or 3 (the direct-load kernel of the small
 * grids). */
int shk_nn_conv3d_cat(int nx, int ny, int nz, const void *coarse, int c1, const void *x, int cin, const float *w,
                      const float *bias, int cout, int relu, int kernel, void *y);
/* dx = input gradient (BF16, NULL to skip), dw / db += weight / bias gradients (FP32). */
int shk_nn_conv3d_backward(int nx, int ny, int nz, const void *x, int cin, const float *w, int cout, const void *dz,
                           void *dx, float *dw, float *db);

/* ------------------------------------------------------------------ scenes */

/* Synthetic scenes for tests and benchmarks, generated on the host into HOST
 * buffers of the layouts above (needs no GPU).
 *   kind 1: pool y < 0.3 plus a sphere r = 0.15 at (0.5, 0.7, 0.5), no solids
 *           (walls at the domain boundary only).
 *   kind 3: 20 seeded solid boxes/spheres (radii 0.03-0.12), pool y < 0.4 plus
 *           8 seeded liquid spheres, liquid clipped by the solid as
 *           macutility3.cpp:352-361 clips it; liquid components with no free
 *           surface are turned to air.
 * Face velocities are uniform in [-1, 1) from a counter-based hash of
 * (seed, component, face index); solid velocities are 0.  dx is the cell size.
 * Returns SHK_OK; *dropped (nullable) receives the number of cells turned to air. */
int shk_scene_generate(int kind, int nx, int ny, int nz, double dx, uint64_t seed,
                       float *fluid, float *solid, float *u, float *v, float *w,
                       float *us, float *vs, float *ws, int64_t *dropped);

/* ABI version (SHK_ABI_VERSION) and a build string. */
int shk_abi_version(void);
const char *shk_build_info(void);

#ifdef __cplusplus
}
#endif
#endif /* SHK_PROJECTION_H */
