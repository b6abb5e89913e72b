// setup.h -- host launchers of the per-projection kernels (setup.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "shk_fractions.h"

namespace shk {

// density: per-face solid density (nullptr entries: no blend)
void launch_fractions(hipStream_t st, Dims d, const float *solid, const float *fluid, double eps_solid,
                      double eps_fluid, int so_solid, int so_fluid, const float *const density[3],
                      float *const area[3], float *const rho[3]);

// Surface-tension jump added to the face velocities (curv: nx*ny*nz scratch).
void launch_surface_tension(hipStream_t st, Dims d, double dt, double dx, double kappa, const float *fluid,
                            float *curv, float *const rho[3], float *const vel[3]);

// Warm start: w = previous dense pressure on the current bricks' unknowns; x += w.
void launch_gather_warm(hipStream_t st, Dims d, int n_active, const int4 *brick_coord, const uint8_t *label,
                        const double *prev, double *w);
void launch_add64(hipStream_t st, int64_t n, double *x, const double *w);

void launch_label(hipStream_t st, Dims d, const float *fluid, float *const area[3], float *const rho[3], uint8_t *label,
                  uint8_t *brick_flag, int nbx, int nby, unsigned long long *count);

// diag (0 on non-unknowns), low-face off-diagonal couplings (0 unless both cells
// are unknowns), rhs, sig = 1/sqrt(diag) (0 on non-unknowns) and optional FP32
// copies of diag and couplings, all in brick layout.
void launch_coeff(hipStream_t st, Dims d, double dt, double dx, int n_active, const int4 *brick_coord,
                  float *const area[3], float *const rho[3], const uint8_t *label, const float *const vel[3],
                  const float *const svel[3], double rhs_correct, double *diag, double *cx, double *cy, double *cz,
                  double *rhs, double *sig, float *diag32, float *cx32, float *cy32, float *cz32);

// Couplings of D^-1/2 A D^-1/2 (FP64 and/or FP32 outputs; any may be nullptr).
// RCMatrix_utility::report statistics of the brick system into part[8 * blocks]
// (blocks <= n_active or any positive count; each block folds its strided bricks)
void launch_report(hipStream_t st, int bx, int n_active, int blocks, const int *nbr, const double *diag,
                   const double *cx, const double *cy, const double *cz, double *part);
void launch_scale_coeff(hipStream_t st, int bx, int n_active, const int *nbr, const double *sig, const double *cx,
                        const double *cy, const double *cz, double *hx, double *hy, double *hz, float *hx32,
                        float *hy32, float *hz32);

// kappa[slot] = the one coupling value of a brick whose stencil is uniform, else 0.
template <class C>
void launch_brick_uniform(hipStream_t st, int bx, int n_active, const int *nbr, const C *cx, const C *cy, const C *cz,
                          C *kappa, const C *diag = nullptr, C *kdiag = nullptr);

// prev (nullable): also the dense FP64 pressure, for the next warm start
void launch_scatter(hipStream_t st, Dims d, const int *brick_slot, int nbx, int nby, const uint8_t *label,
                    const double *x, float *pressure, double *prev);

void launch_update_velocity(hipStream_t st, Dims d, double dt, double dx, const float *pressure, float *const area[3],
                            float *const rho[3], const float *const svel[3], const float *const density[3],
                            float *const vel[3]);

// Writes one partial max per block into `partial`; returns the block count.
int launch_divergence(hipStream_t st, Dims d, double dt, double dx, float *const area[3], float *const rho[3],
                      const uint8_t *label, const float *const vel[3], const float *const svel[3], int use_fp64_update,
                      const int *brick_slot, int nbx, int nby, const double *x, double *partial, int max_blocks);

void launch_export(hipStream_t st, Dims d, const int *brick_slot, int nbx, int nby, const uint8_t *label,
                   const double *diag, const double *cx, const double *cy, const double *cz, const double *rhs,
                   const double *x, uint8_t *o_label, double *o_diag, double *o_cx, double *o_cy, double *o_cz,
                   double *o_rhs, double *o_x);

}  // namespace shk
