/* shk_liquid.h -- C-ABI of the velocity half of the liquid step around the pressure
 * projection, on the same context as shk_projection.h (a 3D context from shk_create;
 * device pointers only).
 *
 * macliquid3::idle (src/liquid/macliquid3.cpp:342-391) runs, per step: the time step
 * from max |u| (:354), extend_both = velocity extrapolation + the solid constraint
 * (:330-340), level-set advection (:366), velocity self-advection (:369-370), gravity
 * (:373), the projection (:382).  Every call below replaces one of those steps for a
 * caller that keeps its MAC velocity on the GPU between steps.
 *
 * Layouts are those of shk_projection.h: dense x-fastest float face arrays
 * u (nx+1) ny nz, v nx (ny+1) nz, w nx ny (nz+1), and per face array one uint8
 * activity mask of the same shape (1 = active).  The masks replace the reference's
 * sparse activity (array3.h active()): an inactive face reads as the background 0.
 * Results are bit-for-bit those of the reference's float/double arithmetic.
 *
 * Configuration keys (shk_configure): MacCormack (default 1), TrimNarrowBand
 * (default 1), WENO (default 0: trilinear; 1: the semi-Lagrangian values by 6th-order
 * WENO interpolation, WENO3.h:92-122) -- macadvection3.cpp:58-60, 280-286.
 *
 * Every call is synchronous at return and returns an shk_status code. */
#ifndef SHK_LIQUID_H
#define SHK_LIQUID_H

#include <stdint.h>

#include "shk_projection.h"

#ifdef __cplusplus
extern "C" {
#endif

/* macarray_extrapolator3::extrapolate(velocity, width)
 * (include/shiokaze/array/macarray_extrapolator3.h:43-47, array_extrapolator3.h:47-76,
 * dilate3.h:43-97): `width` rounds; each round activates every inactive face with an
 * active 6-neighbour and gives it their float average.  The masks are updated in place;
 * *activated (nullable) receives the number of faces activated. */
int shk_extrapolate_velocity(shk_ctx *ctx, float *u, float *v, float *w, uint8_t *mu, uint8_t *mv, uint8_t *mw,
                             int width, int64_t *activated);

/* macutility3_interface::constrain_velocity(solid, velocity) (src/utility/macutility3.cpp:61-92):
 * active faces inside the nodal solid level set ((nx+1)(ny+1)(nz+1) floats, nullable
 * for no solid) lose the velocity component into the solid; v is then clamped to
 * the domain's floor and ceiling. */
int shk_constrain_velocity(shk_ctx *ctx, const float *solid, float *u, float *v, float *w, const uint8_t *mu,
                           const uint8_t *mv, const uint8_t *mw);

/* macutility3_interface::compute_max_u(velocity) (macutility3.cpp:48-60): the largest
 * cell-centred speed over cells whose six faces are active (0 if none). */
int shk_compute_max_u(shk_ctx *ctx, const float *u, const float *v, const float *w, const uint8_t *mu,
                      const uint8_t *mv, const uint8_t *mw, double *max_u);

/* macadvection3_interface::advect_vector(u, velocity, fluid, dt) as macliquid3 calls
 * it (macliquid3.cpp:369-370; macadvection3.cpp:46-54, 70-177): the velocity advected
 * by itself over dt, MacCormack with the min/max limiter unless MacCormack=0, and
 * semi-Lagrangian where the cell-centred level set `fluid` (nx ny nz) at the departed
 * face is above -TrimNarrowBand*dx.  In place; inactive faces come back 0. */
int shk_advect_velocity(shk_ctx *ctx, double dt, float *u, float *v, float *w, const uint8_t *mu, const uint8_t *mv,
                        const uint8_t *mw, const float *fluid);

/* macadvection3_interface::advect_scalar(fluid, velocity, fluid_save, dt) as the level-set
 * tracker calls it (src/surfacetracker/maclevelsetsurfacetracker3.cpp:49-52;
 * macadvection3.cpp:37-45, 178-277): the cells of the cell-centred level set phi
 * (nx ny nz) whose mask mphi is set (the narrow band) are advected by the MAC velocity,
 * MacCormack unless MacCormack=0, semi-Lagrangian where the level set before the step
 * is above -TrimNarrowBand*dx at the departure point; the other cells keep the values
 * they hold (the sparse array's fill and background reads).  `background` is the level
 * set's background value (its half bandwidth, array3.h:261-266): the MacCormack
 * backward pass reads it outside the band (macadvection3.cpp:248).  Redistancing is not
 * part of this call. */
int shk_advect_levelset(shk_ctx *ctx, double dt, float *phi, const uint8_t *mphi, const float *u, const float *v,
                        const float *w, const uint8_t *mu, const uint8_t *mv, const uint8_t *mw, float background);

/* `m_velocity += dt*gravity` of macliquid3::inject_external_force (macliquid3.cpp:195-210):
 * adds (float)(dt*g) to every face and activates every face. */
int shk_add_gravity(shk_ctx *ctx, double dt, double gx, double gy, double gz, float *u, float *v, float *w,
                    uint8_t *mu, uint8_t *mv, uint8_t *mw);

/* redistancer3_interface::redistance(phi, width) of the pderedistancer3 plugin
 * (src/redistancer/pderedistancer3.cpp:39-183) on the cell-centred level set phi
 * (nx*ny*nz floats) and its narrow-band mask (uint8, 1 = active), both in place:
 * the band is trimmed to the cells next to a sign change (gridutility3.cpp:184-216),
 * dilated `width` times (new cells at -/+dx), evolved by ceil(width / IntegrationRate)
 * steps of the redistancing PDE (Euler, or RK2 with RedistTemporalScheme=RK2) and cut at
 * |phi| > width*dx; every cell off the final band is set to -width*dx where its value is
 * negative and +width*dx elsewhere (the level-set fill that the reference's flood fill
 * assigns; the dense values carry each cell's side).  *active receives the final band's
 * cell count (may be NULL).  Keys: IntegrationRate (0.75), RedistTemporalScheme (Euler). */
int shk_redistance(shk_ctx *ctx, float *phi, uint8_t *band, int width, int64_t *active);

/* gridutility3::extrapolate_levelset(solid, fluid) as the level-set tracker calls it after
 * redistancing (maclevelsetsurfacetracker3.cpp:59-61; gridutility3.cpp:87-146) on the
 * cell-centred level set phi (nx*ny*nz floats) and its band mask, in place, with the
 * nodal solid ((nx+1)(ny+1)(nz+1) floats) and its node activity solid_band (uint8, NULL:
 * every node active).  Nothing changes unless an active solid node is below 0.  Band
 * cells whose interpolated solid value s is below 0 take the level set interpolated at
 * the cell moved by -s/dx along the normalised solid gradient (with
 * HorizontalSolidWallExtrapolation, and a normal whose y part is below sin(pi/4), the
 * minimum of that and the value along the normal without its y part); the band is
 * dilated once (new cells >= 0 next to a cell below 0 become dx); then, as
 * combine_levelset(solid, fluid, combined, dx) (:66-86), the band becomes its union with
 * the cells that have an active solid node, each with max(phi, -dx - mean of its 8
 * solid nodes).  Cells off the final band take -background where that expression is
 * negative and +background elsewhere (the flood fill's side), `background` being the
 * level set's half bandwidth.  Arrays are read as given off their bands: pass what the
 * reference reads there (the background with the side's sign, as shk_redistance
 * leaves it).  *active (nullable) receives the final band's cell count.  Keys:
 * SolidWallExtrapolation (1; 0: band cells inside the solid take s),
 * HorizontalSolidWallExtrapolation (1) -- gridutility3.cpp:394-395, 413-414. */
int shk_extrapolate_levelset(shk_ctx *ctx, const float *solid, const uint8_t *solid_band, float *phi, uint8_t *band,
                             float background, int64_t *active);

/* The activity the projection leaves (macpressuresolver3.cpp:269-283, `it.set_off()`):
 * after an shk_project with unknowns, an active face stays active only where its area
 * and liquid fractions of that projection are both non-zero; a projection without
 * unknowns leaves the masks alone. */
int shk_project_activity(shk_ctx *ctx, uint8_t *mu, uint8_t *mv, uint8_t *mw);

#ifdef __cplusplus
}
#endif
#endif
