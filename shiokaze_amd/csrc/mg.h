// mg.h -- aggregation multigrid V-cycle on the brick layout (mg.hip), the
// preconditioner of Preconditioner=mg.
//
// Level 0 is the Jacobi-scaled system A^ = D^-1/2 A D^-1/2 (unit diagonal on the
// unknowns) in the solver's bricks (8x8x8, or 64x1x8 for a 2D grid held as nx x 1 x
// ny).  Level l+1 aggregates the cells of level l 2x2x2 (2x1x2 on a 2D grid: cell
// (X,Y,Z) owns the cells (2X+a, 2Y+b, 2Z+c)) with the Galerkin operator P0' A_l P0 of
// piecewise-constant prolongation P0: again a 7-point (5-point) operator,
//   couplings  c_C = sum of the fine couplings across the aggregates' shared face,
//   diagonal   d_C = sum of the fine diagonals - 2 * (the couplings inside C).
// Coarse levels keep the brick shape of level 0.  Coarsening stops when every axis
// has at most 4 cells; that level (<= 64 cells, one brick) is solved exactly with a
// dense inverse computed on the host.
//
// V-cycle (symmetric, so M is a fixed SPD operator and plain PCG applies):
//   x_l = w D_l^-1 b_l;  t_l = b_l - A_l x_l;  b_{l+1} = P0' t_l;  (recurse)
//   x_l += alpha P0 x_{l+1};  x_l += w D_l^-1 (b_l - A_l x_l)
// with w = 0.9 and alpha = 1.5 (the over-correction that unsmoothed aggregation needs).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

#include "krylov.h"

namespace shk {

constexpr float kMgOmega = 0.9f;  // default damped-Jacobi weight (MgOmega)
constexpr float kMgAlpha = 1.5f;  // default coarse-correction factor (MgAlpha)

struct MgOp {
    const int *nbr;
    const float *d, *cx, *cy, *cz;
    int n_active;
    float omega = kMgOmega, alpha = kMgAlpha;
    // level 0 only: per-brick uniform coupling (unit diagonal, one coupling value; 0 =
    // not uniform), which lets those bricks skip the diagonal and coupling reads
    const float *kappa = nullptr;
    int bx = 8;  // brick shape: 8 (8x8x8) or 64 (64x1x8)
};

struct MgLevelDev {
    int nx = 0, ny = 0, nz = 0;     // cells
    int nbx = 0, nby = 0, nbz = 0;  // bricks
    int n_active = 0;
    int *slot = nullptr;    // dense brick table -> slot or -1
    int4 *coord = nullptr;  // brick coordinates per slot
    int *nbr = nullptr;     // 6 neighbour slots per slot
    int *parent = nullptr;  // slot of the parent brick on the next level (levels < coarsest)
    float *d = nullptr, *cx = nullptr, *cy = nullptr, *cz = nullptr;
    float *b = nullptr, *x = nullptr, *t = nullptr;
    int bx = 8;             // brick shape: 8 (8x8x8) or 64 (64x1x8)
    __host__ __device__ MgOp op(float omega, float alpha) const {
        return MgOp{nbr, d, cx, cy, cz, n_active, omega, alpha, nullptr, bx};
    }
};

constexpr int kMgCoarseMax = 64;  // cells of the coarsest level (<= 4 per axis)

// level-0 diagonal of the scaled system: 1 on unknowns (sig > 0), else 0
void launch_mg_mask(hipStream_t st, int64_t n, const double *sig, float *d0);
// coarse operator of level `c` from level `f`
void launch_mg_coarsen(hipStream_t st, const MgLevelDev &f, const MgLevelDev &c);
// b_c = P0' (b_f - A_f x_f), x_c = w b_c / d_c (or 0 with w = 0): the residual of a
// level >= 1 and its restriction in one pass, the residual never stored
void launch_mg_residual_restrict(hipStream_t st, const MgLevelDev &f, const MgLevelDev &c, float omega,
                                 const SolveState *state);
// x_f += alpha P0 x_c on the fine unknowns (levels >= 1)
void launch_mg_prolong(hipStream_t st, const MgLevelDev &f, const MgLevelDev &c, float alpha, const SolveState *state);
// c.t = c.x + w (c.b - A c.x) / d on a level >= 1; with `fine` (a level >= 1 above c)
// also fine->x += alpha P0 c.t
void launch_mg_smooth(hipStream_t st, const MgLevelDev &c, const MgLevelDev *fine, float omega, float alpha,
                      const SolveState *state);
// level 0 (scaled system, unit diagonal on the unknowns, vectors zero elsewhere):
// the residual t = r - A (w r) after the pre-smoothing step x0 = w r, restricted at
// once to level 1 c1 (c1.b = P0' t, c1.x = w_c c1.b / c1.d; t is not stored); coord
// and parent are level 0's brick coordinates and parent slots
void launch_mg_pre0(hipStream_t st, const MgOp &op, const float *r, const int4 *coord, const int *parent,
                    const MgLevelDev &c1, float omega_c, const SolveState *state);
// y = x0' + w (r - A x0') with x0' = w r + alpha P0 x1 (prolongation fused into the
// post-smoothing), one thread per cell; the r.y sums of its blocks (cell_part, one per
// 256 cells) are folded in order into g.blocks partials `part` for the Krylov kernels
void launch_mg_post0(hipStream_t st, const MgOp &op, const int4 *coord, const int *parent, const float *r,
                     const float *xc, float *y, const SolveState *state, const Grid &g, double *cell_part,
                     double *part);
// coarsest level: x[idx[i]] = sum_j inv[i][j] b[idx[j]] (n <= kMgCoarseMax), other cells 0
void launch_mg_coarse(hipStream_t st, int n, const int *idx, const float *inv, const float *b, float *x,
                      int cells, const SolveState *state);

}  // namespace shk
