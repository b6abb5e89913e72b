// krylov.h -- host launchers of the conjugate gradient kernels (krylov.hip).
//
// The iteration is PCGSolver::solve (local/include/pcgsolver/pcg_solver.h:246-295):
// x0 = 0, r = b, z = M^-1 r, s = z, then per iteration
//   q = A s; alpha = rho / s.q; x += alpha s; r -= alpha q;
//   stop if max|r| <= tol * max|b|; z = M^-1 r; rho' = z.r; s = z + (rho'/rho) s.
// M is the identity (the reference's effective preconditioner, pcg_solver.h:379-384),
// diag(A) (Jacobi, as a symmetric scaling of the system) or a multigrid V-cycle
// (mg.h).  Each kernel publishes one partial per workgroup; every workgroup of the
// next kernel folds them in one fixed order, so results are deterministic, every
// workgroup sees the same scalars, and no host round trip or grid-wide fence is
// needed between iterations.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "shk_common.h"

namespace shk {

struct Reduction {
    double *pa;       // s.q partials (>= blocks)
    double *pb;       // r.r and max|r| partials (>= 2 * blocks)
    SolveState *st;   // device state
    // with a preconditioner z = M r: the r.z partials (>= blocks) written by the last
    // kernel of M; dir_spmv is then given z in place of r and takes rho = r.z
    double *pz = nullptr;
};

// Operator views.  Brick: vectors hold n_active bricks of 8x8x8 cells, x-fastest
// inside a brick; nbr[6*slot + {-x,+x,-y,+y,-z,+z}] is the neighbour brick's slot
// or -1.  The matrix is diag - sum_faces c_face * neighbour with c_face stored at
// the face's high cell (cx/cy/cz = low face of a cell) and zero on every face
// that touches a non-unknown; diag == nullptr means a unit diagonal (the
// symmetrically Jacobi-scaled system).
template <class C>
struct BrickOp {
    const int *nbr;
    const C *diag, *cx, *cy, *cz;
    int n_active;
    // optional: per-brick constant coupling (0 = not uniform); with a diagonal array the
    // operator also needs kdiag, the brick's constant diagonal
    const C *kappa = nullptr;
    int bx = 8;  // brick shape: 8 (8x8x8) or 64 (64x1x8, a 2D grid held as nx x 1 x ny)
    const C *kdiag = nullptr;
};
struct CsrOp {
    const int64_t *rowptr, *cols;
    const double *vals;
    int64_t rows;
};

constexpr int kMaxBlocks = 1024;  // workgroups of a Krylov kernel: 4 of 4 waves per CU

__host__ __device__ inline int blocks_for(int64_t n) { return (int)((n + kBrickCells - 1) / kBrickCells); }

// Launch shape of the Krylov kernels over n elements (tiles of 512).
struct Grid {
    int64_t n = 0;
    int tiles = 0;   // ceil(n / 512)
    int per = 1;     // tiles per workgroup
    int blocks = 1;  // workgroups (<= max_blocks)
};
Grid make_grid(int64_t n, int max_blocks = kMaxBlocks);

// x = 0, r = s = sig*rhs (or rhs); rho = r.r; rnorm0 = max|rhs| over unknowns;
// tol_abs = tol_rel * rnorm0.  sig != nullptr selects the scaled system.
template <class T, class TB>
void launch_cg_init(hipStream_t st, const Grid &g, const TB *rhs, const double *sig, T *x, T *r, T *s, Reduction red,
                    double tol_rel, int64_t max_iter);

// Optional timing of one launch: the start/stop events are recorded by the dispatch
// itself (hipExtLaunchKernelGGL), so their interval is the kernel's own duration.
struct KEvents {
    hipEvent_t start = nullptr, stop = nullptr;
};

// One CG iteration is two kernels (see krylov.hip): dir_spmv (stop test, beta,
// s_out = r + beta s_in, q = A s_out, partial s_out.q) and axpy_norm (alpha,
// x += alpha s, r -= alpha q, partials r.r and max|r|).  The first dir_spmv after
// init takes beta = 0.  s alternates between two buffers.  With sig != nullptr the
// stop test uses the unscaled residual max|r/sig|.
//
// workgroups of the stencil kernel launch_dir_spmv picks for op on g that the device
// holds at once (its register use decides), at most g.blocks: a solve whose stencil
// fits fewer than g.blocks runs every kernel on a grid of this size, so that no
// workgroups of the stencil wait for a second round
template <class T, class C>
int dir_spmv_resident(const Grid &g, const BrickOp<C> &op);

template <class T, class C>
void launch_dir_spmv(hipStream_t st, const Grid &g, const BrickOp<C> &op, const T *r, const T *s_in, T *s_out, T *q,
                     Reduction red, KEvents ev = {});
void launch_dir_spmv_csr(hipStream_t st, const Grid &g, const CsrOp &op, const double *r, const double *s_in,
                         double *s_out, double *q, Reduction red, KEvents ev = {});
template <class T>
void launch_axpy_norm(hipStream_t st, const Grid &g, const double *sig, T *x, T *r, const T *s, const T *q,
                      Reduction red, KEvents ev = {});

// r = b - A x in FP64 (unscaled operator with its diagonal); part holds 4*blocks
// partials; out = {max|r|, sum r^2, sum b^2, max|b|}.  Optionally rs = (float)
// (sig * r) (or (float) r without sig): the next inner solve's right-hand side.
void launch_residual(hipStream_t st, const Grid &g, const BrickOp<double> &op, const double *b, const double *x,
                     double *r, double *part, double *out, const double *sig = nullptr, float *rs = nullptr);
void launch_residual_csr(hipStream_t st, const Grid &g, const CsrOp &op, const double *b, const double *x, double *r,
                         double *part, double *out);

// Flexible CG (FCG(1)) on FP32 brick vectors of a unit-diagonal brick operator with
// device-side scalars (fst: 3 doubles; part1 and part2 2 x blocks each): the direction
// kernel below (launch_fcg_dir_spmv) writes part1 = (p.q, p.r); then
//   axpy: alpha = p.r / p.q; x += alpha p; r -= alpha q; partials r.r and max|r| into
//         part2 (norm_only: the partials of r alone); fst = (p.q, alpha, bad).
// With fs: nothing happens once fs->stop is set (the stop rules run between the axpy
// and the next application of the network: fcg_decide).
void launch_fcg_axpy(hipStream_t st, const Grid &g, float *x, float *r, const float *p, const float *q,
                     const double *part1, double *part2, double *fst, bool norm_only,
                     const FcgState *fs = nullptr);
// *fs = v (one thread: the value travels as a kernel argument)
void launch_fcg_state_init(hipStream_t st, FcgState *fs, const FcgState &v);

// FCG direction and stencil in one pass (the Jacobi path's direction kernel in its
// flexible-CG form): p_out = z + beta p_in with beta = -(z.q) / fst[0] from the nz
// partials zq (zq == nullptr: beta = 0, p_out = z; give p_in = z then), q = A p_out,
// part[b] = p_out.q, part[blocks + b] = p_out.res; nothing once fs->stop is set.
// Unit-diagonal 3D brick operator.
struct FcgDir {
    const FcgState *fs;
    const double *zq;
    int nz;
    const double *fst;
    const float *res;
    double *part;
};
void launch_fcg_dir_spmv(hipStream_t st, const Grid &g, const BrickOp<float> &op, const float *z, const float *p_in,
                         float *p_out, float *q, const FcgDir &fd);

// x64 += sig * (double) d32 (sig may be nullptr): the refinement correction.
void launch_add_correction(hipStream_t st, int64_t n, double *x, const float *d, const double *sig);
// x *= sig: back from the scaled unknowns.
void launch_unscale(hipStream_t st, int64_t n, double *x, const double *sig);
// diag[row] = A[row,row], sig[row] = 1/sqrt(diag) (0 if not positive).
void launch_csr_diag(hipStream_t st, const CsrOp &op, double *diag, double *sig);
// out = values of D^-1/2 A D^-1/2 on A's pattern.
void launch_csr_scale(hipStream_t st, const CsrOp &op, const double *sig, double *out);

}  // namespace shk
